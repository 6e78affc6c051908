// Baldwin-Lomax eddy viscosity: turbulence_model_BL and BL_model_1d
// (sub_turbulence_BL.f90:6-275), ahead of the residual kernel, which then
// takes Amu_t as given (ResArgs::closure = 3).
//
// One workgroup per block.  Amu_t = 0 everywhere (sub_Residual.f90:171; the
// model does it again for blocks with a wall); vorticity at the interior cell
// centres; then the wall subfaces in file order, one thread per wall-normal
// line (a serial pass over the line, as BL_model_1d is), a cell reached by
// lines of several walls keeping the smaller mu_t (first write marked in the
// flag plane); after each wall its ghost rows / columns, statement by
// statement with a barrier between them, over the subface's node range as
// the reference writes them.
#include "ec2d_device.h"

namespace ec2d {

struct BlArgs {
  const BlockDesc* blocks;
  const SubfaceDesc* subs;
  const CaseParams* cps;
  double* arena;
  PlaneMap pm;
};

// x**6 as flang lowers a constant integer power: ((x*x)*x)*...
__device__ __forceinline__ double bl_pow6(double x) {
  double r = x;
#pragma unroll
  for (int k = 1; k < 6; k++) r = r * x;
  return r;
}

// d, u, v and the molecular viscosity of a cell, with the residual kernel's
// expressions (comput_duvtpckw, get_viscous)
__device__ __forceinline__ void bl_cell(const double* const* U, int64_t q, const CaseParams& cp, double& d, double& uu,
                                        double& vv, double& am) {
  d = U[0][q];
  uu = U[1][q] / d;
  vv = U[2][q] / d;
  const double T = (U[3][q] - 0.5 * d * (uu * uu + vv * vv)) / (cp.Cv * d);
  am = 1.0 / cp.Re * (1.0 + cp.Tsb) * sqrt(T * T * T) / (cp.Tsb + T);
}

// cell (i, j) of point p (1-based, from the wall) of line L of a wall on face f
__device__ __forceinline__ void bl_point(int f, int L, int p, int nx, int ny, int& i, int& j) {
  if (f == 1) { i = p; j = L; }
  else if (f == 3) { i = nx - p; j = L; }
  else if (f == 2) { i = L; j = p; }
  else { i = L; j = ny - p; }
}

// BL_model_1d on one line, combined into Amu_t
__device__ void bl_line(const BlockDesc& B, const CaseParams& cp, const double* const* U, const double* x1,
                        const double* y1, const double* om, double* At, double* flag, int f, int L) {
  // real*8 parameters from default-real literals (:222)
  const double AP = (double)26.0f, Ccp = (double)1.6f, Ckleb = (double)0.3f, Cwk = 0.25, AKT = (double)0.4f,
               AK = (double)0.0168f;
  const int nx = B.nx, ny = B.ny;
  const int n1 = (f == 1 || f == 3) ? nx : ny;
  double x0, y0;
  if (f == 1) { x0 = (x1[pidx(B, 1, L)] + x1[pidx(B, 0, L)]) * 0.5; y0 = (y1[pidx(B, 1, L)] + y1[pidx(B, 0, L)]) * 0.5; }
  else if (f == 3) { x0 = (x1[pidx(B, nx - 1, L)] + x1[pidx(B, nx, L)]) * 0.5; y0 = (y1[pidx(B, nx - 1, L)] + y1[pidx(B, nx, L)]) * 0.5; }
  else if (f == 2) { x0 = (x1[pidx(B, L, 1)] + x1[pidx(B, L, 0)]) * 0.5; y0 = (y1[pidx(B, L, 1)] + y1[pidx(B, L, 0)]) * 0.5; }
  else { x0 = (x1[pidx(B, L, ny - 1)] + x1[pidx(B, L, ny)]) * 0.5; y0 = (y1[pidx(B, L, ny - 1)] + y1[pidx(B, L, ny)]) * 0.5; }
  int i, j;
  double d, uu, vv, am;
  // TW = max |Amu omega| over the line, Ret from point 1 (:228-232)
  bl_point(f, L, 1, nx, ny, i, j);
  bl_cell(U, pidx(B, i, j), cp, d, uu, vv, am);
  const double d_1 = d, am_1 = am;
  double TW = fabs(am * om[pidx(B, i, j)]);
  for (int p = 1; p <= n1 - 1; p++) {
    bl_point(f, L, p, nx, ny, i, j);
    const int64_t q = pidx(B, i, j);
    bl_cell(U, q, cp, d, uu, vv, am);
    if (fabs(am * om[q]) > TW) TW = fabs(am * om[q]);
  }
  const double Ret = sqrt(d_1 * TW) / am_1;
  auto yy_of = [&](int64_t q) {
    const double dx = x1[q] - x0, dy = y1[q] - y0;
    return sqrt(dx * dx + dy * dy);
  };
  double Fmax = 0.0, etamax = 0.0, Udif = 0.0;
  for (int p = 1; p <= n1 - 1; p++) {  // Ny_boundary = Ny - 1 (:236)
    bl_point(f, L, p, nx, ny, i, j);
    const int64_t q = pidx(B, i, j);
    bl_cell(U, q, cp, d, uu, vv, am);
    const double u2 = sqrt(uu * uu + vv * vv);
    if (u2 > Udif) Udif = u2;
    const double yy = yy_of(q);
    const double etap = yy * Ret;
    const double FF = yy * fabs(om[q]) * (1.0 - exp(-etap / AP));
    if (FF > Fmax) {
      Fmax = FF;
      etamax = yy;
    }
  }
  const double fa = etamax * Fmax, fb = Cwk * etamax * Udif * Udif / Fmax;
  const double Fwak = fb < fa ? fb : fa;  // dmin1; only Fmax = 0 makes fb NaN, and then every Amu_t is 0
  int iflag = 0;
  for (int p = 1; p <= n1 - 1; p++) {
    bl_point(f, L, p, nx, ny, i, j);
    const int64_t q = pidx(B, i, j);
    bl_cell(U, q, cp, d, uu, vv, am);
    const double yy = yy_of(q);
    const double etap = Ret * yy;
    const double bl = AKT * yy * (1.0 - exp(-etap / AP));
    const double visti = d * bl * bl * fabs(om[q]);
    const double Fkleb = 1.0 / (1.0 + 5.5 * bl_pow6(Ckleb * yy / etamax));
    const double visto = AK * Ccp * d * Fwak * Fkleb;
    if (fabs(visto) < fabs(visti)) iflag = 1;
    const double v = iflag == 0 ? visti : visto;
    if (flag[q] == 0.0) {
      flag[q] = 1.0;
      At[q] = v;
    } else {
      At[q] = fmin(At[q], v);
    }
  }
}

__global__ __launch_bounds__(256) void k_bl(BlArgs a) {
  const BlockDesc B = a.blocks[blockIdx.x];
  const CaseParams cp = a.cps[B.case_id];
  const PlaneMap pm = a.pm;
  double* base = a.arena + B.base;
  auto PL = [&](int slot) { return base + (int64_t)slot * B.plane; };
  const double* U[4] = {PL(pm.U), PL(pm.U + 1), PL(pm.U + 2), PL(pm.U + 3)};
  const double* x1 = PL(pm.x1);
  const double* y1 = PL(pm.y1);
  double* At = PL(pm.Amu_t);
  double* om = PL(pm.bl);
  double* flag = PL(pm.bl + 1);
  const int nx = B.nx, ny = B.ny, tid = threadIdx.x, nt = blockDim.x;
  for (int64_t q = tid; q < B.plane; q += nt) {
    At[q] = 0.0;
    om[q] = 0.0;
    flag[q] = 0.0;
  }
  bool wall = false;
  for (int k = 0; k < B.nsub; k++) wall = wall || a.subs[B.bc_off + k].neighb == -10;
  if (!wall) return;  // uniform over the workgroup
  __syncthreads();
  // vorticity at the interior cell centres (:34-55)
  for (int c = tid; c < (nx - 1) * (ny - 1); c += nt) {
    const int i = 1 + c % (nx - 1), j = 1 + c / (nx - 1);
    const double xi = x1[pidx(B, i + 1, j)] - x1[pidx(B, i - 1, j)];
    const double yi = y1[pidx(B, i + 1, j)] - y1[pidx(B, i - 1, j)];
    const double xj = x1[pidx(B, i, j + 1)] - x1[pidx(B, i, j - 1)];
    const double yj = y1[pidx(B, i, j + 1)] - y1[pidx(B, i, j - 1)];
    const double Jac = 1.0 / (xi * yj - xj * yi);
    const double ix = Jac * yj, iy = -Jac * xj, jx = -Jac * yi, jy = Jac * xi;
    double d, ue, ve, uw, vw, un, vn, us, vs, am;
    bl_cell(U, pidx(B, i + 1, j), cp, d, ue, ve, am);
    bl_cell(U, pidx(B, i - 1, j), cp, d, uw, vw, am);
    bl_cell(U, pidx(B, i, j + 1), cp, d, un, vn, am);
    bl_cell(U, pidx(B, i, j - 1), cp, d, us, vs, am);
    const double ui = ue - uw, vi = ve - vw, uj = un - us, vj = vn - vs;
    const double vx = vi * ix + vj * jx;
    const double uy = ui * iy + uj * jy;
    om[pidx(B, i, j)] = fabs(vx - uy);
  }
  __syncthreads();
  for (int k = 0; k < B.nsub; k++) {
    const SubfaceDesc sf = a.subs[B.bc_off + k];
    if (sf.neighb != -10) continue;
    const int f = sf.face;
    const int lo = (f == 1 || f == 3) ? sf.jst : sf.ist, hi = (f == 1 || f == 3) ? sf.jend - 1 : sf.iend - 1;
    for (int L = lo + tid; L <= hi; L += nt) bl_line(B, cp, U, x1, y1, om, At, flag, f, L);
    __syncthreads();
    // the wall's ghost rows / columns (:81-84, 111-114, 146-149, 180-183)
    if (f == 1 || f == 3) {
      for (int i = -1 + tid; i <= nx + 1; i += nt) At[pidx(B, i, sf.jst - 1)] = At[pidx(B, i, sf.jst)];
      __syncthreads();
      for (int i = -1 + tid; i <= nx + 1; i += nt) At[pidx(B, i, sf.jend)] = At[pidx(B, i, sf.jend - 1)];
      __syncthreads();
      const int ic = f == 1 ? nx : 0, is = f == 1 ? nx - 1 : 1;  // copied side
      const int iw = f == 1 ? 0 : nx, iws = f == 1 ? 1 : nx - 1;  // wall side, negated
      for (int j = sf.jst + tid; j <= sf.jend; j += nt) At[pidx(B, ic, j)] = At[pidx(B, is, j)];
      __syncthreads();
      for (int j = sf.jst + tid; j <= sf.jend; j += nt) At[pidx(B, iw, j)] = -At[pidx(B, iws, j)];
      __syncthreads();
    } else {
      const int jw = f == 2 ? 0 : ny, jws = f == 2 ? 1 : ny - 1;  // wall side, negated
      const int jc = f == 2 ? ny : 0, js = f == 2 ? ny - 1 : 1;   // copied side
      for (int i = sf.ist + tid; i <= sf.iend; i += nt) At[pidx(B, i, jw)] = -At[pidx(B, i, jws)];
      __syncthreads();
      for (int i = sf.ist + tid; i <= sf.iend; i += nt) At[pidx(B, i, jc)] = At[pidx(B, i, js)];
      __syncthreads();
      for (int j = -1 + tid; j <= ny + 1; j += nt) At[pidx(B, sf.ist - 1, j)] = At[pidx(B, sf.ist, j)];
      __syncthreads();
      for (int j = -1 + tid; j <= ny + 1; j += nt) At[pidx(B, sf.iend, j)] = At[pidx(B, sf.iend - 1, j)];
      __syncthreads();
    }
  }
}

}  // namespace ec2d
