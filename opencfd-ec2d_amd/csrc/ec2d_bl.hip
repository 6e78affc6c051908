// Baldwin-Lomax eddy viscosity: turbulence_model_BL and BL_model_1d
// (sub_turbulence_BL.f90:6-275), ahead of the residual kernel, which then
// takes Amu_t as given (ResArgs::closure = 3).
//
// k_bl_init: Amu_t = 0 everywhere (sub_Residual.f90:171; the model does it
// again for blocks with a wall) and the vorticity at the interior cell
// centres.  Then, wall by wall in file order (wall w of every block per
// launch pair): k_bl_lines, one wave per wall-normal line, a cell reached by
// lines of several walls keeping the smaller mu_t (first write marked in the
// flag plane); k_bl_ghost, the wall's ghost rows / columns statement by
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

// max over the wave's lanes (exact in any order)
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

// the wall subface w (0-based, in file order) of a block, or -1
__device__ __forceinline__ int bl_wall(const BlArgs& a, const BlockDesc& B, int w) {
  for (int k = 0; k < B.nsub; k++)
    if (a.subs[B.bc_off + k].neighb == -10 && w-- == 0) return B.bc_off + k;
  return -1;
}

// Amu_t = 0, flags = 0 and the vorticity at the interior cell centres
// (:31-55), for every block (sub_Residual.f90:171 zeroes Amu_t of blocks
// without a wall too).  Grid: (plane tiles, blocks).
__global__ __launch_bounds__(256) void k_bl_init(BlArgs a) {
  const BlockDesc B = a.blocks[blockIdx.y];
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= B.plane) return;
  const CaseParams cp = a.cps[B.case_id];
  const PlaneMap pm = a.pm;
  double* base = a.arena + B.base;
  auto PL = [&](int slot) { return base + (int64_t)slot * B.plane; };
  PL(pm.Amu_t)[q] = 0.0;
  PL(pm.bl + 1)[q] = 0.0;
  const int i = (int)(q % B.P) - 1, j = (int)(q / B.P) - 1;
  double w = 0.0;
  if (i >= 1 && i <= B.nx - 1 && j >= 1 && j <= B.ny - 1) {
    const double* U[4] = {PL(pm.U), PL(pm.U + 1), PL(pm.U + 2), PL(pm.U + 3)};
    const double* x1 = PL(pm.x1);
    const double* y1 = PL(pm.y1);
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
    w = fabs(vx - uy);
  }
  PL(pm.bl)[q] = w;
}

// BL_model_1d (:218-275) on the lines of wall w of every block, one wave per
// line, the line's points spread over the lanes.  TW, Udif and Fmax are
// maxima, etamax the distance of the first point where F reaches Fmax (the
// serial loop's strict > update), and the switch from the inner to the outer
// viscosity a prefix OR along the line: all exact in any order, so the result
// equals the serial model's.  Grid: (line groups of 4, blocks).
__global__ __launch_bounds__(256) void k_bl_lines(BlArgs a, int w) {
  const BlockDesc B = a.blocks[blockIdx.y];
  const int sidx = bl_wall(a, B, w);
  if (sidx < 0) return;
  const SubfaceDesc sf = a.subs[sidx];
  const int f = sf.face;
  const int lo = (f == 1 || f == 3) ? sf.jst : sf.ist, hi = (f == 1 || f == 3) ? sf.jend - 1 : sf.iend - 1;
  const int L = lo + (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  if (L > hi) return;  // a whole wave
  const int lane = threadIdx.x & 63;
  const CaseParams cp = a.cps[B.case_id];
  const PlaneMap pm = a.pm;
  double* base = a.arena + B.base;
  auto PL = [&](int slot) { return base + (int64_t)slot * B.plane; };
  const double* U[4] = {PL(pm.U), PL(pm.U + 1), PL(pm.U + 2), PL(pm.U + 3)};
  const double* x1 = PL(pm.x1);
  const double* y1 = PL(pm.y1);
  const double* om = PL(pm.bl);
  double* At = PL(pm.Amu_t);
  double* flag = PL(pm.bl + 1);
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
  auto qof = [&](int p) {
    int i, j;
    bl_point(f, L, p, nx, ny, i, j);
    return pidx(B, i, j);
  };
  auto yy_of = [&](int64_t q) {
    const double dx = x1[q] - x0, dy = y1[q] - y0;
    return sqrt(dx * dx + dy * dy);
  };
  double d, uu, vv, am;
  // TW = max |Amu omega|, Ret from point 1 (:228-232)
  double TW = 0.0;
  for (int p = 1 + lane; p <= n1 - 1; p += 64) {
    const int64_t q = qof(p);
    bl_cell(U, q, cp, d, uu, vv, am);
    TW = fmax(TW, fabs(am * om[q]));
  }
  TW = wave_max(TW);
  bl_cell(U, qof(1), cp, d, uu, vv, am);
  const double Ret = sqrt(d * TW) / am;
  // Udif, Fmax and the first point reaching Fmax (:238-251)
  double Udif = 0.0, Fm = 0.0;
  for (int p = 1 + lane; p <= n1 - 1; p += 64) {
    const int64_t q = qof(p);
    bl_cell(U, q, cp, d, uu, vv, am);
    Udif = fmax(Udif, sqrt(uu * uu + vv * vv));
    const double yy = yy_of(q);
    const double etap = yy * Ret;
    Fm = fmax(Fm, yy * fabs(om[q]) * (1.0 - exp(-etap / AP)));
  }
  Udif = wave_max(Udif);
  const double Fmax = wave_max(Fm);
  int pfirst = 0x7fffffff;
  if (Fmax > 0.0) {
    for (int p = 1 + lane; p <= n1 - 1; p += 64) {
      const int64_t q = qof(p);
      const double yy = yy_of(q);
      const double etap = yy * Ret;
      if (yy * fabs(om[q]) * (1.0 - exp(-etap / AP)) == Fmax) {
        pfirst = p;
        break;  // this lane's points come in increasing order
      }
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) pfirst = min(pfirst, __shfl_xor(pfirst, o));
  const double etamax = Fmax > 0.0 ? yy_of(qof(pfirst)) : 0.0;
  const double fa = etamax * Fmax, fb = Cwk * etamax * Udif * Udif / Fmax;
  const double Fwak = fb < fa ? fb : fa;  // dmin1; only Fmax = 0 makes fb NaN, and then every Amu_t is 0
  // the switch to the outer viscosity is sticky along the line (:254-266):
  // chunks of 64 points in order, a prefix OR inside each by ballot
  bool carry = false;
  for (int p0 = 1; p0 <= n1 - 1; p0 += 64) {
    const int p = p0 + lane;
    const bool on = p <= n1 - 1;
    double visti = 0.0, visto = 0.0;
    int64_t q = 0;
    if (on) {
      q = qof(p);
      bl_cell(U, q, cp, d, uu, vv, am);
      const double yy = yy_of(q);
      const double etap = Ret * yy;
      const double bl = AKT * yy * (1.0 - exp(-etap / AP));
      visti = d * bl * bl * fabs(om[q]);
      const double Fkleb = 1.0 / (1.0 + 5.5 * bl_pow6(Ckleb * yy / etamax));
      visto = AK * Ccp * d * Fwak * Fkleb;
    }
    const unsigned long long m = __ballot(on && fabs(visto) < fabs(visti));
    const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);
    const bool iflag = carry || (m & upto) != 0;
    carry = carry || m != 0;
    if (on) {
      const double v = iflag ? visto : visti;
      if (flag[q] == 0.0) {
        flag[q] = 1.0;
        At[q] = v;
      } else {
        At[q] = fmin(At[q], v);
      }
    }
  }
}

// the ghost rows / columns of wall w of every block (:81-84, 111-114,
// 146-149, 180-183), statement by statement.  Grid: blocks.
__global__ __launch_bounds__(256) void k_bl_ghost(BlArgs a, int w) {
  const BlockDesc B = a.blocks[blockIdx.x];
  const int sidx = bl_wall(a, B, w);
  if (sidx < 0) return;  // uniform over the workgroup
  const SubfaceDesc sf = a.subs[sidx];
  double* At = a.arena + B.base + (int64_t)a.pm.Amu_t * B.plane;
  const int nx = B.nx, ny = B.ny, tid = threadIdx.x, nt = blockDim.x, f = sf.face;
  if (f == 1 || f == 3) {
    for (int i = -1 + tid; i <= nx + 1; i += nt) At[pidx(B, i, sf.jst - 1)] = At[pidx(B, i, sf.jst)];
    __syncthreads();
    for (int i = -1 + tid; i <= nx + 1; i += nt) At[pidx(B, i, sf.jend)] = At[pidx(B, i, sf.jend - 1)];
    __syncthreads();
    const int ic = f == 1 ? nx : 0, is = f == 1 ? nx - 1 : 1;   // copied side
    const int iw = f == 1 ? 0 : nx, iws = f == 1 ? 1 : nx - 1;  // wall side, negated
    for (int j = sf.jst + tid; j <= sf.jend; j += nt) At[pidx(B, ic, j)] = At[pidx(B, is, j)];
    __syncthreads();
    for (int j = sf.jst + tid; j <= sf.jend; j += nt) At[pidx(B, iw, j)] = -At[pidx(B, iws, j)];
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
  }
}

}  // namespace ec2d
