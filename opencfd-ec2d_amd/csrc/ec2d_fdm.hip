// Finite-difference WENO5 interior of blocks listed in FDM.in:
// Residual_FDM (sub_Finite_Difference.f90:4-150).
//
// The interior (more than LFDM = 4 cells from the block edge) uses
// Steger-Warming flux-vector splitting at cell centres (:154-202) and 5th-order
// WENO-JS along grid lines (:207-275), plus central viscous fluxes (:86-134).
// The fused FVM residual kernel skips these faces and cells; two launches
// here produce the interior residual:
//   k_fdm_lines  one workgroup per grid line, i-lines then j-lines: the split
//                fluxes of the whole line in LDS, WENO5 at every interface,
//                interface fluxes F(i+1/2) to scratch planes; viscous cases
//                add one workgroup per row for Ev1 / Ev2 at cell centres
//                (scratch planes, fdm_visc_row);
//   k_fdm_res    Res = F(i+1/2)-F(i-1/2) + G(j+1/2)-G(j-1/2) for the interior.
// Line length is at most 992 cells (checked on the host).
#include "ec2d_device.h"
#include "ec2d_physics.h"

namespace ec2d {

struct FdmArgs {
  const BlockDesc* blocks;
  const int2* lines;   // (block, line index): j for i-lines, i for j-lines
  const int2* rows;    // (block, j) of every interior row
  const CaseParams* cps;
  double* arena;
  PlaneMap pm;
  int n_i;             // k_fdm_lines: lines[0, n_i) are i-lines, lines[n_i, n_lines) j-lines,
  int n_lines;         // workgroups from n_lines on take the viscous fluxes of rows[blockIdx.x - n_lines]
  int fx, fy;          // first scratch plane of the i- and j-direction interface fluxes (4 each)
  int ev;              // first scratch plane of Ev1 (3) and Ev2 (3)
  int viscous;
};

constexpr int FDM_MAXN = 992;  // 8 line arrays in LDS stay below 64 KiB
constexpr int LFDM = 4;

// sub_Finite_Difference.f90:154-202
__device__ __forceinline__ void split_sw(double d, double u, double v, double cc, double* fp, double* fm, double A1,
                                         double A2, double g) {
  double tmp1 = 2.0 * (g - 1.0);
  double tmp3 = (3.0 - g) / (2.0 * (g - 1.0));
  double vs = A1 * u + A2 * v;
  double ss = sqrt(A1 * A1 + A2 * A2);
  double ak1 = A1 / ss, ak2 = A2 / ss;
  double E1 = vs, E2 = vs - cc * ss, E3 = vs + cc * ss;
  double E1P = (E1 + fabs(E1)) * 0.5, E2P = (E2 + fabs(E2)) * 0.5, E3P = (E3 + fabs(E3)) * 0.5;
  double E1M = E1 - E1P, E2M = E2 - E2P, E3M = E3 - E3P;
  double tmp0 = d / (2.0 * g);
  double uc1 = u - cc * ak1, uc2 = u + cc * ak1, vc1 = v - cc * ak2, vc2 = v + cc * ak2;
  double vvc1 = (uc1 * uc1 + vc1 * vc1) / 2.0, vvc2 = (uc2 * uc2 + vc2 * vc2) / 2.0;
  double vv = (g - 1.0) * (u * u + v * v);
  double W2 = tmp3 * cc * cc;
  fp[0] = tmp0 * (tmp1 * E1P + E2P + E3P);
  fp[1] = tmp0 * (tmp1 * E1P * u + E2P * uc1 + E3P * uc2);
  fp[2] = tmp0 * (tmp1 * E1P * v + E2P * vc1 + E3P * vc2);
  fp[3] = tmp0 * (E1P * vv + E2P * vvc1 + E3P * vvc2 + W2 * (E2P + E3P));
  fm[0] = tmp0 * (tmp1 * E1M + E2M + E3M);
  fm[1] = tmp0 * (tmp1 * E1M * u + E2M * uc1 + E3M * uc2);
  fm[2] = tmp0 * (tmp1 * E1M * v + E2M * vc1 + E3M * vc2);
  fm[3] = tmp0 * (E1M * vv + E2M * vvc1 + E3M * vvc2 + W2 * (E2M + E3M));
}

// WENO5-JS for the positive flux at k+1/2 from v(k-2..k+2): fp_weno5 (:207-237)
__device__ __forceinline__ double weno_p(double vm2, double vm1, double v0, double vp1, double vp2) {
  const double ep = 1.0e-6, C03 = 3.0 / 10.0, C13 = 3.0 / 5.0, C23 = 1.0 / 10.0;
  double a = v0 - 2.0 * vp1 + vp2, b = 3.0 * v0 - 4.0 * vp1 + vp2;
  double S0 = 13.0 / 12.0 * (a * a) + 1.0 / 4.0 * (b * b);
  a = vm1 - 2.0 * v0 + vp1; b = vm1 - vp1;
  double S1 = 13.0 / 12.0 * (a * a) + 1.0 / 4.0 * (b * b);
  a = vm2 - 2.0 * vm1 + v0; b = vm2 - 4.0 * vm1 + 3.0 * v0;
  double S2 = 13.0 / 12.0 * (a * a) + 1.0 / 4.0 * (b * b);
  double e0 = ep + S0, e1 = ep + S1, e2 = ep + S2;
  double a0 = C03 / (e0 * e0), a1 = C13 / (e1 * e1), a2 = C23 / (e2 * e2);
  double W0 = a0 / (a0 + a1 + a2), W1 = a1 / (a0 + a1 + a2), W2 = a2 / (a0 + a1 + a2);
  double q03 = 1.0 / 3.0 * v0 + 5.0 / 6.0 * vp1 - 1.0 / 6.0 * vp2;
  double q13 = -(1.0 / 6.0 * vm1) + 5.0 / 6.0 * v0 + 1.0 / 3.0 * vp1;
  double q23 = 1.0 / 3.0 * vm2 - 7.0 / 6.0 * vm1 + 11.0 / 6.0 * v0;
  return W0 * q03 + W1 * q13 + W2 * q23;
}

// WENO5-JS for the negative flux at k-1/2 from v(k-2..k+2): fm_weno5 (:240-275)
__device__ __forceinline__ double weno_m(double vm2, double vm1, double v0, double vp1, double vp2) {
  const double ep = 1.0e-6, C03 = 3.0 / 10.0, C13 = 3.0 / 5.0, C23 = 1.0 / 10.0;
  double a = v0 - 2.0 * vm1 + vm2, b = 3.0 * v0 - 4.0 * vm1 + vm2;
  double S0 = 13.0 / 12.0 * (a * a) + 1.0 / 4.0 * (b * b);
  a = vp1 - 2.0 * v0 + vm1; b = vp1 - vm1;
  double S1 = 13.0 / 12.0 * (a * a) + 1.0 / 4.0 * (b * b);
  a = vp2 - 2.0 * vp1 + v0; b = vp2 - 4.0 * vp1 + 3.0 * v0;
  double S2 = 13.0 / 12.0 * (a * a) + 1.0 / 4.0 * (b * b);
  double e0 = ep + S0, e1 = ep + S1, e2 = ep + S2;
  double a0 = C03 / (e0 * e0), a1 = C13 / (e1 * e1), a2 = C23 / (e2 * e2);
  double W0 = a0 / (a0 + a1 + a2), W1 = a1 / (a0 + a1 + a2), W2 = a2 / (a0 + a1 + a2);
  double q03 = 1.0 / 3.0 * v0 + 5.0 / 6.0 * vm1 - 1.0 / 6.0 * vm2;
  double q13 = -(1.0 / 6.0 * vp1) + 5.0 / 6.0 * v0 + 1.0 / 3.0 * vm1;
  double q23 = 1.0 / 3.0 * vp2 - 7.0 / 6.0 * vp1 + 11.0 / 6.0 * v0;
  return W0 * q03 + W1 * q13 + W2 * q23;
}

__device__ void fdm_visc_row(const FdmArgs& a, int2 bl);

__global__ __launch_bounds__(256) void k_fdm_lines(FdmArgs a) {
  __shared__ double sfp[4][FDM_MAXN + 1], sfm[4][FDM_MAXN + 1];
  if ((int)blockIdx.x >= a.n_lines) {
    const int2 r = a.rows[blockIdx.x - a.n_lines];
    if (!case_frozen(a.blocks[r.x].case_id)) fdm_visc_row(a, r);
    return;
  }
  const int2 bl = a.lines[blockIdx.x];
  const int dir = (int)blockIdx.x < a.n_i ? 0 : 1;
  const BlockDesc B = a.blocks[bl.x];
  if (case_frozen(B.case_id)) return;
  const CaseParams cp = a.cps[B.case_id];
  const PlaneMap pm = a.pm;
  double* base = a.arena + B.base;
  auto PL = [&](int slot) { return base + (int64_t)slot * B.plane; };
  const int nxp = B.nx - 1, nyp = B.ny - 1;
  const int n = dir == 0 ? nxp : nyp;
  const int line = bl.y;
  // split fluxes at every point k = 1..n of the line (sub_Finite_Difference.f90:37-46, 59-68)
  for (int k = 1 + threadIdx.x; k <= n; k += blockDim.x) {
    int i = dir == 0 ? k : line, j = dir == 0 ? line : k;
    int64_t q = pidx(B, i, j);
    double d = PL(pm.U)[q];
    double uu = PL(pm.U + 1)[q] / d;
    double vv = PL(pm.U + 2)[q] / d;
    double T = (PL(pm.U + 3)[q] - 0.5 * d * (uu * uu + vv * vv)) / (cp.Cv * d);
    double cc = sqrt(T) / cp.Ma;
    double jac = PL(pm.Ajac)[q];
    double A1 = dir == 0 ? PL(pm.Akx)[q] / jac : PL(pm.Aix)[q] / jac;
    double A2 = dir == 0 ? PL(pm.Aky)[q] / jac : PL(pm.Aiy)[q] / jac;
    double fp[4], fm[4];
    split_sw(d, uu, vv, cc, fp, fm, A1, A2, cp.gamma);
    for (int m = 0; m < 4; m++) {
      sfp[m][k] = fp[m];
      sfm[m][k] = fm[m];
    }
  }
  __syncthreads();
  // interface fluxes F(k+1/2) = -(h+(k) + h-(k)) for k = LFDM .. n-LFDM
  double* F = PL(dir == 0 ? a.fx : a.fy);
  for (int k = LFDM + threadIdx.x; k <= n - LFDM; k += blockDim.x) {
    int i = dir == 0 ? k : line, j = dir == 0 ? line : k;
    int64_t q = pidx(B, i, j);
    for (int m = 0; m < 4; m++) {
      const double* p = sfp[m];
      const double* w = sfm[m];
      double hp = weno_p(p[k - 2], p[k - 1], p[k], p[k + 1], p[k + 2]);
      double hm = weno_m(w[k - 1], w[k], w[k + 1], w[k + 2], w[k + 3]);  // h-(k) is the WENO of point k+1
      F[(int64_t)m * B.plane + q] = -(hp + hm);
    }
  }
}

// viscous fluxes at cell centres 2..n-1 (sub_Finite_Difference.f90:86-117), one workgroup per row j
__device__ void fdm_visc_row(const FdmArgs& a, int2 bl) {
  const BlockDesc B = a.blocks[bl.x];
  const CaseParams cp = a.cps[B.case_id];
  const PlaneMap pm = a.pm;
  double* base = a.arena + B.base;
  auto PL = [&](int slot) { return base + (int64_t)slot * B.plane; };
  const int nxp = B.nx - 1, nyp = B.ny - 1;
  const int j = bl.y;
  if (j < 2 || j > nyp - 1) return;
  auto prim = [&](int i, int jj, double& uu, double& vv, double& T) {
    int64_t q = pidx(B, i, jj);
    double d = PL(pm.U)[q];
    uu = PL(pm.U + 1)[q] / d;
    vv = PL(pm.U + 2)[q] / d;
    T = (PL(pm.U + 3)[q] - 0.5 * d * (uu * uu + vv * vv)) / (cp.Cv * d);
  };
  for (int i = 2 + threadIdx.x; i <= nxp - 1; i += blockDim.x) {
    int64_t q = pidx(B, i, j);
    double mu0 = PL(pm.Amu)[q] + PL(pm.Amu_t)[q];
    double Ak0 = cp.Cp * (PL(pm.Amu)[q] / cp.Pr + PL(pm.Amu_t)[q] / EC2D_PRT);
    double ue, ve, Te, uw, vw, Tw, un, vn, Tn, us, vs, Ts, uc, vc, Tc;
    prim(i + 1, j, ue, ve, Te);
    prim(i - 1, j, uw, vw, Tw);
    prim(i, j + 1, un, vn, Tn);
    prim(i, j - 1, us, vs, Ts);
    prim(i, j, uc, vc, Tc);
    double ui = 0.5 * (ue - uw), vi = 0.5 * (ve - vw), Ti = 0.5 * (Te - Tw);
    double uj = 0.5 * (un - us), vj = 0.5 * (vn - vs), Tj = 0.5 * (Tn - Ts);
    double kx = PL(pm.Akx)[q], ky = PL(pm.Aky)[q], ix = PL(pm.Aix)[q], iy = PL(pm.Aiy)[q], jac = PL(pm.Ajac)[q];
    double ux = ui * kx + uj * ix, vx = vi * kx + vj * ix, Tx = Ti * kx + Tj * ix;
    double uy = ui * ky + uj * iy, vy = vi * ky + vj * iy, Ty = Ti * ky + Tj * iy;
    double t11 = (4.0 / 3.0 * ux - 2.0 / 3.0 * vy) * mu0;
    double t22 = (4.0 / 3.0 * vy - 2.0 / 3.0 * ux) * mu0;
    double t12 = (uy + vx) * mu0;
    double E1 = uc * t11 + vc * t12 + Ak0 * Tx;
    double E2 = uc * t12 + vc * t22 + Ak0 * Ty;
    double* ev = PL(a.ev);
    ev[0 * B.plane + q] = (kx * t11 + ky * t12) / jac;
    ev[1 * B.plane + q] = (kx * t12 + ky * t22) / jac;
    ev[2 * B.plane + q] = (kx * E1 + ky * E2) / jac;
    ev[3 * B.plane + q] = (ix * t11 + iy * t12) / jac;
    ev[4 * B.plane + q] = (ix * t12 + iy * t22) / jac;
    ev[5 * B.plane + q] = (ix * E1 + iy * E2) / jac;
  }
}

// interior residual (sub_Finite_Difference.f90:122-145), one workgroup per row j
__global__ __launch_bounds__(256) void k_fdm_res(FdmArgs a) {
  const int2 bl = a.rows[blockIdx.x];
  const BlockDesc B = a.blocks[bl.x];
  const PlaneMap pm = a.pm;
  double* base = a.arena + B.base;
  auto PL = [&](int slot) { return base + (int64_t)slot * B.plane; };
  const int nxp = B.nx - 1, nyp = B.ny - 1;
  const int j = bl.y;
  if (j < LFDM + 1 || j > nyp - LFDM || case_frozen(B.case_id)) return;
  const double* F = PL(a.fx);
  const double* G = PL(a.fy);
  const double* ev = PL(a.ev);
  // viscous additions apply to interfaces with 3 <= i, j <= n-2 (:123-133)
  auto fi = [&](int m, int i, int jj) {
    int64_t q = pidx(B, i, jj);
    double f = F[(int64_t)m * B.plane + q];
    if (a.viscous && m >= 1 && i >= 3 && i <= nxp - 2 && jj >= 3 && jj <= nyp - 2)
      f = f + 0.5 * (ev[(int64_t)(m - 1) * B.plane + q] + ev[(int64_t)(m - 1) * B.plane + pidx(B, i + 1, jj)]);
    return f;
  };
  auto gj = [&](int m, int i, int jj) {
    int64_t q = pidx(B, i, jj);
    double f = G[(int64_t)m * B.plane + q];
    if (a.viscous && m >= 1 && i >= 3 && i <= nxp - 2 && jj >= 3 && jj <= nyp - 2)
      f = f + 0.5 * (ev[(int64_t)(m + 2) * B.plane + q] + ev[(int64_t)(m + 2) * B.plane + pidx(B, i, jj + 1)]);
    return f;
  };
  for (int i = LFDM + 1 + threadIdx.x; i <= nxp - LFDM; i += blockDim.x) {
    int64_t q = pidx(B, i, j);
    for (int m = 0; m < 4; m++)
      PL(pm.Res + m)[q] = (fi(m, i, j) - fi(m, i - 1, j) + gj(m, i, j) - gj(m, i, j - 1));
  }
}

// Res_max / sum Res^2 of every tile from the Res planes (used when FDM blocks
// exist, because their interior residual is only known after k_fdm_res)
template <int NV>
__global__ __launch_bounds__(256) void k_tile_norms(const TileDesc* tiles, const BlockDesc* blocks, double* arena,
                                                    PlaneMap pm, double* partials, int tile_cols, int tile_rows) {
  const int lt = blockIdx.x;
  const TileDesc t = tiles[lt];
  const BlockDesc B = blocks[t.block];
  if (case_frozen(B.case_id)) return;  // a frozen case keeps the partials of its last residual
  double* base = arena + B.base;
  __shared__ double sred[2 * NV][4];
  double mx[NV], sm[NV];
  for (int m = 0; m < NV; m++) { mx[m] = 0.0; sm[m] = 0.0; }
  const int tid = threadIdx.x;
  const int ncell = tile_cols * tile_rows;
  for (int c = tid; c < ncell; c += blockDim.x) {
    int i = t.i0 + c % tile_cols, j = t.j0 + c / tile_cols;
    if (i > B.nx - 1 || j > B.ny - 1) continue;
    int64_t q = pidx(B, i, j);
    for (int m = 0; m < NV; m++) {
      double r = base[(int64_t)(pm.Res + m) * B.plane + q];
      mx[m] = fmax(mx[m], fabs(r));
      sm[m] = sm[m] + r * r;
    }
  }
  for (int m = 0; m < NV; m++) {
    double a = mx[m], b = sm[m];
    for (int off = 32; off > 0; off >>= 1) {
      a = fmax(a, __shfl_xor(a, off));
      b = b + __shfl_xor(b, off);
    }
    if ((tid & 63) == 0) {
      sred[2 * m][tid >> 6] = a;
      sred[2 * m + 1][tid >> 6] = b;
    }
  }
  __syncthreads();
  if (tid < 2 * NV) {
    double acc = sred[tid][0];
    for (int w = 1; w < 4; w++) acc = (tid & 1) ? acc + sred[tid][w] : fmax(acc, sred[tid][w]);
    partials[(int64_t)lt * 2 * NV + tid] = acc;
  }
}

}  // namespace ec2d
