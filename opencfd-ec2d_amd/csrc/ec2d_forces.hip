// Wall forces and force coefficients on the finest mesh.  The reference has
// no force integration (only wall Cp in its separate post-processing tool,
// read_flow-1.2.f90:36-51); the definition, shared with the CPU oracle:
//   every wall face (subfaces with neighb = -10, in block / file / index
//   order), with n its unit normal pointing into the fluid and s its length,
//   contributes
//     pressure  -(p - p_inf) n s, p of the adjacent interior cell;
//     viscous   (tau . n) s, tau = mu (grad u + grad u^T - 2/3 div u I) from
//               the face-centred gradient stencil of Residual_FVM
//               (sub_Residual.f90:301-343, 408-450), mu = Amu + Amu_t of the
//               adjacent interior cell (that routine's wall-face rule).
// One workgroup per case; each thread sums every 256th face in order, then a
// fixed tree combines the threads: the result depends only on the input.
#include "ec2d_device.h"

namespace ec2d {

struct ForceArgs {
  const BlockDesc* blocks;
  const int4* faces;      // (block, face, i, j): face index of the wall face
  const int* case_face0;  // faces of case c: [case_face0[c], case_face0[c+1])
  const CaseParams* cps;
  const double* arena;
  PlaneMap pm;
  int viscous;
  double* out;            // 4 per case: pressure Fx, Fy, viscous Fx, Fy
};

__global__ __launch_bounds__(256) void k_forces(ForceArgs a) {
  __shared__ double sh[4][256];
  const int c = blockIdx.x;
  const int f0 = a.case_face0[c], f1 = a.case_face0[c + 1];
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int f = f0 + threadIdx.x; f < f1; f += 256) {
    const int4 w = a.faces[f];
    const BlockDesc B = a.blocks[w.x];
    const CaseParams cp = a.cps[B.case_id];
    const double* base = a.arena + B.base;
    auto PL = [&](int slot) { return base + (int64_t)slot * B.plane; };
    const PlaneMap pm = a.pm;
    const int face = w.y, i = w.z, j = w.w, nx = B.nx, ny = B.ny;
    const bool iface = face == 1 || face == 3;
    const double sgn = (face == 3 || face == 4) ? -1.0 : 1.0;
    const int ic = face == 3 ? nx - 1 : i, jc = face == 4 ? ny - 1 : j;
    const int64_t q = pidx(B, i, j), qc = pidx(B, ic, jc);
    const double sl = PL(iface ? pm.si : pm.sj)[q];
    const double n1 = sgn * PL(iface ? pm.ni1 : pm.nj1)[q];
    const double n2 = sgn * PL(iface ? pm.ni2 : pm.nj2)[q];
    const double *U0 = PL(pm.U), *U1 = PL(pm.U + 1), *U2 = PL(pm.U + 2), *U3 = PL(pm.U + 3);
    const double g = cp.gamma, p_inf = 1.0 / (g * cp.Ma * cp.Ma);
    const double d = U0[qc], u = U1[qc] / d, v = U2[qc] / d;
    const double p = (g - 1.0) * (U3[qc] - 0.5 * d * (u * u + v * v));
    const double dp = p - p_inf;
    acc[0] = acc[0] - dp * n1 * sl;
    acc[1] = acc[1] - dp * n2 * sl;
    if (a.viscous) {
      const double *x1 = PL(pm.x1), *y1 = PL(pm.y1);
      auto UU = [&](int ii, int jj) { const int64_t p2 = pidx(B, ii, jj); return U1[p2] / U0[p2]; };
      auto VV = [&](int ii, int jj) { const int64_t p2 = pidx(B, ii, jj); return U2[p2] / U0[p2]; };
      auto X = [&](int ii, int jj) { return x1[pidx(B, ii, jj)]; };
      auto Y = [&](int ii, int jj) { return y1[pidx(B, ii, jj)]; };
      double Dix, Diy, Djx, Djy, Diu, Div, Dju, Djv;
      if (iface) {
        Dix = X(i, j) - X(i - 1, j);
        Diy = Y(i, j) - Y(i - 1, j);
        Djx = (X(i - 1, j + 1) + X(i, j + 1) - X(i - 1, j - 1) - X(i, j - 1)) * 0.25;
        Djy = (Y(i - 1, j + 1) + Y(i, j + 1) - Y(i - 1, j - 1) - Y(i, j - 1)) * 0.25;
        Diu = UU(i, j) - UU(i - 1, j);
        Div = VV(i, j) - VV(i - 1, j);
        Dju = (UU(i - 1, j + 1) + UU(i, j + 1) - UU(i - 1, j - 1) - UU(i, j - 1)) * 0.25;
        Djv = (VV(i - 1, j + 1) + VV(i, j + 1) - VV(i - 1, j - 1) - VV(i, j - 1)) * 0.25;
      } else {
        Dix = (X(i + 1, j - 1) + X(i + 1, j) - X(i - 1, j - 1) - X(i - 1, j)) * 0.25;
        Diy = (Y(i + 1, j - 1) + Y(i + 1, j) - Y(i - 1, j - 1) - Y(i - 1, j)) * 0.25;
        Djx = X(i, j) - X(i, j - 1);
        Djy = Y(i, j) - Y(i, j - 1);
        Diu = (UU(i + 1, j - 1) + UU(i + 1, j) - UU(i - 1, j - 1) - UU(i - 1, j)) * 0.25;
        Div = (VV(i + 1, j - 1) + VV(i + 1, j) - VV(i - 1, j - 1) - VV(i - 1, j)) * 0.25;
        Dju = UU(i, j) - UU(i, j - 1);
        Djv = VV(i, j) - VV(i, j - 1);
      }
      const double Ds = 1.0 / (Dix * Djy - Djx * Diy);
      const double ux = (Diu * Djy - Dju * Diy) * Ds, vx = (Div * Djy - Djv * Diy) * Ds;
      const double uy = (-Diu * Djx + Dju * Dix) * Ds, vy = (-Div * Djx + Djv * Dix) * Ds;
      const double mu = PL(pm.Amu)[qc] + PL(pm.Amu_t)[qc];
      const double div = ux + vy;
      const double t11 = mu * (2.0 * ux - 2.0 / 3.0 * div);
      const double t12 = mu * (uy + vx);
      const double t22 = mu * (2.0 * vy - 2.0 / 3.0 * div);
      acc[2] = acc[2] + (t11 * n1 + t12 * n2) * sl;
      acc[3] = acc[3] + (t12 * n1 + t22 * n2) * sl;
    }
  }
#pragma unroll
  for (int k = 0; k < 4; k++) sh[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
#pragma unroll
      for (int k = 0; k < 4; k++) sh[k][threadIdx.x] = sh[k][threadIdx.x] + sh[k][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x < 4) a.out[4 * c + threadIdx.x] = sh[threadIdx.x][0];
}

}  // namespace ec2d
