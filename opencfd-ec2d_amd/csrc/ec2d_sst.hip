// Menter SST k-omega model: SST_kw (sub_turbulence_SST.f90:6-271) with
// Amut_boundary (sub_turbulence_SA.f90:180-220), in two launches over the
// rows of a mesh level, before the mean-flow residual kernel.
//
//   k_sst_cell  per interior cell: velocity / k / omega gradients by central
//               differences in index space (:32-70), vorticity, the blending
//               functions F2 and F1, mu_t with the a1 limiter and the 1e5 cap,
//               the production / destruction sources Qk, Qw (:71-113); writes
//               mu*Re, the raw mu_t, F1, Qk and Qw to scratch planes;
//   k_sst_res   per interior cell: its four faces' k and omega fluxes
//               (first-order upwind convection plus blended diffusion with
//               face-centred gradients, :116-242), Res(5:6) in the
//               reference's summation order (:245-262), mu = (mu*Re)/Re and
//               mu_t = mu_t/Re into the Amu / Amu_t planes, and the wall ghost
//               cells' mu_t = -mu_t of their interior neighbour.
//
// The reference scales Amu by Re in place and divides it back at the end
// (:85, :259); (mu*Re)/Re is not always mu, and the mean-flow viscous flux
// reads the round-tripped value, so the residual kernel repeats the round
// trip for interior cells.  Ghost mu_t other than at walls keeps whatever the
// Amu_t plane holds, as the reference's array does.
#include "ec2d_device.h"
#include "ec2d_physics.h"

namespace ec2d {

struct SstArgs {
  const BlockDesc* blocks;
  const int2* rows;  // (block, j) per workgroup
  const CaseParams* cps;
  double* arena;
  PlaneMap pm;
};

// scratch planes of the model, from pm.sst on
enum { SST_AMS = 0, SST_ATR = 1, SST_F1 = 2, SST_QK = 3, SST_QW = 4, SST_NPLANES = 5 };

__global__ __launch_bounds__(256) void k_sst_cell(SstArgs a) {
  const int2 rj = a.rows[blockIdx.x];
  const BlockDesc B = a.blocks[rj.x];
  const int j = rj.y;
  const int nx = B.nx, ny = B.ny;
  if (j < 1 || j > ny - 1 || case_frozen(B.case_id)) return;
  const CaseParams cp = a.cps[B.case_id];
  const PlaneMap pm = a.pm;
  double* base = a.arena + B.base;
  auto PL = [&](int slot) { return base + (int64_t)slot * B.plane; };
  const double *U0 = PL(pm.U), *U1 = PL(pm.U + 1), *U2 = PL(pm.U + 2), *U3 = PL(pm.U + 3), *U4 = PL(pm.U + 4),
               *U5 = PL(pm.U + 5), *x1 = PL(pm.x1), *y1 = PL(pm.y1);
  const double sw2 = 0.856, beta1 = 0.075, Cw1s = 0.533, beta2 = 0.0828, Cw2s = 0.440;
  const double a1s = 0.31, betas = 0.09;
  const double Re = cp.Re;
  for (int i = 1 + threadIdx.x; i <= nx - 1; i += blockDim.x) {
    const int64_t q = pidx(B, i, j), qe = pidx(B, i + 1, j), qw = pidx(B, i - 1, j), qn = pidx(B, i, j + 1),
                  qs = pidx(B, i, j - 1);
    // primitives and k, omega as Flow_Var / SST_kw hold them (d, uu, v; Kt = U5/d, Wt = U6/d)
    auto uu = [&](int64_t p) { return U1[p] / U0[p]; };
    auto vv = [&](int64_t p) { return U2[p] / U0[p]; };
    auto Kt = [&](int64_t p) { return U4[p] / U0[p]; };
    auto Wt = [&](int64_t p) { return U5[p] / U0[p]; };
    const double d = U0[q];
    const double uc = uu(q), vc = vv(q);
    const double T = (U3[q] - 0.5 * d * (uc * uc + vc * vc)) / (cp.Cv * d);
    const double amu = (1.0 / cp.Re * (1.0 + cp.Tsb) * sqrt(T * T * T) / (cp.Tsb + T)) * Re;
    const double kc = Kt(q), wc = Wt(q);
    const double Dix = (x1[qe] - x1[qw]) * 0.5;
    const double Diy = (y1[qe] - y1[qw]) * 0.5;
    const double Djx = (x1[qn] - x1[qs]) * 0.5;
    const double Djy = (y1[qn] - y1[qs]) * 0.5;
    const double Ds = 1.0 / (Dix * Djy - Djx * Diy);
    const double Diu = (uu(qe) - uu(qw)) * 0.5;
    const double Div = (vv(qe) - vv(qw)) * 0.5;
    const double Dik = (Kt(qe) - Kt(qw)) * 0.5;
    const double Diw = (Wt(qe) - Wt(qw)) * 0.5;
    const double Dju = (uu(qn) - uu(qs)) * 0.5;
    const double Djv = (vv(qn) - vv(qs)) * 0.5;
    const double Djk = (Kt(qn) - Kt(qs)) * 0.5;
    const double Djw = (Wt(qn) - Wt(qs)) * 0.5;
    const double vx = (Div * Djy - Djv * Diy) * Ds;
    const double kx = (Dik * Djy - Djk * Diy) * Ds;
    const double wx = (Diw * Djy - Djw * Diy) * Ds;
    const double uy = (-Diu * Djx + Dju * Dix) * Ds;
    const double ky = (-Dik * Djx + Djk * Dix) * Ds;
    const double wy = (-Diw * Djx + Djw * Dix) * Ds;
    const double omega = vx - uy;
    const double dw = PL(pm.dw)[q];
    const double dw2 = dw * dw, Re2 = Re * Re;
    const double arg2 = fmax(2.0 * sqrt(fabs(kc)) / (EC2D_SP(0.09) * wc * dw * Re), 500.0 * amu / (d * wc * dw2 * Re2));
    const double f2 = tanh(arg2 * arg2);
    const double at = fmin(fmin(d * kc / wc, a1s * d * kc * Re / (f2 * fabs(omega))), EC2D_SP(100000.0));
    const double Kws = 2.0 * (kx * wx + ky * wy) * d * sw2 / (wc + 1.0e-20);
    const double CD_kw = fmax(Kws, 1.0e-20);
    const double arg3 = fmax(sqrt(fabs(kc)) / (EC2D_SP(0.09) * wc * dw * Re), 500.0 * amu / (d * wc * dw2 * Re2));
    const double arg1 = fmin(arg3, 4.0 * d * sw2 * kc / (CD_kw * dw2));
    const double f1 = tanh(((arg1 * arg1) * arg1) * arg1);
    const double Pk = at * omega * omega;
    const double Pk0 = fmin(Pk, 20.0 * betas * kc * wc * Re2);
    const double Qk = Pk0 / Re - betas * d * wc * kc * Re;
    const double Cw_ = f1 * Cw1s + (1.0 - f1) * Cw2s;
    const double beta_ = f1 * beta1 + (1.0 - f1) * beta2;
    const double Qw = Cw_ * d * omega * omega / Re - beta_ * d * (wc * wc) * Re + (1.0 - f1) * Kws / Re;
    PL(pm.sst + SST_AMS)[q] = amu;
    PL(pm.sst + SST_ATR)[q] = at;
    PL(pm.sst + SST_F1)[q] = f1;
    PL(pm.sst + SST_QK)[q] = Qk;
    PL(pm.sst + SST_QW)[q] = Qw;
  }
}

// k and omega fluxes through one i-face (between cells (i-1,j) and (i,j),
// :116-176) or one j-face (between (i,j-1) and (i,j), :182-242).
template <bool IFACE>
__device__ __forceinline__ void sst_face(double& Fk, double& Fw, const BlockDesc& B, const double* base,
                                         const PlaneMap& pm, int i, int j, double Re) {
  const double sk1 = 0.85, sw1 = 0.5, sk2 = 1.0, sw2 = 0.856;
  auto PL = [&](int slot) { return base + (int64_t)slot * B.plane; };
  const double *U0 = PL(pm.U), *U1 = PL(pm.U + 1), *U2 = PL(pm.U + 2), *U4 = PL(pm.U + 4), *U5 = PL(pm.U + 5),
               *x1 = PL(pm.x1), *y1 = PL(pm.y1), *AmS = PL(pm.sst + SST_AMS), *AtR = PL(pm.sst + SST_ATR);
  auto Kt = [&](int ii, int jj) { const int64_t p = pidx(B, ii, jj); return U4[p] / U0[p]; };
  auto Wt = [&](int ii, int jj) { const int64_t p = pidx(B, ii, jj); return U5[p] / U0[p]; };
  const int nx = B.nx, ny = B.ny;
  const int64_t q = pidx(B, i, j);
  const int ia = IFACE ? i - 1 : i, ja = IFACE ? j : j - 1;  // the other cell of the face
  const int64_t qa = pidx(B, ia, ja);
  const double s = PL(IFACE ? pm.si : pm.sj)[q], n1 = PL(IFACE ? pm.ni1 : pm.nj1)[q], n2 = PL(IFACE ? pm.ni2 : pm.nj2)[q];
  const double un1 = U1[qa] / U0[qa] * n1 + U2[qa] / U0[qa] * n2;
  const double un2 = U1[q] / U0[q] * n1 + U2[q] / U0[q] * n2;
  const double ka = Kt(ia, ja), kb = Kt(i, j), wa = Wt(ia, ja), wb = Wt(i, j);
  Fk = -0.5 * ((un1 + fabs(un1)) * ka + (un2 - fabs(un2)) * kb) * s;
  Fw = -0.5 * ((un1 + fabs(un1)) * wa + (un2 - fabs(un2)) * wb) * s;
  // F1 of the face's own cell (i, j); outside the interior the model never set it
  const bool own_in = i >= 1 && i <= nx - 1 && j >= 1 && j <= ny - 1;
  const double f1 = own_in ? PL(pm.sst + SST_F1)[q] : 0.0;
  const double sK = f1 * sk1 + (1.0 - f1) * sk2;
  const double sW = f1 * sw1 + (1.0 - f1) * sw2;
  const bool first = IFACE ? i == 1 : j == 1, last = IFACE ? i == nx : j == ny;
  double A1, A2;
  if (first) {
    A1 = AmS[q] + sK * AtR[q];
    A2 = AmS[q] + sW * AtR[q];
  } else if (last) {
    A1 = AmS[qa] + sK * AtR[qa];
    A2 = AmS[qa] + sW * AtR[qa];
  } else if (IFACE) {
    A1 = (AmS[qa] + AmS[q] + sK * (AtR[qa] + AtR[q])) * 0.5;
    A2 = (AmS[qa] + AmS[q] + sW * (AtR[qa] + AtR[q])) * 0.5;
  } else {
    A1 = (AmS[q] + AmS[qa] + sK * (AtR[q] + AtR[qa])) * 0.5;
    A2 = (AmS[q] + AmS[qa] + sW * (AtR[q] + AtR[qa])) * 0.5;
  }
  double Dix, Diy, Djx, Djy, Dik, Diw, Djk, Djw;
  if (IFACE) {
    Dix = x1[q] - x1[qa];
    Diy = y1[q] - y1[qa];
    Djx = (x1[pidx(B, i - 1, j + 1)] + x1[pidx(B, i, j + 1)] - x1[pidx(B, i - 1, j - 1)] - x1[pidx(B, i, j - 1)]) * 0.25;
    Djy = (y1[pidx(B, i - 1, j + 1)] + y1[pidx(B, i, j + 1)] - y1[pidx(B, i - 1, j - 1)] - y1[pidx(B, i, j - 1)]) * 0.25;
    Dik = kb - ka;
    Diw = wb - wa;
    Djk = (Kt(i - 1, j + 1) + Kt(i, j + 1) - Kt(i - 1, j - 1) - Kt(i, j - 1)) * 0.25;
    Djw = (Wt(i - 1, j + 1) + Wt(i, j + 1) - Wt(i - 1, j - 1) - Wt(i, j - 1)) * 0.25;
  } else {
    Dix = (x1[pidx(B, i + 1, j - 1)] + x1[pidx(B, i + 1, j)] - x1[pidx(B, i - 1, j - 1)] - x1[pidx(B, i - 1, j)]) * 0.25;
    Diy = (y1[pidx(B, i + 1, j - 1)] + y1[pidx(B, i + 1, j)] - y1[pidx(B, i - 1, j - 1)] - y1[pidx(B, i - 1, j)]) * 0.25;
    Djx = x1[q] - x1[qa];
    Djy = y1[q] - y1[qa];
    Dik = (Kt(i + 1, j - 1) + Kt(i + 1, j) - Kt(i - 1, j - 1) - Kt(i - 1, j)) * 0.25;
    Diw = (Wt(i + 1, j - 1) + Wt(i + 1, j) - Wt(i - 1, j - 1) - Wt(i - 1, j)) * 0.25;
    Djk = kb - ka;
    Djw = wb - wa;
  }
  const double Ds = 1.0 / (Dix * Djy - Djx * Diy);
  const double Kx = (Dik * Djy - Djk * Diy) * Ds, Wx = (Diw * Djy - Djw * Diy) * Ds;
  const double Ky = (-Dik * Djx + Djk * Dix) * Ds, Wy = (-Diw * Djx + Djw * Dix) * Ds;
  Fk = Fk + A1 * (Kx * n1 + Ky * n2) * s / Re;
  Fw = Fw + A2 * (Wx * n1 + Wy * n2) * s / Re;
}

__global__ __launch_bounds__(256) void k_sst_res(SstArgs a) {
  const int2 rj = a.rows[blockIdx.x];
  const BlockDesc B = a.blocks[rj.x];
  const int j = rj.y;
  const int nx = B.nx, ny = B.ny;
  if (j < 1 || j > ny - 1 || case_frozen(B.case_id)) return;
  const CaseParams cp = a.cps[B.case_id];
  const PlaneMap pm = a.pm;
  double* base = a.arena + B.base;
  auto PL = [&](int slot) { return base + (int64_t)slot * B.plane; };
  const double Re = cp.Re;
  const double* flag = PL(pm.flag);
  for (int i = 1 + threadIdx.x; i <= nx - 1; i += blockDim.x) {
    const int64_t q = pidx(B, i, j);
    double kw, ww, ke, we, ks, ws, kn, wn;
    sst_face<true>(kw, ww, B, base, pm, i, j, Re);
    sst_face<true>(ke, we, B, base, pm, i + 1, j, Re);
    sst_face<false>(ks, ws, B, base, pm, i, j, Re);
    sst_face<false>(kn, wn, B, base, pm, i, j + 1, Re);
    const double vol = PL(pm.vol)[q];
    double r5 = ke - kw;
    double r6 = we - ww;
    r5 = r5 + kn - ks + PL(pm.sst + SST_QK)[q] * vol;
    r6 = r6 + wn - ws + PL(pm.sst + SST_QW)[q] * vol;
    PL(pm.Res + 4)[q] = r5;
    PL(pm.Res + 5)[q] = r6;
    PL(pm.Amu)[q] = PL(pm.sst + SST_AMS)[q] / Re;
    const double at = PL(pm.sst + SST_ATR)[q] / Re;
    double* Amu_t = PL(pm.Amu_t);
    Amu_t[q] = at;
    // Amut_boundary: a wall ghost of layer 0 takes -mu_t of this cell
    if (i == 1 && flag[pidx(B, 0, j)] != 0.0) Amu_t[pidx(B, 0, j)] = -at;
    if (i == nx - 1 && flag[pidx(B, nx, j)] != 0.0) Amu_t[pidx(B, nx, j)] = -at;
    if (j == 1 && flag[pidx(B, i, 0)] != 0.0) Amu_t[pidx(B, i, 0)] = -at;
    if (j == ny - 1 && flag[pidx(B, i, ny)] != 0.0) Amu_t[pidx(B, i, ny)] = -at;
  }
}

}  // namespace ec2d
