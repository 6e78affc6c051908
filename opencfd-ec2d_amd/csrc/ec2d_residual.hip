// Fused residual kernel: Comput_Residual_one_mesh (sub_Residual.f90:13-107)
// for all blocks of a batch in one launch.
//
// One workgroup computes one TX x TY tile of cells of one block:
//   1. stage U of the tile plus a 2-cell halo from HBM into LDS and turn it
//      into primitives (comput_duvtpckw, sub_Residual.f90:114-152), Sutherland
//      viscosity (get_viscous, :670-689) and, for SA, nu_t
//      (sub_turbulence_SA.f90:24-36); the wall ghost override of Amut_boundary
//      (:180-220) is applied in LDS;
//   2. every i-face and j-face of the tile: reconstruction, rotation, flux,
//      back-rotation, viscous flux (Residual_FVM, :200-455) and the SA face
//      fluxes (sub_turbulence_SA.f90:43-115); results stay in LDS;
//   3. per cell: Res = Fi(i+1)-Fi(i)+Fj(j+1)-Fj(j) (:459-469), the SA
//      residual and source (sub_turbulence_SA.f90:77-168), spectral radii
//      (comput_Lvc, sub_NS_singlegrid.f90:352-388), local time step (comput_dt,
//      :324-348) and this tile's share of Res_max / sum(Res^2).
// HBM traffic per cell: U read once (+halo), Res, dt, Lci, Lcj (Lvi, Lvj)
// written once; faces shared by two tiles are recomputed, never re-read.
#include "ec2d_device.h"
#include "ec2d_physics.h"

namespace ec2d {

struct ResArgs {
  const TileDesc* tiles;
  const int* tile_perm;       // launch order -> logical tile
  const BlockDesc* blocks;
  const CaseParams* cps;
  double* arena;
  double* partials;           // [logical tile][2*NV]: max |Res|, sum Res^2
  unsigned long long* err;    // first cell with T < 1e-5 (block<<40 | j<<20 | i), ~0 if none
  int* cold;                  // per case, in per-case runs: set for a cell with T < 1e-5 (in place of err)
  MeshOpts o;
  PlaneMap pm;
  int clip;                   // apply force_vt_kw to interior nu_t on load
  double* sk_out;             // or null: the changing inputs of the fused LU-SGS in hyperplane order (the
                              // copy k_lusgs_skew makes), written here from the values the epilogue holds
  int sk_only;                // with sk_out: Res, dt, Lci, Lcj, Lvi, Lvj go to the hyperplane copy alone,
                              // not to their planes (k_lusgs_skew's reverse copy restores the planes)
  int write_visc;             // store Amu / Amu_t planes
  int force;                  // coarse multigrid levels, interior cells, variables 1..4, after the norms' values:
                              // 1: QF = QF - Res (comput_force_function), 2: Res = Res + QF (Add_force_function)
  int set_un;                 // also Un = U, variables 1..4, cells 0..nx x 0..ny (Set_Un)
  int closure;                // 1 (SA): take mu_t of interior cells from the closure planes where in range;
                              // 2 (4 equations): mu_t of every ring cell from the closure plane;
                              // 3 (4 equations, Baldwin-Lomax): mu_t of every ring cell from Amu_t as k_bl left it
};

// FIX: the level's scheme, flux and reconstruction are MUSCL3 with the van
// Albada limiter on both face directions, Roe and the original (primitive)
// reconstruction, as compile-time constants, so that the compiler drops the
// other branches of every face; the launch picks it when the options match.
template <int TX, int TY, int NV, bool VISC, bool SA, bool FIX>
__global__ __launch_bounds__(TX* TY) void k_residual(ResArgs a) {
  constexpr int RW = TX + 4, RH = TY + 4, NT = TX * TY;
  const int o_scheme = FIX ? 5 : a.o.scheme, o_gscheme = FIX ? 5 : a.o.global_scheme;
  const int o_flux = FIX ? 4 : a.o.flux, o_recon = FIX ? 0 : a.o.recon;
  // correctly rounded reciprocals of the constant divisors; div_rcp with
  // them gives the reference's quotients (x / Pr, x / Pr_t, x / sigma, x / k^2)
  constexpr double yPrt = 1.0 / EC2D_PRT, ySig = 1.0 / (2.0 / 3.0), yK2 = 1.0 / (0.41 * 0.41);
  constexpr int NFI = (TX + 1) * TY, NFJ = TX * (TY + 1);
  // viscous cases stage T and rebuild p = p00 d T from it (the staging's own
  // expression, so the same value); the cell centres of the viscous
  // gradients are read from their planes, not staged
  __shared__ double sd[RH][RW], su[RH][RW], sv[RH][RW], sp[VISC ? 1 : RH][VISC ? 1 : RW];
  __shared__ double sT[VISC ? RH : 1][VISC ? RW : 1];
  __shared__ double sAm[VISC ? RH : 1][VISC ? RW : 1], sAt[VISC ? RH : 1][VISC ? RW : 1];
  __shared__ double svt[SA ? RH : 1][SA ? RW : 1];
  // face fluxes: the i-faces' (sFi, SA: sFv) and, after every cell has taken
  // its i-face difference, the j-faces' (sFj, sFw) in the same LDS, so that
  // four workgroups of a 32 x 8 SA tile fit on a CU
  constexpr int NIF = 4 * TY * (TX + 1) + (SA ? 2 * TY * (TX + 1) : 1);
  constexpr int NJF = 4 * (TY + 1) * TX + (SA ? 2 * (TY + 1) * TX : 1);
  __shared__ double sF[NIF > NJF ? NIF : NJF];
  auto& sFi = *reinterpret_cast<double(*)[4][TY][TX + 1]>(sF);
  auto& sFv = *reinterpret_cast<double(*)[SA ? 2 : 1][SA ? TY : 1][SA ? TX + 1 : 1]>(sF + 4 * TY * (TX + 1));
  auto& sFj = *reinterpret_cast<double(*)[4][TY + 1][TX]>(sF);
  auto& sFw = *reinterpret_cast<double(*)[SA ? 2 : 1][SA ? TY + 1 : 1][SA ? TX : 1]>(sF + 4 * (TY + 1) * TX);
  __shared__ double sred[2 * NV][NT / 64];

  const int lt = a.tile_perm[blockIdx.x];
  const TileDesc t = a.tiles[lt];
  const BlockDesc B = a.blocks[t.block];
  if (case_frozen(B.case_id)) return;
  const CaseParams cp = a.cps[B.case_id];
  const double yPr = 1.0 / cp.Pr;
  const PlaneMap pm = a.pm;
  const int nx = B.nx, ny = B.ny;
  const double g = cp.gamma;
  double* base = a.arena + B.base;
  auto PL = [&](int slot) { return base + (int64_t)slot * B.plane; };
  const int tid = threadIdx.x;
  const int i0 = t.i0, j0 = t.j0;
  const double* U[NV];
#pragma unroll
  for (int m = 0; m < NV; m++) U[m] = PL(pm.U + m);
  // staged values of local cell (y, x): pressure, and the cell centres
  const double* X1p = PL(pm.x1);
  const double* Y1p = PL(pm.y1);
  auto SX = [&](int y, int x) { return X1p[pidx(B, i0 - 2 + x, j0 - 2 + y)]; };
  auto SY = [&](int y, int x) { return Y1p[pidx(B, i0 - 2 + x, j0 - 2 + y)]; };
  auto SP = [&](int y, int x) -> double {
    if constexpr (VISC) return cp.p00 * sd[y][x] * sT[y][x];
    else return sp[y][x];
  };
  auto SD = [&](int y, int x) { return sd[y][x]; };
  auto SU = [&](int y, int x) { return su[y][x]; };
  auto SV = [&](int y, int x) { return sv[y][x]; };

  // ---- 1. stage tile + halo ------------------------------------------------
  for (int idx = tid; idx < RW * RH; idx += NT) {
    int li = idx % RW, lj = idx / RW;
    int i = i0 - 2 + li, j = j0 - 2 + lj;
    if (i < -1 || i > nx + 1 || j < -1 || j > ny + 1) continue;
    int64_t q = pidx(B, i, j);
    double d = U[0][q];
    double uu = U[1][q] / d;
    double vv = U[2][q] / d;
    double T = (U[3][q] - 0.5 * d * (uu * uu + vv * vv)) / (cp.Cv * d);
    sd[lj][li] = d;
    su[lj][li] = uu;
    sv[lj][li] = vv;
    if constexpr (!VISC) sp[lj][li] = cp.p00 * d * T;
    bool interior = (i >= 1 && i <= nx - 1 && j >= 1 && j <= ny - 1);
    if (interior && T < 1.0e-5) {
      if (a.cold) {
        atomicOr(a.cold + B.case_id, 1);
      } else {
        unsigned long long key = ((unsigned long long)t.block << 40) | ((unsigned long long)j << 20) | (unsigned)i;
        atomicMin(a.err, key);
      }
    }
    if constexpr (VISC) {
      sT[lj][li] = T;
      bool ring = (i >= 0 && i <= nx && j >= 0 && j <= ny);
      double am = 0.0;
      if (ring) am = 1.0 / cp.Re * (1.0 + cp.Tsb) * sqrt(T * T * T) / (cp.Tsb + T);
      // SST: the model scaled interior Amu by Re and back (sub_turbulence_SST.f90:85, 259)
      if constexpr (NV == 6) {
        if (interior) am = (am * cp.Re) / cp.Re;
      }
      sAm[lj][li] = am;
      double at = 0.0;
      if constexpr (SA) {
        double vt = U[4][q];
        if (a.clip && interior && vt < 0) {
          vt = 1.0e-10;
          const_cast<double*>(U[4])[q] = vt;
        }
        svt[lj][li] = vt;
        if (ring) {
          const double Cv1_3 = 7.1 * 7.1 * 7.1;
          double X = d * vt / am;
          double X3 = X * X * X;
          double fv1 = X3 / (X3 + Cv1_3);
          at = fv1 * d * vt;
        }
        if (a.closure && interior && PL(pm.clo + 1)[q] != 0.0) at = PL(pm.clo)[q];
      }
      // SST: mu_t as the model left it (interior and wall ghosts set by
      // k_sst_res, other ghosts untouched)
      if constexpr (NV == 6) at = ring ? PL(pm.Amu_t)[q] : 0.0;
      // the learned closure as the eddy-viscosity model
      if constexpr (NV == 4) {
        if (a.closure == 2 && ring) at = PL(pm.clo)[q];
        // Baldwin-Lomax (k_bl wrote every cell, ghosts included, by its own rules)
        if (a.closure == 3 && ring) at = PL(pm.Amu_t)[q];
      }
      sAt[lj][li] = at;
    }
  }
  __syncthreads();
  if (SA || (NV == 4 && VISC && a.closure == 2)) {
    // Amut_boundary: wall ghost cells of layer 0 take -Amu_t of the adjacent interior cell
    // (sub_turbulence_SA.f90:200-216; the closure's mu_t follows the same rule)
    const double* flag = PL(pm.flag);
    for (int idx = tid; idx < RW * RH; idx += NT) {
      int li = idx % RW, lj = idx / RW;
      int i = i0 - 2 + li, j = j0 - 2 + lj;
      if (i < 0 || i > nx || j < 0 || j > ny) continue;
      bool gi = (i == 0 || i == nx) && j >= 1 && j <= ny - 1;
      bool gj = (j == 0 || j == ny) && i >= 1 && i <= nx - 1;
      if (!(gi || gj)) continue;
      if (flag[pidx(B, i, j)] == 0.0) continue;
      int ii = li, jj = lj;
      if (i == 0 && gi) ii = li + 1;
      else if (gi) ii = li - 1;
      else if (j == 0) jj = lj + 1;
      else jj = lj - 1;
      sAt[lj][li] = -sAt[jj][ii];
    }
    __syncthreads();
  }

  const bool fdm = B.fdm != 0;
  const int LF = 4;
  // ---- 2a. i-faces ---------------------------------------------------------
  for (int f = tid; f < NFI; f += NT) {
    int fi = f % (TX + 1), fj = f / (TX + 1);
    int i = i0 + fi, j = j0 + fj;
    if (i > nx || j > ny - 1) continue;
    // faces of the finite-difference interior carry no FVM flux, but the SA
    // model's own face loops (sub_turbulence_SA.f90:43-73) cover every face
    const bool skip_fvm = fdm && (i >= LF + 2 && i <= nx - LF - 1) && (j >= LF + 2 && j <= ny - LF - 1);
    if (skip_fvm && !SA) continue;
    int64_t q = pidx(B, i, j);
    double si = PL(pm.si)[q], ni1 = PL(pm.ni1)[q], ni2 = PL(pm.ni2)[q];
    int lj = fj + 2, li = fi + 2;  // local index of cell (i, j); stencil li-2 .. li+1
    double UL[4], UR[4];
    double F1 = 0.0, F2 = 0.0, F3 = 0.0, F4 = 0.0;
    if (!skip_fvm) {
    if (o_recon == 0) {
#define RECON(ARR, m)                                                                                           \
  UL[m] = scheme_fP(ARR(lj, li - 2), ARR(lj, li - 1), ARR(lj, li), ARR(lj, li + 1), o_scheme);              \
  UR[m] = scheme_fm(ARR(lj, li - 2), ARR(lj, li - 1), ARR(lj, li), ARR(lj, li + 1), o_scheme);
      RECON(SD, 0) RECON(SU, 1) RECON(SV, 2) RECON(SP, 3)
#undef RECON
    } else {
      double U0[4][4];
      for (int k = 0; k < 4; k++) {
        int64_t qk = pidx(B, i - 2 + k, j);
        for (int m = 0; m < 4; m++) U0[k][m] = U[m][qk];
      }
      if (o_recon == 1) {
        double QL[4], QR[4];
        for (int m = 0; m < 4; m++) {
          QL[m] = scheme_fP(U0[0][m], U0[1][m], U0[2][m], U0[3][m], o_scheme);
          QR[m] = scheme_fm(U0[0][m], U0[1][m], U0[2][m], U0[3][m], o_scheme);
        }
        cons_to_prim(QL, UL, g);
        cons_to_prim(QR, UR, g);
      } else {
        recon_characteristic(U0, UL, UR, g, o_scheme);
      }
    }
    double QL[4], QR[4], F0[4];
    QL[0] = UL[0]; QL[1] = UL[1] * ni1 + UL[2] * ni2; QL[2] = -UL[1] * ni2 + UL[2] * ni1; QL[3] = UL[3];
    QR[0] = UR[0]; QR[1] = UR[1] * ni1 + UR[2] * ni2; QR[2] = -UR[1] * ni2 + UR[2] * ni1; QR[3] = UR[3];
    face_flux(o_flux, QL, QR, F0, g);
    F1 = -F0[0] * si;
    F2 = -(F0[1] * ni1 - F0[2] * ni2) * si;
    F3 = -(F0[1] * ni2 + F0[2] * ni1) * si;
    F4 = -F0[3] * si;
    }
    if constexpr (VISC) {
      double Amu1, Amk1;
      if (i == 1) {
        Amu1 = sAm[lj][li] + sAt[lj][li];
        Amk1 = cp.Cp * (div_rcp(sAm[lj][li], cp.Pr, yPr) + div_rcp(sAt[lj][li], EC2D_PRT, yPrt));
      } else if (i == nx) {
        Amu1 = sAm[lj][li - 1] + sAt[lj][li - 1];
        Amk1 = cp.Cp * (div_rcp(sAm[lj][li - 1], cp.Pr, yPr) + div_rcp(sAt[lj][li - 1], EC2D_PRT, yPrt));
      } else {
        Amu1 = (sAm[lj][li - 1] + sAm[lj][li] + sAt[lj][li - 1] + sAt[lj][li]) * 0.5;
        Amk1 = cp.Cp * (div_rcp(sAm[lj][li - 1] + sAm[lj][li], cp.Pr, yPr) + div_rcp(sAt[lj][li - 1] + sAt[lj][li], EC2D_PRT, yPrt)) * 0.5;
      }
      double Dix = SX(lj, li) - SX(lj, li - 1);
      double Diy = SY(lj, li) - SY(lj, li - 1);
      double Djx = (SX(lj + 1, li - 1) + SX(lj + 1, li) - SX(lj - 1, li - 1) - SX(lj - 1, li)) * 0.25;
      double Djy = (SY(lj + 1, li - 1) + SY(lj + 1, li) - SY(lj - 1, li - 1) - SY(lj - 1, li)) * 0.25;
      double Ds = 1.0 / (Dix * Djy - Djx * Diy);
      double Diu = su[lj][li] - su[lj][li - 1];
      double Div = sv[lj][li] - sv[lj][li - 1];
      double DiT = sT[lj][li] - sT[lj][li - 1];
      double Dju = (su[lj + 1][li - 1] + su[lj + 1][li] - su[lj - 1][li - 1] - su[lj - 1][li]) * 0.25;
      double Djv = (sv[lj + 1][li - 1] + sv[lj + 1][li] - sv[lj - 1][li - 1] - sv[lj - 1][li]) * 0.25;
      double DjT = (sT[lj + 1][li - 1] + sT[lj + 1][li] - sT[lj - 1][li - 1] - sT[lj - 1][li]) * 0.25;
      double ux = (Diu * Djy - Dju * Diy) * Ds, vx = (Div * Djy - Djv * Diy) * Ds, Tx = (DiT * Djy - DjT * Diy) * Ds;
      double uy = (-Diu * Djx + Dju * Dix) * Ds, vy = (-Div * Djx + Djv * Dix) * Ds, Ty = (-DiT * Djx + DjT * Dix) * Ds;
      double u1 = (su[lj][li] + su[lj][li - 1]) * 0.5;
      double v1 = (sv[lj][li] + sv[lj][li - 1]) * 0.5;
      double t11 = ((4.0 / 3.0) * ux - (2.0 / 3.0) * vy) * Amu1;
      double t22 = ((4.0 / 3.0) * vy - (2.0 / 3.0) * ux) * Amu1;
      double t12 = (uy + vx) * Amu1;
      double E1 = u1 * t11 + v1 * t12 + Amk1 * Tx;
      double E2 = u1 * t12 + v1 * t22 + Amk1 * Ty;
      if (!skip_fvm) {
        F2 = F2 + (t11 * ni1 + t12 * ni2) * si;
        F3 = F3 + (t12 * ni1 + t22 * ni2) * si;
        F4 = F4 + (E1 * ni1 + E2 * ni2) * si;
      }
      if constexpr (SA) {
        const double Cb2 = 0.622, SA_sigma = 2.0 / 3.0;
        const double c_diff = 0.5 * (1.0 + Cb2) / SA_sigma;
        double vn1 = su[lj][li - 1] * ni1 + sv[lj][li - 1] * ni2;
        double vn2 = su[lj][li] * ni1 + sv[lj][li] * ni2;
        double vfi = -0.5 * ((vn1 + fabs(vn1)) * svt[lj][li - 1] + (vn2 - fabs(vn2)) * svt[lj][li]) * si;
        double v0 = c_diff * (svt[lj][li] + sAm[lj][li] / sd[lj][li] + svt[lj][li - 1] + sAm[lj][li - 1] / sd[lj][li - 1]);
        double Dv = svt[lj][li] - svt[lj][li - 1];
        double Djvt = (svt[lj + 1][li - 1] + svt[lj + 1][li] - svt[lj - 1][li - 1] - svt[lj - 1][li]) * 0.25;
        double vtx = (Dv * Djy - Djvt * Diy) * Ds;
        double vty = (-Dv * Djx + Djvt * Dix) * Ds;
        sFv[0][fj][fi] = vfi + v0 * (vtx * ni1 + vty * ni2) * si;
        sFv[1][fj][fi] = (vtx * ni1 + vty * ni2) * si;
      }
    }
    sFi[0][fj][fi] = F1;
    sFi[1][fj][fi] = F2;
    sFi[2][fj][fi] = F3;
    sFi[3][fj][fi] = F4;
  }
  __syncthreads();

  // The cell of this thread in phases 3a and 3: row-major, or, when the
  // epilogue also writes the hyperplane-order copy (sk_out), along the
  // tile's hyperplanes: TY consecutive lanes take the TY rows of one
  // hyperplane, so that the copy's stores form runs of TY positions.
  const int kd = tid / TY, cjd = tid % TY;
  const int cell_i = a.sk_out ? (kd - cjd + TX) % TX : tid % TX;
  const int cell_j = a.sk_out ? cjd : tid / TX;

  // ---- 3a. each cell's i-face difference (the first terms of Res, in the
  // reference's order), before the j-faces take over the LDS --------------
  double R[NV];
  double r5 = 0.0, v0c = 0.0;  // SA: the cell's v0 of sub_turbulence_SA.f90, kept for phase 3
  {
    const int ci = cell_i, cj = cell_j;
    const int i = i0 + ci, j = j0 + cj;
    if (i <= nx - 1 && j <= ny - 1) {
      const bool fdm_cell = fdm && (i >= LF + 1 && i <= nx - LF - 1) && (j >= LF + 1 && j <= ny - LF - 1);
      if (!fdm_cell) {
#pragma unroll
        for (int m = 0; m < 4; m++) R[m] = sFi[m][cj][ci + 1] - sFi[m][cj][ci];
      }
      if constexpr (SA) {
        const double SA_sigma = 2.0 / 3.0, Cb2 = 0.622;
        const int li = ci + 2, lj = cj + 2;
        v0c = div_rcp((svt[lj][li] + sAm[lj][li] / sd[lj][li]) * Cb2, SA_sigma, ySig);
        r5 = sFv[0][cj][ci + 1] - sFv[0][cj][ci] - v0c * (sFv[1][cj][ci + 1] - sFv[1][cj][ci]);
      }
    }
  }
  __syncthreads();
  // ---- 2b. j-faces ---------------------------------------------------------
  // NFJ = NT + 64: the 64 faces beyond one per thread go to the last wave
  // (the first wave already takes the TY i-faces beyond one per thread), so
  // that no wave evaluates more than three faces
  for (int f = (tid + 64) % NT; f < NFJ; f += NT) {
    int fi = f % TX, fj = f / TX;
    int i = i0 + fi, j = j0 + fj;
    if (i > nx - 1 || j > ny) continue;
    const bool skip_fvm = fdm && (i >= LF + 2 && i <= nx - LF - 1) && (j >= LF + 2 && j <= ny - LF - 1);
    if (skip_fvm && !SA) continue;
    int64_t q = pidx(B, i, j);
    double sj = PL(pm.sj)[q], nj1 = PL(pm.nj1)[q], nj2 = PL(pm.nj2)[q];
    int lj = fj + 2, li = fi + 2;
    double UL[4], UR[4];
    double F1 = 0.0, F2 = 0.0, F3 = 0.0, F4 = 0.0;
    if (!skip_fvm) {
    if (o_recon == 0) {
      // the j-direction original-variable reconstruction uses the global
      // scheme option, not the mesh's own (sub_Residual.f90:367)
#define RECON(ARR, m)                                                                                              \
  UL[m] = scheme_fP(ARR(lj - 2, li), ARR(lj - 1, li), ARR(lj, li), ARR(lj + 1, li), o_gscheme);          \
  UR[m] = scheme_fm(ARR(lj - 2, li), ARR(lj - 1, li), ARR(lj, li), ARR(lj + 1, li), o_gscheme);
      RECON(SD, 0) RECON(SU, 1) RECON(SV, 2) RECON(SP, 3)
#undef RECON
    } else {
      double U0[4][4];
      for (int k = 0; k < 4; k++) {
        int64_t qk = pidx(B, i, j - 2 + k);
        for (int m = 0; m < 4; m++) U0[k][m] = U[m][qk];
      }
      if (o_recon == 1) {
        double QL[4], QR[4];
        for (int m = 0; m < 4; m++) {
          QL[m] = scheme_fP(U0[0][m], U0[1][m], U0[2][m], U0[3][m], o_scheme);
          QR[m] = scheme_fm(U0[0][m], U0[1][m], U0[2][m], U0[3][m], o_scheme);
        }
        cons_to_prim(QL, UL, g);
        cons_to_prim(QR, UR, g);
      } else {
        recon_characteristic(U0, UL, UR, g, o_scheme);
      }
    }
    double QL[4], QR[4], F0[4];
    QL[0] = UL[0]; QL[1] = UL[1] * nj1 + UL[2] * nj2; QL[2] = -UL[1] * nj2 + UL[2] * nj1; QL[3] = UL[3];
    QR[0] = UR[0]; QR[1] = UR[1] * nj1 + UR[2] * nj2; QR[2] = -UR[1] * nj2 + UR[2] * nj1; QR[3] = UR[3];
    face_flux(o_flux, QL, QR, F0, g);
    F1 = -F0[0] * sj;
    F2 = -(F0[1] * nj1 - F0[2] * nj2) * sj;
    F3 = -(F0[1] * nj2 + F0[2] * nj1) * sj;
    F4 = -F0[3] * sj;
    }
    if constexpr (VISC) {
      double Amu1, Amk1;
      if (j == 1) {
        Amu1 = sAm[lj][li] + sAt[lj][li];
        Amk1 = cp.Cp * (div_rcp(sAm[lj][li], cp.Pr, yPr) + div_rcp(sAt[lj][li], EC2D_PRT, yPrt));
      } else if (j == ny) {
        Amu1 = sAm[lj - 1][li] + sAt[lj - 1][li];
        Amk1 = cp.Cp * (div_rcp(sAm[lj - 1][li], cp.Pr, yPr) + div_rcp(sAt[lj - 1][li], EC2D_PRT, yPrt));
      } else {
        Amu1 = (sAm[lj][li] + sAm[lj - 1][li] + sAt[lj][li] + sAt[lj - 1][li]) * 0.5;
        Amk1 = cp.Cp * (div_rcp(sAm[lj - 1][li] + sAm[lj][li], cp.Pr, yPr) + div_rcp(sAt[lj - 1][li] + sAt[lj][li], EC2D_PRT, yPrt)) * 0.5;
      }
      double Dix = (SX(lj - 1, li + 1) + SX(lj, li + 1) - SX(lj - 1, li - 1) - SX(lj, li - 1)) * 0.25;
      double Diy = (SY(lj - 1, li + 1) + SY(lj, li + 1) - SY(lj - 1, li - 1) - SY(lj, li - 1)) * 0.25;
      double Djx = SX(lj, li) - SX(lj - 1, li);
      double Djy = SY(lj, li) - SY(lj - 1, li);
      double Ds = 1.0 / (Dix * Djy - Djx * Diy);
      double Diu = (su[lj - 1][li + 1] + su[lj][li + 1] - su[lj - 1][li - 1] - su[lj][li - 1]) * 0.25;
      double Div = (sv[lj - 1][li + 1] + sv[lj][li + 1] - sv[lj - 1][li - 1] - sv[lj][li - 1]) * 0.25;
      double DiT = (sT[lj - 1][li + 1] + sT[lj][li + 1] - sT[lj - 1][li - 1] - sT[lj][li - 1]) * 0.25;
      double Dju = su[lj][li] - su[lj - 1][li];
      double Djv = sv[lj][li] - sv[lj - 1][li];
      double DjT = sT[lj][li] - sT[lj - 1][li];
      double ux = (Diu * Djy - Dju * Diy) * Ds, vx = (Div * Djy - Djv * Diy) * Ds, Tx = (DiT * Djy - DjT * Diy) * Ds;
      double uy = (-Diu * Djx + Dju * Dix) * Ds, vy = (-Div * Djx + Djv * Dix) * Ds, Ty = (-DiT * Djx + DjT * Dix) * Ds;
      double t11 = ((4.0 / 3.0) * ux - (2.0 / 3.0) * vy) * Amu1;
      double t22 = ((4.0 / 3.0) * vy - (2.0 / 3.0) * ux) * Amu1;
      double t12 = (uy + vx) * Amu1;
      double u1 = (su[lj][li] + su[lj - 1][li]) * 0.5;
      double v1 = (sv[lj][li] + sv[lj - 1][li]) * 0.5;
      double E1 = u1 * t11 + v1 * t12 + Amk1 * Tx;
      double E2 = u1 * t12 + v1 * t22 + Amk1 * Ty;
      if (!skip_fvm) {
        F2 = F2 + (t11 * nj1 + t12 * nj2) * sj;
        F3 = F3 + (t12 * nj1 + t22 * nj2) * sj;
        F4 = F4 + (E1 * nj1 + E2 * nj2) * sj;
      }
      if constexpr (SA) {
        const double Cb2 = 0.622, SA_sigma = 2.0 / 3.0;
        const double c_diff = 0.5 * (1.0 + Cb2) / SA_sigma;
        double vn1 = su[lj - 1][li] * nj1 + sv[lj - 1][li] * nj2;
        double vn2 = su[lj][li] * nj1 + sv[lj][li] * nj2;
        double vfi = -0.5 * ((vn1 + fabs(vn1)) * svt[lj - 1][li] + (vn2 - fabs(vn2)) * svt[lj][li]) * sj;
        double v0 = c_diff * (svt[lj][li] + sAm[lj][li] / sd[lj][li] + svt[lj - 1][li] + sAm[lj - 1][li] / sd[lj - 1][li]);
        double Div2 = (svt[lj - 1][li + 1] + svt[lj][li + 1] - svt[lj - 1][li - 1] - svt[lj][li - 1]) * 0.25;
        double Djv2 = svt[lj][li] - svt[lj - 1][li];
        double vtx = (Div2 * Djy - Djv2 * Diy) * Ds;
        double vty = (-Div2 * Djx + Djv2 * Dix) * Ds;
        sFw[0][fj][fi] = vfi + v0 * (vtx * nj1 + vty * nj2) * sj;
        sFw[1][fj][fi] = (vtx * nj1 + vty * nj2) * sj;
      }
    }
    sFj[0][fj][fi] = F1;
    sFj[1][fj][fi] = F2;
    sFj[2][fj][fi] = F3;
    sFj[3][fj][fi] = F4;
  }
  __syncthreads();

  // ---- 3. cells ------------------------------------------------------------
  double rmax[NV], rsum[NV];
#pragma unroll
  for (int m = 0; m < NV; m++) { rmax[m] = 0.0; rsum[m] = 0.0; }
  {
    int ci = cell_i, cj = cell_j;
    int i = i0 + ci, j = j0 + cj;
    if (i <= nx - 1 && j <= ny - 1) {
      int64_t q = pidx(B, i, j);
      int li = ci + 2, lj = cj + 2;
      bool fdm_cell = fdm && (i >= LF + 1 && i <= nx - LF - 1) && (j >= LF + 1 && j <= ny - LF - 1);
      if (!fdm_cell) {
#pragma unroll
        for (int m = 0; m < 4; m++) R[m] = R[m] + sFj[m][cj + 1][ci] - sFj[m][cj][ci];
      }
      if constexpr (SA) {
        const double SA_sigma = 2.0 / 3.0, Cv1 = 7.1, Cb1 = 0.1355, Cb2 = 0.622, SA_k = 0.41;
        const double Cw1 = Cb1 / (SA_k * SA_k) + (1.0 + Cb2) / SA_sigma, Cw2 = 0.3, Ct3 = 1.3, Ct4 = 0.5;
        const double Cv1_3 = Cv1 * Cv1 * Cv1, Cw3_6 = 2.0 * 2.0 * 2.0 * 2.0 * 2.0 * 2.0;
        double vtc = svt[lj][li];
        const double v0 = v0c;
        r5 = r5 + sFw[0][cj + 1][ci] - sFw[0][cj][ci] - v0 * (sFw[1][cj + 1][ci] - sFw[1][cj][ci]);
        double Dix = (SX(lj, li + 1) - SX(lj, li - 1)) * 0.5;
        double Diy = (SY(lj, li + 1) - SY(lj, li - 1)) * 0.5;
        double Djx = (SX(lj + 1, li) - SX(lj - 1, li)) * 0.5;
        double Djy = (SY(lj + 1, li) - SY(lj - 1, li)) * 0.5;
        double Ds = 1.0 / (Dix * Djy - Djx * Diy);
        double Diu = (su[lj][li + 1] - su[lj][li - 1]) * 0.5;
        double Div = (sv[lj][li + 1] - sv[lj][li - 1]) * 0.5;
        double Dju = (su[lj + 1][li] - su[lj - 1][li]) * 0.5;
        double Djv = (sv[lj + 1][li] - sv[lj - 1][li]) * 0.5;
        double vx = (Div * Djy - Djv * Diy) * Ds;
        double uy = (-Diu * Djx + Dju * Dix) * Ds;
        double S = fabs(vx - uy);
        double dwc = PL(pm.dw)[q];
        double X = sd[lj][li] * vtc / sAm[lj][li];
        double X3 = X * X * X;
        double fv1 = X3 / (X3 + Cv1_3);
        double fv2 = 1.0 - X / (1.0 + X * fv1);
        double kd = SA_k * dwc;
        double S1 = fmax(S + fv2 * vtc / (kd * kd), EC2D_SP(0.3) * S);
        double r = fmin(vtc / (S1 * SA_k * SA_k * dwc * dwc), 10.0);
        double gg = r + Cw2 * (r * r * r * r * r * r - r);
        double fw = gg * pow((1.0 + Cw3_6) / (gg * gg * gg * gg * gg * gg + Cw3_6), 1.0 / 6.0);
        double ft2 = Ct3 * exp(-(Ct4 * X * X));
        double qv = vtc / dwc;
        double Q_SA = Cb1 * (1.0 - ft2) * S1 * vtc - (Cw1 * fw - div_rcp(Cb1 * ft2, SA_k * SA_k, yK2)) * (qv * qv);
        r5 = r5 + Q_SA * PL(pm.vol)[q];
        R[4] = r5;
      }
      double Rs[NV];  // Res as stored (with the forcing function on coarse levels)
      if (!fdm_cell) {
#pragma unroll
        for (int m = 0; m < 4; m++) {
          double r = R[m];
          if (a.force == 1) PL(pm.QF + m)[q] = PL(pm.QF + m)[q] - r;
          else if (a.force == 2) r = r + PL(pm.QF + m)[q];
          if (!a.sk_only) PL(pm.Res + m)[q] = r;
          Rs[m] = r;
        }
      }
      if constexpr (SA) {
        if (!a.sk_only) PL(pm.Res + 4)[q] = R[4];
      }
      if constexpr (NV == 6) {
        // Res(5:6) came from k_sst_res; read for the norms
        R[4] = PL(pm.Res + 4)[q];
        R[5] = PL(pm.Res + 5)[q];
      }
#pragma unroll
      for (int m = 4; m < NV; m++) Rs[m] = R[m];
      // comput_Lvc
      int64_t qi1 = pidx(B, i + 1, j), qj1 = pidx(B, i, j + 1);
      double s0 = PL(pm.vol)[q];
      double si = 0.5 * (PL(pm.si)[q] + PL(pm.si)[qi1]);
      double ni1 = 0.5 * (PL(pm.ni1)[q] + PL(pm.ni1)[qi1]);
      double ni2 = 0.5 * (PL(pm.ni2)[q] + PL(pm.ni2)[qi1]);
      double sj = 0.5 * (PL(pm.sj)[q] + PL(pm.sj)[qj1]);
      double nj1 = 0.5 * (PL(pm.nj1)[q] + PL(pm.nj1)[qj1]);
      double nj2 = 0.5 * (PL(pm.nj2)[q] + PL(pm.nj2)[qj1]);
      double uni = su[lj][li] * ni1 + sv[lj][li] * ni2;
      double unj = su[lj][li] * nj1 + sv[lj][li] * nj2;
      double T = (U[3][q] - 0.5 * sd[lj][li] * (su[lj][li] * su[lj][li] + sv[lj][li] * sv[lj][li])) / (cp.Cv * sd[lj][li]);
      double cc = sqrt(T) / cp.Ma;
      double Lci = (fabs(uni) + cc) * si;
      double Lcj = (fabs(unj) + cc) * sj;
      double Lvi = 0.0, Lvj = 0.0;
      if constexpr (VISC) {
        double tmp1 = g / sd[lj][li] * (div_rcp(sAm[lj][li], cp.Pr, yPr) + div_rcp(sAt[lj][li], EC2D_PRT, yPrt));
        Lvi = tmp1 * si * si / s0;
        Lvj = tmp1 * sj * sj / s0;
      }
      if (!a.sk_only) {
        if constexpr (VISC) {
          PL(pm.Lvi)[q] = Lvi;
          PL(pm.Lvj)[q] = Lvj;
        }
        PL(pm.Lci)[q] = Lci;
        PL(pm.Lcj)[q] = Lcj;
      }
      double dt;
      if (cp.Iflag_local_dt == 0) {
        dt = cp.dt_global;
      } else {
        dt = cp.CFL * s0 / (Lci + Lcj + 1.0 * (Lvi + Lvj));
        if (dt > cp.dtmax) dt = cp.dtmax;
        if (dt < cp.dtmin) dt = cp.dtmin;
      }
      if (!a.sk_only) PL(pm.dt)[q] = dt;
      if (a.sk_out) {
        // logical plane k of cell (i, j) at k n + (i-1 + j-1) R + j-1 (the
        // layout of k_lusgs_skew): U (nu_t as clipped in the staging), Res,
        // dt, Lci, Lcj and, viscous, Lvi, Lvj
        using I = RecIn;
        double* so = a.sk_out + B.sk_base + (int64_t)(i - 1 + j - 1) * B.sk_R + (j - 1);
        const int64_t n = B.sk_n;
#pragma unroll
        for (int m = 0; m < NV; m++) {
          double u = U[m][q];
          if constexpr (SA) {
            if (m == 4) u = svt[lj][li];
          }
          so[(I::U + m) * n] = u;
          so[(I::RES<NV> + m) * n] = Rs[m];
        }
        so[I::DT<NV> * n] = dt;
        so[I::LCI<NV> * n] = Lci;
        so[I::LCJ<NV> * n] = Lcj;
        if constexpr (VISC) {
          so[I::LVI<NV> * n] = Lvi;
          so[I::LVJ<NV> * n] = Lvj;
        }
      }
      if (!fdm_cell) {
#pragma unroll
        for (int m = 0; m < NV; m++) {
          rmax[m] = fabs(R[m]);
          rsum[m] = R[m] * R[m];
        }
      } else {
#pragma unroll
        for (int m = 4; m < NV; m++) {
          rmax[m] = fabs(R[m]);
          rsum[m] = R[m] * R[m];
        }
      }
      if constexpr (VISC) {
        if (a.write_visc) {
          PL(pm.Amu)[q] = sAm[lj][li];
          PL(pm.Amu_t)[q] = sAt[lj][li];
        }
      }
    }
  }
  if constexpr (VISC) {
    // mu_t of the first ghost layer (i = 0 or nx, j = 0 or ny, corners
    // included), as the reference's array holds it after Amut_boundary; the
    // output files carry it.  A ring cell belongs to the tile whose interior
    // holds the cell next to it (its indices clamped to 1..nx-1, 1..ny-1).
    // SST and Baldwin-Lomax read these cells from the plane above: the store
    // writes back the value read.
    if (a.write_visc) {
      for (int idx = tid; idx < RW * RH; idx += NT) {
        int li = idx % RW, lj = idx / RW;
        int i = i0 - 2 + li, j = j0 - 2 + lj;
        if (i < 0 || i > nx || j < 0 || j > ny) continue;
        if (i != 0 && i != nx && j != 0 && j != ny) continue;
        int ci = min(max(i, 1), nx - 1), cj = min(max(j, 1), ny - 1);
        if (ci < i0 || ci >= i0 + TX || cj < j0 || cj >= j0 + TY) continue;
        PL(pm.Amu_t)[pidx(B, i, j)] = sAt[lj][li];
      }
    }
  }
  if (a.set_un) {
    // Set_Un over the cells and the first ghost layer; a ghost cell belongs
    // to the tile that holds the interior cell next to it
    for (int idx = tid; idx < RW * RH; idx += NT) {
      int li = idx % RW, lj = idx / RW;
      int i = i0 - 2 + li, j = j0 - 2 + lj;
      if (i < 0 || i > nx || j < 0 || j > ny) continue;
      int ci = min(max(i, 1), nx - 1), cj = min(max(j, 1), ny - 1);
      if (ci < i0 || ci >= i0 + TX || cj < j0 || cj >= j0 + TY) continue;
      int64_t q = pidx(B, i, j);
#pragma unroll
      for (int m = 0; m < 4; m++) PL(pm.Un + m)[q] = U[m][q];
    }
  }
  // ---- tile reduction (fixed tree: wave butterfly, then waves in order) ----
#pragma unroll
  for (int m = 0; m < NV; m++) {
    double mx = rmax[m], sm = rsum[m];
    for (int off = 32; off > 0; off >>= 1) {
      mx = fmax(mx, __shfl_xor(mx, off));
      sm = sm + __shfl_xor(sm, off);
    }
    if ((tid & 63) == 0) {
      sred[2 * m][tid >> 6] = mx;
      sred[2 * m + 1][tid >> 6] = sm;
    }
  }
  __syncthreads();
  if (tid < 2 * NV) {
    int m = tid >> 1;
    double acc = sred[tid][0];
    for (int w = 1; w < NT / 64; w++) acc = (tid & 1) ? acc + sred[tid][w] : fmax(acc, sred[tid][w]);
    a.partials[(int64_t)lt * 2 * NV + tid] = acc;
    (void)m;
  }
}

// Deterministic second stage: one workgroup per case reduces its tiles in
// logical order into Res_max(1:NV) and Res_rms(1:NV) = sqrt(sum/Num_Cell).
template <int NV>
__global__ void k_norms(const double* partials, const int* case_tile0, const double* ncell, double* norms) {
  int c = blockIdx.x;
  int t0 = case_tile0[c], t1 = case_tile0[c + 1];
  __shared__ double sh[256];
  for (int k = 0; k < 2 * NV; k++) {
    double acc = 0.0;
    bool is_sum = k & 1;
    for (int t = t0 + threadIdx.x; t < t1; t += blockDim.x) {
      double v = partials[(int64_t)t * 2 * NV + k];
      acc = is_sum ? acc + v : fmax(acc, v);
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) sh[threadIdx.x] = is_sum ? sh[threadIdx.x] + sh[threadIdx.x + s] : fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      int m = k >> 1;
      if (is_sum) norms[c * 12 + 6 + m] = sqrt(sh[0] / ncell[c]);
      else norms[c * 12 + m] = sh[0];
    }
    __syncthreads();
  }
}

}  // namespace ec2d

// Run to a residual drop (ec2d_run_to_drop): after the first fine-level
// residual of every main-loop pass, one workgroup per case reduces the
// density's per-tile sums of Res^2 as k_norms does (the same order, so the
// value equals Res_rms(1) of ec2d_res_norms) and decides whether the case
// goes on.  The first check of a run records the reference value r0.  A case
// stops (active = 0, its blocks frozen by every later kernel) when its rms
// falls to drop * r0 (status 1), when the rms of any of its equations is not
// finite (status 2: the NaN guard), or
// has been checked max_iter times (status 3).  iters counts the checks, i.e.
// the residual evaluations of the case's state; the state it stops in has had
// iters - 1 updates in this run.
struct ConvState {
  double r0, last;
  int iters, status;
};

__global__ __launch_bounds__(256) void k_converge(const double* partials, const int* case_tile0, const double* ncell,
                                                  int nv, int* active, ConvState* st, double drop, int max_iter,
                                                  const int* cold) {
  const int c = blockIdx.x;
  if (active[c] == 0) return;
  const int t0 = case_tile0[c], t1 = case_tile0[c + 1];
  __shared__ double sh[256], sh_all[256];
  // acc: the density sum of squares; all: every equation's (non-negative
  // terms, so a NaN or an overflow in any equation leaves it non-finite)
  double acc = 0.0, all = 0.0;
  for (int t = t0 + threadIdx.x; t < t1; t += blockDim.x) {
    const double* pt = partials + (int64_t)t * 2 * nv;
    acc = acc + pt[1];
    for (int m = 0; m < nv; m++) all = all + pt[2 * m + 1];
  }
  sh[threadIdx.x] = acc;
  sh_all[threadIdx.x] = all;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + s];
      sh_all[threadIdx.x] = sh_all[threadIdx.x] + sh_all[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double rms = sqrt(sh[0] / ncell[c]);
    ConvState s = st[c];
    s.iters += 1;
    if (s.iters == 1) s.r0 = rms;
    s.last = rms;
    if (cold[c]) s.status = 4;  // the reference's T < 1e-5 stop, for this case only
    else if (!isfinite(rms) || !isfinite(sh_all[0])) s.status = 2;
    else if (rms <= drop * s.r0) s.status = 1;
    else if (s.iters >= max_iter) s.status = 3;
    if (s.status != 0) active[c] = 0;
    st[c] = s;
  }
}
