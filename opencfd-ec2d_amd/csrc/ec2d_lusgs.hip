// LU-SGS implicit update: du_LU_SGS_2D (sub_LU_SGS.f90:4-108) followed by
// U = U + dU (sub_NS_singlegrid.f90:203-211), for all blocks in one launch.
//
// The forward sweep needs dU(i-1,j) and dU(i,j-1); the backward sweep
// dU(i+1,j) and dU(i,j+1).  Any order that respects these dependencies gives
// the reference's result bit for bit.  Here one workgroup owns one block and
// one lane owns one row: at step s, lane r handles the cell of row r whose
// column is s-r, i.e. the workgroup walks the i+j hyperplanes of the
// reference (sub_LU_SGS.f90:25-29, 70-75) with every cell of a plane on its
// own lane.
//   * the state of the neighbour in i is the lane's own result of the
//     previous step (registers);
//   * the state of the neighbour in j is the previous result of lane r-1,
//     moved one lane up with a DPP wave shift; between 64-row waves it goes
//     through LDS with one barrier per step.
// The sweep is a chain of dependent steps, so its speed is set by the latency
// of one step.  Every operand a cell needs from memory is known in advance
// (its address depends only on the step), so each lane loads the operands of
// the cell it will handle D steps later into a ring of registers while it
// computes the current one: the loads are unconditional (indices are clamped
// into the block for steps outside it) so that they stay in flight across
// steps instead of being waited for at branch joins.
// The backward sweep runs the same schedule on reversed indices and adds dU
// to U in place; the old U values its neighbours still need travel in the
// carries, so U is read once and written once.
#include "ec2d_device.h"
#include "ec2d_physics.h"

namespace ec2d {

struct LusgsArgs {
  const BlockDesc* blocks;
  const int* block_list;  // blocks of this launch
  const CaseParams* cps;
  double* arena;
  PlaneMap pm;
  double alfa1;
  int update;             // 1: U += dU fused into the backward sweep
};

// state of one finished cell as its neighbours need it
template <int NV>
struct Carry {
  double dU[NV], U[NV];
  double s, Lc, Lv;  // face area, inviscid and viscous spectral radius in the carry's direction
};

// operands of one cell, loaded ahead of its step
template <int NV>
struct Pre {
  double U[NV], X[NV];  // X: Res in the forward sweep, the forward dU in the backward sweep
  double vol, dt, Lci, Lcj, Lvi, Lvj, si, sj, ni1, ni2, nj1, nj2;
};

// lane r-1 -> lane r (lane 0 receives 0)
__device__ __forceinline__ double shr1(double v) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_update_dpp(0, lo, 0x138, 0xF, 0xF, false);
  hi = __builtin_amdgcn_update_dpp(0, hi, 0x138, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

template <int NV>
__device__ __forceinline__ void shift_carry(Carry<NV>& dst, const Carry<NV>& src) {
#pragma unroll
  for (int m = 0; m < NV; m++) {
    dst.dU[m] = shr1(src.dU[m]);
    dst.U[m] = shr1(src.U[m]);
  }
  dst.s = shr1(src.s);
  dst.Lc = shr1(src.Lc);
  dst.Lv = shr1(src.Lv);
}

template <int NV>
__device__ __forceinline__ void put_carry(double* x, const Carry<NV>& c) {
#pragma unroll
  for (int m = 0; m < NV; m++) { x[m] = c.dU[m]; x[NV + m] = c.U[m]; }
  x[2 * NV] = c.s; x[2 * NV + 1] = c.Lc; x[2 * NV + 2] = c.Lv;
}

template <int NV>
__device__ __forceinline__ void get_carry(Carry<NV>& c, const double* x) {
#pragma unroll
  for (int m = 0; m < NV; m++) { c.dU[m] = x[m]; c.U[m] = x[NV + m]; }
  c.s = x[2 * NV]; c.Lc = x[2 * NV + 1]; c.Lv = x[2 * NV + 2];
}

template <int NV>
struct Planes {
  double* U[NV];
  double* dU[NV];
  const double* R[NV];
  const double *vol, *dt, *Lci, *Lcj, *Lvi, *Lvj, *si, *sj, *ni1, *ni2, *nj1, *nj2;
};

template <int NV, bool VISC, bool FWD>
__device__ __forceinline__ void load_pre(Pre<NV>& p, const Planes<NV>& P, const BlockDesc& B, int s, int r, int ncol,
                                         int nrow) {
  int i = FWD ? s - r + 1 : ncol - (s - r);
  int j = FWD ? r + 1 : nrow - r;
  i = min(max(i, 1), ncol);
  j = min(max(j, 1), nrow);
  const int64_t q = pidx(B, i, j);
#pragma unroll
  for (int m = 0; m < NV; m++) {
    p.U[m] = P.U[m][q];
    p.X[m] = FWD ? P.R[m][q] : P.dU[m][q];
  }
  p.vol = P.vol[q];
  p.dt = P.dt[q];
  p.Lci = P.Lci[q];
  p.Lcj = P.Lcj[q];
  if constexpr (VISC) {
    p.Lvi = P.Lvi[q];
    p.Lvj = P.Lvj[q];
  } else {
    p.Lvi = 0.0;
    p.Lvj = 0.0;
  }
  p.si = P.si[q];
  p.sj = P.sj[q];
  p.ni1 = P.ni1[q];
  p.ni2 = P.ni2[q];
  p.nj1 = P.nj1[q];
  p.nj2 = P.nj2[q];
}

// one cell of one sweep; returns through the carries
template <int NV, bool VISC, bool FWD>
__device__ __forceinline__ void lusgs_cell(const Pre<NV>& p, Carry<NV>& side, Carry<NV>& up, const Carry<NV>& low,
                                           bool act, bool has_side, bool has_low, double g, double alfa1,
                                           const Planes<NV>& P, int64_t q, int update) {
  // diagonal (sub_LU_SGS.f90:31-39, 76-83)
  double alfa = p.vol / p.dt + 1.0 * (p.Lci + p.Lcj) + p.vol * alfa1;
  if constexpr (VISC) alfa = alfa + 2.0 * (p.Lvi + p.Lvj);
  double al[NV];
#pragma unroll
  for (int m = 0; m < NV; m++) al[m] = alfa;
  if constexpr (NV == 5) al[4] = al[4] + p.Lvi + p.Lvj;
  double dui[NV], duj[NV], DFi[NV], DFj[NV];
  comput_dfn<NV>(DFi, side.U, side.dU, p.ni1, p.ni2, g);
  comput_dfn<NV>(DFj, low.U, low.dU, p.nj1, p.nj2, g);
  // forward: the areas are the cell's own si, sj; backward: the neighbour's
  const double ai = FWD ? p.si : side.s;
  const double aj = FWD ? p.sj : low.s;
#pragma unroll
  for (int m = 0; m < NV; m++) {
    double di, dj;
    if (FWD) {
      di = 0.5 * (DFi[m] * ai + 1.0 * side.Lc * side.dU[m]);
      dj = 0.5 * (DFj[m] * aj + 1.0 * low.Lc * low.dU[m]);
    } else {
      di = -0.5 * (DFi[m] * ai - 1.0 * side.Lc * side.dU[m]);
      dj = -0.5 * (DFj[m] * aj - 1.0 * low.Lc * low.dU[m]);
    }
    if constexpr (VISC) {
      di = di + side.Lv * side.dU[m];
      dj = dj + low.Lv * low.dU[m];
    }
    dui[m] = has_side ? di : 0.0;
    duj[m] = has_low ? dj : 0.0;
  }
  double du[NV];
#pragma unroll
  for (int m = 0; m < NV; m++) {
    if (FWD) du[m] = (p.X[m] + dui[m] + duj[m]) / al[m];
    else du[m] = p.X[m] + (dui[m] + duj[m]) / al[m];
  }
  if (act) {
#pragma unroll
    for (int m = 0; m < NV; m++) {
      P.dU[m][q] = du[m];
      if (!FWD && update) P.U[m][q] = p.U[m] + du[m];
    }
#pragma unroll
    for (int m = 0; m < NV; m++) {
      side.dU[m] = du[m];
      side.U[m] = p.U[m];
      up.dU[m] = du[m];
      up.U[m] = p.U[m];
    }
    side.s = p.si;
    side.Lc = p.Lci;
    side.Lv = p.Lvi;
    up.s = p.sj;
    up.Lc = p.Lcj;
    up.Lv = p.Lvj;
  }
}

template <int NV, bool VISC, bool FWD, int D>
__device__ __forceinline__ void sweep(const Planes<NV>& P, const BlockDesc& B, int ncol, int nrow, int r, bool row_ok,
                                      int lane, int wave, bool multi, double g, double alfa1, int update,
                                      double (*xfer)[16][2 * NV + 3]) {
  const int nsteps = ncol + nrow - 1;
  const int npad = (nsteps + D - 1) / D * D;  // steps past the last plane are empty for every lane
  Carry<NV> side = {}, up = {}, low = {};
  Pre<NV> ring[D];
#pragma unroll
  for (int d = 0; d < D; d++) load_pre<NV, VISC, FWD>(ring[d], P, B, d, r, ncol, nrow);
  for (int s0 = 0; s0 < npad; s0 += D) {
#pragma unroll
    for (int d = 0; d < D; d++) {
      const int s = s0 + d;
      shift_carry<NV>(low, up);
      if (multi && lane == 0 && wave > 0) get_carry<NV>(low, xfer[(s + 1) & 1][wave - 1]);
      const int i = FWD ? s - r + 1 : ncol - (s - r);
      const int j = FWD ? r + 1 : nrow - r;
      const bool act = row_ok && i >= 1 && i <= ncol;
      const bool has_side = FWD ? (i != 1) : (i != ncol);
      const bool has_low = FWD ? (j != 1) : (j != nrow);
      const int64_t q = pidx(B, min(max(i, 1), ncol), min(max(j, 1), nrow));
      lusgs_cell<NV, VISC, FWD>(ring[d], side, up, low, act, has_side, has_low, g, alfa1, P, q, update);
      load_pre<NV, VISC, FWD>(ring[d], P, B, s + D, r, ncol, nrow);
      if (multi) {
        if (lane == 63) put_carry<NV>(xfer[s & 1][wave], up);
        __syncthreads();
      }
    }
  }
}

template <int NV, bool VISC, int D, int MAXT>
__global__ __launch_bounds__(MAXT) void k_lusgs(LusgsArgs a) {
  __shared__ double xfer[2][16][2 * NV + 3];
  const BlockDesc B = a.blocks[a.block_list[blockIdx.x]];
  const CaseParams cp = a.cps[B.case_id];
  const PlaneMap pm = a.pm;
  double* base = a.arena + B.base;
  auto PL = [&](int slot) { return base + (int64_t)slot * B.plane; };
  const int ncol = B.nx - 1, nrow = B.ny - 1;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nwave = (nrow + 63) >> 6;
  const bool multi = nwave > 1;
  const bool row_ok = tid < nrow;
  Planes<NV> P;
#pragma unroll
  for (int m = 0; m < NV; m++) {
    P.U[m] = PL(pm.U + m);
    P.dU[m] = PL(pm.dU + m);
    P.R[m] = PL(pm.Res + m);
  }
  P.vol = PL(pm.vol); P.dt = PL(pm.dt); P.Lci = PL(pm.Lci); P.Lcj = PL(pm.Lcj);
  P.Lvi = PL(pm.Lvi); P.Lvj = PL(pm.Lvj); P.si = PL(pm.si); P.sj = PL(pm.sj);
  P.ni1 = PL(pm.ni1); P.ni2 = PL(pm.ni2); P.nj1 = PL(pm.nj1); P.nj2 = PL(pm.nj2);
  sweep<NV, VISC, true, D>(P, B, ncol, nrow, tid, row_ok, lane, wave, multi, cp.gamma, a.alfa1, a.update, xfer);
  __threadfence();
  __syncthreads();
  sweep<NV, VISC, false, D>(P, B, ncol, nrow, tid, row_ok, lane, wave, multi, cp.gamma, a.alfa1, a.update, xfer);
}

}  // namespace ec2d
