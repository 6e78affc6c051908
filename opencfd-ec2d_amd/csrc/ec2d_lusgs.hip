// LU-SGS implicit update: du_LU_SGS_2D (sub_LU_SGS.f90:4-108), optionally
// followed by U = U + dU (sub_NS_singlegrid.f90:203-211), for all blocks of
// one mesh level.  Three launches:
//
//   k_lusgs_prep     every cell in parallel: writes, for each of the two
//                    sweeps, a record of all the cell's update needs that does
//                    not depend on dU - Res, U, the diagonal, the neighbours'
//                    face coefficients and F(U_nb).n, which is the half of
//                    comput_DFn that depends on the old state only;
//   k_lusgs_sweep    one workgroup per block: the two dependent sweeps;
//   k_lusgs_scatter  every cell in parallel: dU into the block planes and,
//                    when asked, U = U + dU.
//
// Sweep order.  The forward sweep needs dU(i-1,j) and dU(i,j-1), the backward
// sweep dU(i+1,j) and dU(i,j+1).  Any order that respects these dependencies
// gives the reference's result bit for bit.  One lane owns one row of cells
// and step s handles hyperplane s (sub_LU_SGS.f90:25-29, 70-75): lane r takes
// the cell of its row at column offset c = s - r.  The i-neighbour's result
// is the lane's own from the step before (registers); the j-neighbour's is
// that of lane r-1 from the step before, moved one lane up with a DPP wave
// shift, or through LDS between 64-row waves.
//
// Record layout ("skewed").  A block's records are stored hyperplane by
// hyperplane: field f of the cell that lane r handles at step s sits at
// f*n + s*R + r, with R the row count rounded up to 64 and n = S*R for
// S = ncol + nrow - 1 steps.  Every load of a step is then one contiguous run
// per wave instead of 64 scattered lines, and the sweep is bound by its
// dependent arithmetic, not by address divergence.  The forward record of
// (s, r) belongs to cell (c+1, r+1), the backward record to cell
// (ncol-c, nrow-r): both sweeps use the same neighbour tests (c != 0 for the
// i-neighbour, r != 0 for the j-neighbour) and the same index arithmetic.
//
// Per block the records occupy (2*NF + NV) * n doubles: the forward set, the
// backward set (whose X fields receive the forward dU from the forward
// sweep), and NV output fields for the backward dU.
#include "ec2d_device.h"
#include "ec2d_physics.h"

namespace ec2d {

struct LusgsArgs {
  const BlockDesc* blocks;
  const int* block_list;  // blocks of this launch
  const CaseParams* cps;
  double* arena;
  double* sk;             // skewed records of the level
  PlaneMap pm;
  double alfa1;
  int update;             // scatter: 1 also adds dU to U
};

// field offsets of one record (in units of n)
template <int NV, bool VISC>
struct RecF {
  static constexpr int X = 0, U = NV, FI = 2 * NV, FJ = 3 * NV;
  static constexpr int N1I = 4 * NV, N2I = N1I + 1, N1J = N1I + 2, N2J = N1I + 3;
  static constexpr int AI = N1I + 4, AJ = AI + 1, LCI = AI + 2, LCJ = AI + 3;
  static constexpr int LVI = AI + 4, LVJ = AI + 5;  // VISC only
  static constexpr int AL0 = VISC ? AI + 6 : AI + 4;
  static constexpr int AL4 = AL0 + 1;               // NV == 5 only
  static constexpr int NF = AL0 + (NV == 5 ? 2 : 1);
};

// host-side record size, equal to RecF<nv, visc>::NF
inline int lusgs_nf(int nv, bool visc) { return 4 * nv + 8 + (visc ? 2 : 0) + (nv == 5 ? 2 : 1); }

template <int NV>
struct Rec {
  double X[NV], U[NV], Fi[NV], Fj[NV];
  double n1i, n2i, n1j, n2j, ai, aj, Lci, Lcj, Lvi, Lvj, al0, al4;
};

// result of one cell as its neighbours need it: dU and U + dU
template <int NV>
struct Out {
  double du[NV], W[NV];
};

// lane r-1 -> lane r (lane 0 receives 0)
__device__ __forceinline__ double shr1(double v) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_update_dpp(0, lo, 0x138, 0xF, 0xF, false);
  hi = __builtin_amdgcn_update_dpp(0, hi, 0x138, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

template <int NV>
__device__ __forceinline__ void shift_out(Out<NV>& dst, const Out<NV>& src) {
#pragma unroll
  for (int m = 0; m < NV; m++) {
    dst.du[m] = shr1(src.du[m]);
    dst.W[m] = shr1(src.W[m]);
  }
}

template <int NV, bool VISC>
__device__ __forceinline__ void load_rec(Rec<NV>& p, const double* __restrict__ rb, int64_t n, int64_t pos) {
  using F = RecF<NV, VISC>;
  const double* b = rb + pos;
#pragma unroll
  for (int m = 0; m < NV; m++) {
    p.X[m] = b[(F::X + m) * n];
    p.U[m] = b[(F::U + m) * n];
    p.Fi[m] = b[(F::FI + m) * n];
    p.Fj[m] = b[(F::FJ + m) * n];
  }
  p.n1i = b[F::N1I * n]; p.n2i = b[F::N2I * n]; p.n1j = b[F::N1J * n]; p.n2j = b[F::N2J * n];
  p.ai = b[F::AI * n]; p.aj = b[F::AJ * n]; p.Lci = b[F::LCI * n]; p.Lcj = b[F::LCJ * n];
  if constexpr (VISC) {
    p.Lvi = b[F::LVI * n];
    p.Lvj = b[F::LVJ * n];
  } else {
    p.Lvi = p.Lvj = 0.0;
  }
  p.al0 = b[F::AL0 * n];
  if constexpr (NV == 5) p.al4 = b[F::AL4 * n];
  else p.al4 = p.al0;
}

// ----------------------------------------------------------------- prep
// One record.  FWD: cell (c+1, r+1), neighbours (i-1,j) and (i,j-1); the
// face areas are the cell's own (sub_LU_SGS.f90:43-56).  Backward: cell
// (ncol-c, nrow-r), neighbours (i+1,j) and (i,j+1), areas of the neighbour,
// normals of the cell itself (sub_LU_SGS.f90:86-99).
template <int NV, bool VISC, bool FWD>
__device__ __forceinline__ void write_rec(double* __restrict__ rb, int64_t n, const BlockDesc& B, const double* base,
                                          const PlaneMap& pm, int c, int r, double g, double alfa1) {
  using F = RecF<NV, VISC>;
  const int ncol = B.nx - 1, nrow = B.ny - 1;
  auto PL = [&](int slot) { return base + (int64_t)slot * B.plane; };
  if (r >= nrow || c < 0 || c >= ncol) {
    // padding position: never active in a sweep
#pragma unroll 4
    for (int f = FWD ? 0 : NV; f < F::NF; f++) rb[f * n] = 0.0;
    if (!FWD) {
#pragma unroll
      for (int m = 0; m < NV; m++) rb[(F::X + m) * n] = 0.0;
    }
    return;
  }
  const int i = FWD ? c + 1 : ncol - c;
  const int j = FWD ? r + 1 : nrow - r;
  const int d = FWD ? -1 : 1;
  const bool hs = c != 0, hl = r != 0;
  const int64_t q = pidx(B, i, j);
  const int64_t qs = pidx(B, hs ? i + d : i, j);
  const int64_t ql = pidx(B, i, hl ? j + d : j);
  double U[NV];
#pragma unroll
  for (int m = 0; m < NV; m++) {
    U[m] = PL(pm.U + m)[q];
    rb[(F::U + m) * n] = U[m];
    if (FWD) rb[(F::X + m) * n] = PL(pm.Res + m)[q];
  }
  const double n1i = PL(pm.ni1)[q], n2i = PL(pm.ni2)[q], n1j = PL(pm.nj1)[q], n2j = PL(pm.nj2)[q];
  rb[F::N1I * n] = n1i; rb[F::N2I * n] = n2i; rb[F::N1J * n] = n1j; rb[F::N2J * n] = n2j;
  double Us[NV], Ul[NV], Fi[NV], Fj[NV];
#pragma unroll
  for (int m = 0; m < NV; m++) {
    Us[m] = PL(pm.U + m)[qs];
    Ul[m] = PL(pm.U + m)[ql];
  }
  flux_n_old<NV>(Fi, Us, n1i, n2i, g);
  flux_n_old<NV>(Fj, Ul, n1j, n2j, g);
#pragma unroll
  for (int m = 0; m < NV; m++) {
    rb[(F::FI + m) * n] = hs ? Fi[m] : 0.0;
    rb[(F::FJ + m) * n] = hl ? Fj[m] : 0.0;
  }
  const double* si = PL(pm.si);
  const double* sj = PL(pm.sj);
  rb[F::AI * n] = hs ? (FWD ? si[q] : si[qs]) : 0.0;
  rb[F::AJ * n] = hl ? (FWD ? sj[q] : sj[ql]) : 0.0;
  const double* Lci = PL(pm.Lci);
  const double* Lcj = PL(pm.Lcj);
  rb[F::LCI * n] = hs ? Lci[qs] : 0.0;
  rb[F::LCJ * n] = hl ? Lcj[ql] : 0.0;
  const double* Lvi = PL(pm.Lvi);
  const double* Lvj = PL(pm.Lvj);
  if constexpr (VISC) {
    rb[F::LVI * n] = hs ? Lvi[qs] : 0.0;
    rb[F::LVJ * n] = hl ? Lvj[ql] : 0.0;
  }
  // diagonal (sub_LU_SGS.f90:31-39, 76-83)
  const double vol = PL(pm.vol)[q], dt = PL(pm.dt)[q];
  double alfa = vol / dt + 1.0 * (Lci[q] + Lcj[q]) + vol * alfa1;
  if constexpr (VISC) alfa = alfa + 2.0 * (Lvi[q] + Lvj[q]);
  rb[F::AL0 * n] = alfa;
  if constexpr (NV == 5) rb[F::AL4 * n] = alfa + Lvi[q] + Lvj[q];
}

template <int NV, bool VISC>
__global__ __launch_bounds__(256) void k_lusgs_prep(LusgsArgs a) {
  const BlockDesc B = a.blocks[a.block_list[blockIdx.y]];
  const int64_t n = B.sk_n;
  const int64_t pos = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (pos >= n) return;
  const int s = (int)(pos / B.sk_R), r = (int)(pos - (int64_t)s * B.sk_R);
  const double g = a.cps[B.case_id].gamma;
  const double* base = a.arena + B.base;
  double* fw = a.sk + B.sk_base + pos;
  write_rec<NV, VISC, true>(fw, n, B, base, a.pm, s - r, r, g, a.alfa1);
  write_rec<NV, VISC, false>(fw + RecF<NV, VISC>::NF * n, n, B, base, a.pm, s - r, r, g, a.alfa1);
}

// ----------------------------------------------------------------- sweeps
// one cell; FWD: dU = (Res + dui + duj) / alfa, backward:
// dU = dU + (dui + duj) / alfa (sub_LU_SGS.f90:41-62, 85-105)
template <int NV, bool VISC, bool FWD>
__device__ __forceinline__ void lusgs_cell(const Rec<NV>& p, const Out<NV>& side, const Out<NV>& low, bool has_side,
                                           bool has_low, double g, Out<NV>& o) {
  double DFi[NV], DFj[NV];
  dfn_from_old<NV>(DFi, side.W, p.Fi, p.n1i, p.n2i, g);
  dfn_from_old<NV>(DFj, low.W, p.Fj, p.n1j, p.n2j, g);
  double dui[NV], duj[NV];
#pragma unroll
  for (int m = 0; m < NV; m++) {
    double di, dj;
    if (FWD) {
      di = 0.5 * (DFi[m] * p.ai + 1.0 * p.Lci * side.du[m]);
      dj = 0.5 * (DFj[m] * p.aj + 1.0 * p.Lcj * low.du[m]);
    } else {
      di = -0.5 * (DFi[m] * p.ai - 1.0 * p.Lci * side.du[m]);
      dj = -0.5 * (DFj[m] * p.aj - 1.0 * p.Lcj * low.du[m]);
    }
    if constexpr (VISC) {
      di = di + p.Lvi * side.du[m];
      dj = dj + p.Lvj * low.du[m];
    }
    dui[m] = has_side ? di : 0.0;
    duj[m] = has_low ? dj : 0.0;
  }
#pragma unroll
  for (int m = 0; m < NV; m++) {
    const double al = (NV == 5 && m == 4) ? p.al4 : p.al0;
    double du;
    if (FWD) du = (p.X[m] + dui[m] + duj[m]) / al;
    else du = p.X[m] + (dui[m] + duj[m]) / al;
    o.du[m] = du;
    o.W[m] = p.U[m] + du;
  }
}

// Barrier between the waves of one block inside a sweep.  Only LDS data
// (the carries in xfer) crosses waves here, so wait for this wave's LDS
// operations (lgkmcnt(0)) and not for its global loads: __syncthreads() would
// also drain vmcnt and with it the prefetch ring.  s_waitcnt encoding
// (gfx9): vmcnt = 63, expcnt = 7, lgkmcnt = 0.
__device__ __forceinline__ void lds_barrier() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0xC07F);
  __builtin_amdgcn_s_barrier();
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// One sweep over the records rec (forward or backward set).  Each lane
// loads the record it will need D steps later into a ring of registers while
// it computes the current one; loads past the last step are clamped to it.
// The cell's dU goes to dst: for the forward sweep the X field of the
// backward record of the same cell, for the backward sweep the output fields.
template <int NV, bool VISC, bool FWD, int D>
__device__ __forceinline__ void sweep(const double* __restrict__ rec, double* __restrict__ dst, int64_t n, int S,
                                      int R, int ncol, int nrow, int r, int lane, int wave, bool multi, double g,
                                      double (*xfer)[16][2 * NV]) {
  const int npad = (S + D - 1) / D * D;
  const int rl = min(r, R - 1);  // loads of lanes beyond the block stay in its records
  Out<NV> side = {}, low = {}, o;
  Rec<NV> ring[D];
#pragma unroll
  for (int d = 0; d < D; d++) load_rec<NV, VISC>(ring[d], rec, n, (int64_t)min(d, S - 1) * R + rl);
  for (int s0 = 0; s0 < npad; s0 += D) {
#pragma unroll
    for (int d = 0; d < D; d++) {
      const int s = s0 + d;
      shift_out<NV>(low, side);
      if (multi && lane == 0 && wave > 0) {
        const double* x = xfer[(s + 1) & 1][wave - 1];
#pragma unroll
        for (int m = 0; m < NV; m++) { low.du[m] = x[m]; low.W[m] = x[NV + m]; }
      }
      const int c = s - r;
      const bool act = r < nrow && c >= 0 && c < ncol;
      lusgs_cell<NV, VISC, FWD>(ring[d], side, low, c != 0, r != 0, g, o);
      if (act) {
        const int64_t w = FWD ? (int64_t)(S - 1 - s) * R + (nrow - 1 - r) : (int64_t)s * R + r;
#pragma unroll
        for (int m = 0; m < NV; m++) dst[m * n + w] = o.du[m];
        side = o;
      }
      load_rec<NV, VISC>(ring[d], rec, n, (int64_t)min(s + D, S - 1) * R + rl);
      if (multi) {
        if (lane == 63) {
          double* x = xfer[s & 1][wave];
#pragma unroll
          for (int m = 0; m < NV; m++) { x[m] = side.du[m]; x[NV + m] = side.W[m]; }
        }
        lds_barrier();
      }
    }
  }
}

template <int NV, bool VISC, int D, int MAXT>
__global__ __launch_bounds__(MAXT) void k_lusgs_sweep(LusgsArgs a) {
  using F = RecF<NV, VISC>;
  __shared__ double xfer[2][16][2 * NV];
  const BlockDesc B = a.blocks[a.block_list[blockIdx.x]];
  const double g = a.cps[B.case_id].gamma;
  const int ncol = B.nx - 1, nrow = B.ny - 1;
  const int S = B.sk_S, R = B.sk_R;
  const int64_t n = B.sk_n;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const bool multi = R > 64;
  double* fw = a.sk + B.sk_base;
  double* bw = fw + F::NF * n;
  double* out = bw + F::NF * n;
  sweep<NV, VISC, true, D>(fw, bw + F::X * n, n, S, R, ncol, nrow, tid, lane, wave, multi, g, xfer);
  __threadfence();
  __syncthreads();
  sweep<NV, VISC, false, D>(bw, out, n, S, R, ncol, nrow, tid, lane, wave, multi, g, xfer);
}

// ----------------------------------------------------------------- scatter
template <int NV, bool VISC>
__global__ __launch_bounds__(256) void k_lusgs_scatter(LusgsArgs a) {
  using F = RecF<NV, VISC>;
  const BlockDesc B = a.blocks[a.block_list[blockIdx.y]];
  const int ncol = B.nx - 1, nrow = B.ny - 1;
  const int cidx = blockIdx.x * 256 + threadIdx.x;
  if (cidx >= ncol * nrow) return;
  const int j = cidx / ncol + 1, i = cidx - (j - 1) * ncol + 1;
  const int64_t n = B.sk_n;
  const double* out = a.sk + B.sk_base + 2 * F::NF * n + (int64_t)(ncol - i + nrow - j) * B.sk_R + (nrow - j);
  double* base = a.arena + B.base;
  const int64_t q = pidx(B, i, j);
#pragma unroll
  for (int m = 0; m < NV; m++) {
    const double du = out[m * n];
    base[(int64_t)(a.pm.dU + m) * B.plane + q] = du;
    if (a.update) {
      double* U = base + (int64_t)(a.pm.U + m) * B.plane;
      U[q] = U[q] + du;
    }
  }
}

}  // namespace ec2d
