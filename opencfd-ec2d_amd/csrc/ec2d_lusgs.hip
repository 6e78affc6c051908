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
// gives the reference's result bit for bit.  Two lanes own one row of cells
// and step s handles hyperplane s (sub_LU_SGS.f90:25-29, 70-75): row t takes
// its cell at column offset c = s - t.  Of the pair, lane 2t evaluates the
// j-neighbour's flux difference and lane 2t+1 the i-neighbour's; they swap
// the two halves with a DPP quad permute and both finish the cell with the
// same bits.  Every lane then finds its neighbour one lane down (DPP wave
// shift): lane 2t the result of row t-1 (the j-neighbour), lane 2t+1 the
// result of its own row at the step before (the i-neighbour); the first lane
// of a wave gets row t-1 through LDS.
//
// The pair never selects by lane.  The backward sum dui + duj is one addition
// and commutes.  The forward sum (Res + dui) + duj becomes pre + swap(pre)
// with pre = X_k + (own term): lane 2t+1 loads Res as X_k, lane 2t loads -0.0
// (the unwritten trash pad of the forward set's X fields holds it), and
// -0.0 + x == x for every x, so both lanes add (Res + dui) and duj.
//
// Division by the diagonal.  The prep stores y = 1/alfa (correctly rounded);
// the sweep forms q = a*y and refines it twice with FMA residuals (Markstein):
// the first refinement makes q faithful, the second, with y = RN(1/alfa),
// returns RN(a/alfa) - the quotient the reference's division gives, in five
// dependent operations instead of the eleven of a full division.
//
// Record layout ("skewed").  A block's records are stored hyperplane by
// hyperplane: the record of the cell that row t handles at step s sits at
// position s*R + t, with R the row count rounded up to whole waves of the
// sweep (the fused launch, one sweep wave: the row count itself),
// S = ncol + nrow - 1 steps and n = S*R + 64 positions per field.  Fields go
// in pairs of doubles, so that every load and store moves 16 bytes per lane:
// the vector memory address path costs the same per lane for 8 and for 16
// bytes, and with 8-byte loads it, not the arithmetic, bounds a block of four
// waves.  Every load of a step is one contiguous run per wave (buffer loads
// with one lane offset and a scalar offset per pair).  The forward record of
// (s, t) belongs to cell (c+1, t+1), the backward record to cell
// (ncol-c, nrow-t): both sweeps use the same neighbour tests (c != 0 for the
// i-neighbour, t != 0 for the j-neighbour) and the same index arithmetic.
//
// Per block the records occupy (2*NREC + 2*XP) * n doubles: the forward set,
// the backward set (whose X pairs receive the forward dU from the forward
// sweep), and XP output pairs for the backward dU.  The last 64 positions of
// every field are a trash pad that lanes without a cell to store write to,
// so the sweep's stores are unconditional (a store under a branch makes the
// compiler drain every outstanding load, the prefetch ring included, before
// the next use of a loaded register).
#include "ec2d_device.h"
#include "ec2d_physics.h"
#include <type_traits>

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
  int split;              // sweep: workgroups per block (1, 2 or 4: a block's rows split between CUs)
  int dir;                // sweep: 0 both sweeps, 1 forward only, 2 backward only (split launches)
  unsigned long long* gring;  // sweep, split: one ring per boundary between workgroups and direction
  int64_t gstride;        // sweep, split: slots of one ring (at least npad * 2 * NV)
  int* err;               // sweep, split: set when a cross-workgroup wait ran out of its budget
  int spin_max;           // the budget of a bounded wait, in polls
  int nblk;               // skew: blocks of this launch
  int sk_tiles;           // skew: 16 x 16 tiles per block (the largest block's count)
};

// Field layout of one record set (one sweep direction) of a block.  A
// record has a part shared by the two lanes of a cell and a part per side
// (k = 0: the j-neighbour, k = 1: the i-neighbour).  Both are sequences of
// doubles stored in pairs; in doubles from the set's start:
//   shared double d of position pos:   (d/2)*2n + 2*pos + d%2
//   side double e of (pos, k):         NSH*n + (e/2)*4n + 2*(2*pos + k) + e%2
// so that a wave's pair loads are contiguous and the two lanes of a cell
// read adjacent pairs of their sides.
template <int NV, bool VISC>
struct RecF {
  // shared: X = Res (forward) or the forward dU (backward), in whole pairs
  // of their own (the forward pair lanes read them from another position);
  // then U, the diagonal and its correctly rounded reciprocal (a separate one
  // for the SA equation)
  static constexpr int XP = (NV + 1) / 2;
  static constexpr int U = 2 * XP, AL0 = U + NV, Y0 = AL0 + 1, AL4 = AL0 + 2, Y4 = AL0 + 3;
  static constexpr int NSH = 2 * XP + (NV + 2 + (NV == 5 ? 2 : 0) + 1) / 2 * 2;
  // per side: F(U_nb).n, the normal, the face area, the neighbour's inviscid
  // and viscous spectral radius (padded to whole pairs)
  static constexpr int F = 0, N1 = NV, N2 = NV + 1, A = NV + 2, LC = NV + 3, LV = NV + 4;
  static constexpr int NSD = (NV + 4 + (VISC ? 1 : 0) + 1) / 2 * 2;
  static constexpr int NREC = NSH + 2 * NSD;  // doubles per position
};

// The inputs of one record, as logical planes (the block planes they come
// from, or their skewed copies of k_lusgs_skew); U and Res take NV planes.
struct RecIn {
  static constexpr int U = 0;
  template <int NV> static constexpr int RES = NV;
  template <int NV> static constexpr int VOL = 2 * NV;
  template <int NV> static constexpr int DT = 2 * NV + 1;
  template <int NV> static constexpr int LCI = 2 * NV + 2;
  template <int NV> static constexpr int LCJ = 2 * NV + 3;
  template <int NV> static constexpr int LVI = 2 * NV + 4;
  template <int NV> static constexpr int LVJ = 2 * NV + 5;
  template <int NV> static constexpr int NI1 = 2 * NV + 6;
  template <int NV> static constexpr int NI2 = 2 * NV + 7;
  template <int NV> static constexpr int NJ1 = 2 * NV + 8;
  template <int NV> static constexpr int NJ2 = 2 * NV + 9;
  template <int NV> static constexpr int SI = 2 * NV + 10;
  template <int NV> static constexpr int SJ = 2 * NV + 11;
  template <int NV> static constexpr int COUNT = 2 * NV + 12;  // at most NREC of RecF: the skewed copy fits a record set
};

// host side: doubles per position of one record set (RecF<nv, visc>::NREC)
// and of the output pairs
inline int lusgs_rec_doubles(int nv, bool visc) {
  return 2 * ((nv + 1) / 2) + (nv + 2 + (nv == 5 ? 2 : 0) + 1) / 2 * 2 + 2 * ((nv + 4 + (visc ? 1 : 0) + 1) / 2 * 2);
}
inline int lusgs_out_doubles(int nv) { return 2 * ((nv + 1) / 2); }

// operands of one lane for one cell: the shared part and NS sides (NS = 1:
// the lane's own side k, NS = 2: side 0 (j) and side 1 (i))
template <int NV, int NS>
struct Rec {
  double X[NV], U[NV];
  double F[NS][NV], n1[NS], n2[NS], a[NS], Lc[NS], Lv[NS];
  double al0, al4, y0, y4;
};

// one cell as its neighbours need it: dU and U + dU
template <int NV>
struct Out {
  double du[NV], W[NV];
};

// lane l-1 -> lane l (lane 0 receives 0: bound_ctrl)
__device__ __forceinline__ double shr1(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x138, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x138, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}

// lane l <-> lane l^1 (quad_perm [1,0,3,2])
__device__ __forceinline__ double swap_pair(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0xB1, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0xB1, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}

typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
struct D2 {
  double x, y;
};

// one pair of doubles (16 bytes) at voff + soff
__device__ __forceinline__ D2 bload2(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(D2, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}

__device__ __forceinline__ void bstore2(double x, double y, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  const D2 v = {x, y};
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, v), r, voff, soff, 0);
}

// one double (the odd equation's field of the fused launch's X, 8 bytes a position)
__device__ __forceinline__ double bload1(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

__device__ __forceinline__ void bstore1(double x, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, x), r, voff, soff, 0);
}

// v16: byte offset 16*pos of the lane's position, vx16 that of its X pairs
// (its own, or a trash-pad slot); the field offsets are scalar and fixed.
// base: byte offset of the record set in the block's record region, n16 =
// 16n (bytes of one pair field); NS = 1 loads side k, NS = 2 both sides.
template <int NV, bool VISC, int NS>
__device__ __forceinline__ void load_rec(Rec<NV, NS>& p, __amdgpu_buffer_rsrc_t r, uint32_t base, uint32_t n16,
                                         uint32_t v16, uint32_t vx16, int k) {
  using F = RecF<NV, VISC>;
  double sh[F::NSH];
#pragma unroll
  for (int j = 0; j < F::NSH / 2; j++) {
    const D2 v = bload2(r, j < F::XP ? vx16 : v16, base + (uint32_t)j * n16);
    sh[2 * j] = v.x;
    sh[2 * j + 1] = v.y;
  }
#pragma unroll
  for (int m = 0; m < NV; m++) {
    p.X[m] = sh[m];
    p.U[m] = sh[F::U + m];
  }
  p.al0 = sh[F::AL0];
  p.y0 = sh[F::Y0];
  if constexpr (NV == 5) {
    p.al4 = sh[F::AL4];
    p.y4 = sh[F::Y4];
  } else {
    p.al4 = p.al0;
    p.y4 = p.y0;
  }
  const uint32_t sb = base + (uint32_t)(F::NSH / 2) * n16;  // side part, at 2*pos + k
#pragma unroll
  for (int h = 0; h < NS; h++) {
    const uint32_t vd = 2u * v16 + (uint32_t)(NS == 1 ? k : h) * 16u;
    double sd[F::NSD];
#pragma unroll
    for (int j = 0; j < F::NSD / 2; j++) {
      const D2 v = bload2(r, vd, sb + (uint32_t)j * 2u * n16);
      sd[2 * j] = v.x;
      sd[2 * j + 1] = v.y;
    }
#pragma unroll
    for (int m = 0; m < NV; m++) p.F[h][m] = sd[F::F + m];
    p.n1[h] = sd[F::N1];
    p.n2[h] = sd[F::N2];
    p.a[h] = sd[F::A];
    p.Lc[h] = sd[F::LC];
    if constexpr (VISC) p.Lv[h] = sd[F::LV];
    else p.Lv[h] = 0.0;
  }
}

// ----------------------------------------------------------------- prep
// The arithmetic of one record (sub_LU_SGS.f90:31-58 forward, :76-100
// backward) from a loader LD of its inputs; SH(d) and SD(k, e) give the
// destination of shared double d and of side double e of side k (the global
// skewed records of k_lusgs_prep, or an LDS slot of k_lusgs_fused).
template <int NV, bool VISC, bool FWD, class SHF, class SDF, class LDF>
__device__ __forceinline__ void rec_body(SHF SH, SDF SD, LDF LD, bool has_j, bool has_i, double g, double alfa1) {
  using F = RecF<NV, VISC>;
  using I = RecIn;
  // LD(plane, who): who 0 the cell, 1 its j-neighbour, 2 its i-neighbour
#pragma unroll
  for (int m = 0; m < NV; m++) {
    SH(F::U + m) = LD(I::U + m, 0);
    if (FWD) SH(m) = LD(I::RES<NV> + m, 0);
  }
  // diagonal (sub_LU_SGS.f90:31-39, 76-83)
  const double vol = LD(I::VOL<NV>, 0), dt = LD(I::DT<NV>, 0);
  double alfa = vol / dt + 1.0 * (LD(I::LCI<NV>, 0) + LD(I::LCJ<NV>, 0)) + vol * alfa1;
  if constexpr (VISC) alfa = alfa + 2.0 * (LD(I::LVI<NV>, 0) + LD(I::LVJ<NV>, 0));
  SH(F::AL0) = alfa;
  SH(F::Y0) = 1.0 / alfa;
  if constexpr (NV == 5) {
    const double alfa4 = alfa + LD(I::LVI<NV>, 0) + LD(I::LVJ<NV>, 0);
    SH(F::AL4) = alfa4;
    SH(F::Y4) = 1.0 / alfa4;
  }
  // sides: k = 0 the j-neighbour, k = 1 the i-neighbour
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const bool has = k == 0 ? has_j : has_i;
    if (!has) continue;
    const int who = 1 + k;
    const double n1 = LD(k == 0 ? I::NJ1<NV> : I::NI1<NV>, 0), n2 = LD(k == 0 ? I::NJ2<NV> : I::NI2<NV>, 0);
    double Un[NV], Fo[NV];
#pragma unroll
    for (int m = 0; m < NV; m++) Un[m] = LD(I::U + m, who);
    flux_n_old<NV>(Fo, Un, n1, n2, g);
#pragma unroll
    for (int m = 0; m < NV; m++) SD(k, F::F + m) = Fo[m];
    SD(k, F::N1) = n1;
    SD(k, F::N2) = n2;
    SD(k, F::A) = LD(k == 0 ? I::SJ<NV> : I::SI<NV>, FWD ? 0 : who);
    SD(k, F::LC) = LD(k == 0 ? I::LCJ<NV> : I::LCI<NV>, who);
    if constexpr (VISC) SD(k, F::LV) = LD(k == 0 ? I::LVJ<NV> : I::LVI<NV>, who);
  }
}

// block plane slot of a logical input plane
template <int NV>
__device__ __forceinline__ int rec_in_slot(const PlaneMap& pm, int plane) {
  using I = RecIn;
  if (plane < NV) return pm.U + plane;
  if (plane < 2 * NV) return pm.Res + plane - NV;
  switch (plane - 2 * NV) {
    case 0: return pm.vol;
    case 1: return pm.dt;
    case 2: return pm.Lci;
    case 3: return pm.Lcj;
    case 4: return pm.Lvi;
    case 5: return pm.Lvj;
    case 6: return pm.ni1;
    case 7: return pm.ni2;
    case 8: return pm.nj1;
    case 9: return pm.nj2;
    case 10: return pm.si;
    default: return pm.sj;
  }
  (void)sizeof(I);
}

// One record set at position (c, r) from the block planes.  Forward: cell
// (c+1, r+1), neighbours (i-1,j) and (i,j-1); the face areas are the cell's
// own (sub_LU_SGS.f90:43-56).  Backward: cell (ncol-c, nrow-r), neighbours
// (i+1,j) and (i,j+1), areas of the neighbour, normals of the cell itself
// (sub_LU_SGS.f90:86-99).  SH(d) and SD(k, e) give the destination of shared
// double d and of side double e of side k.
template <int NV, bool VISC, bool FWD, class SHF, class SDF>
__device__ __forceinline__ void make_rec(SHF SH, SDF SD, const BlockDesc& B, const double* base, const PlaneMap& pm,
                                         int c, int r, double g, double alfa1) {
  const int ncol = B.nx - 1, nrow = B.ny - 1;
  // A padding position is never active in a sweep, and a side without a
  // neighbour is never used (the sweep selects 0 for it): neither is
  // written.  They keep the zeros the record arena starts with, which no
  // sweep overwrites (sweep stores go to the X pairs of cells and the trash
  // pad), so an inactive lane only ever reads finite values.
  if (r >= nrow || c < 0 || c >= ncol) return;
  const int i = FWD ? c + 1 : ncol - c;
  const int j = FWD ? r + 1 : nrow - r;
  const int d = FWD ? -1 : 1;
  const int64_t qw[3] = {pidx(B, i, j), pidx(B, i, j + d), pidx(B, i + d, j)};
  auto LD = [&](int plane, int who) { return base[(int64_t)rec_in_slot<NV>(pm, plane) * B.plane + qw[who]]; };
  rec_body<NV, VISC, FWD>(SH, SD, LD, r != 0, c != 0, g, alfa1);
}

// The same record from the skewed copy of the inputs (k_lusgs_skew: logical
// plane k of the block at sk + k n, a cell at its forward position
// (i-1 + j-1) R + j-1).  Backward position (c, r) is the cell whose forward
// position is (ncol-1-c + nrow-1-r) R + nrow-1-r; its neighbours (i+1, j),
// (i, j+1) sit R and R + 1 further, the forward neighbours (i-1, j), (i, j-1)
// R and R + 1 before.  Lanes of consecutive rows read consecutive positions.
// Every input is loaded (one batch of independent loads) before the
// arithmetic starts: a producer of the fused launch then waits for memory
// once per record.
template <int NV, bool VISC, bool FWD, class SHF, class SDF>
__device__ __forceinline__ void make_rec_sk(SHF SH, SDF SD, const BlockDesc& B, const double* sk, int c, int r,
                                            double g, double alfa1) {
  using I = RecIn;
  const int ncol = B.nx - 1, nrow = B.ny - 1;
  if (r >= nrow || c < 0 || c >= ncol) return;
  const int64_t R = B.sk_R, n = B.sk_n;
  const int cf = FWD ? c : ncol - 1 - c, rf = FWD ? r : nrow - 1 - r;
  const int64_t p = (int64_t)(cf + rf) * R + rf;
  const int64_t pw[3] = {p, FWD ? p - R - 1 : p + R + 1, FWD ? p - R : p + R};
  const bool has[2] = {r != 0, c != 0};
  // the cell's planes; per side the neighbour's U, Lc, Lv and face area
  constexpr int NK = I::COUNT<NV>, NB = NV + 3;
  double own[NK], nb[2][NB];
#pragma unroll
  for (int k = 0; k < NK; k++) {
    const bool need = (FWD || k < I::RES<NV> || k >= I::VOL<NV>) &&
                      (VISC || (k != I::LVI<NV> && k != I::LVJ<NV>));
    own[k] = need ? sk[(int64_t)k * n + p] : 0.0;
  }
#pragma unroll
  for (int kk = 0; kk < 2; kk++) {
    const int64_t q = pw[1 + kk];
#pragma unroll
    for (int m = 0; m < NB; m++) nb[kk][m] = 0.0;
    if (has[kk]) {
#pragma unroll
      for (int m = 0; m < NV; m++) nb[kk][m] = sk[(int64_t)(I::U + m) * n + q];
      nb[kk][NV] = sk[(int64_t)(kk == 0 ? I::LCJ<NV> : I::LCI<NV>) * n + q];
      if constexpr (VISC) nb[kk][NV + 1] = sk[(int64_t)(kk == 0 ? I::LVJ<NV> : I::LVI<NV>) * n + q];
      if constexpr (!FWD) nb[kk][NV + 2] = sk[(int64_t)(kk == 0 ? I::SJ<NV> : I::SI<NV>) * n + q];
    }
  }
  auto LD = [&](int plane, int who) -> double {
    if (who == 0) return own[plane];
    const double* v = nb[who - 1];
    if (plane < NV) return v[plane];
    if (plane == I::LCI<NV> || plane == I::LCJ<NV>) return v[NV];
    if (plane == I::LVI<NV> || plane == I::LVJ<NV>) return v[NV + 1];
    return v[NV + 2];  // the face area
  };
  rec_body<NV, VISC, FWD>(SH, SD, LD, has[0], has[1], g, alfa1);
}

template <int NV, bool VISC, bool FWD>
__device__ __forceinline__ void write_rec(double* __restrict__ rb, int64_t n, int64_t pos, const BlockDesc& B,
                                          const double* base, const PlaneMap& pm, int c, int r, double g,
                                          double alfa1) {
  using F = RecF<NV, VISC>;
  // shared double d, side double e of side k (layout of RecF)
  auto SH = [&](int d) -> double& { return rb[(int64_t)(d / 2) * 2 * n + 2 * pos + (d & 1)]; };
  auto SD = [&](int k, int e) -> double& {
    return rb[(int64_t)F::NSH * n + (int64_t)(e / 2) * 4 * n + 2 * (2 * pos + k) + (e & 1)];
  };
  make_rec<NV, VISC, FWD>(SH, SD, B, base, pm, c, r, g, alfa1);
}

// One thread per cell, a wave per 8 x 8 tile of cells (a workgroup per 32
// columns x 8 rows): a load of the block planes touches 8 cache lines per
// wave and a record store about 15 (one per hyperplane through the tile);
// one thread per record position would make the loads walk a diagonal, one
// cache line per lane.
template <int NV, bool VISC>
__global__ __launch_bounds__(256) void k_lusgs_prep(LusgsArgs a) {
  using F = RecF<NV, VISC>;
  const BlockDesc B = a.blocks[a.block_list[blockIdx.y]];
  if (case_frozen(B.case_id)) return;
  const int64_t n = B.sk_n;
  const int ncol = B.nx - 1, nrow = B.ny - 1;
  // tiles of the skewed (step, row) space: a workgroup covers 32 steps x 8
  // rows, a wave 8 x 8, eight lanes one step; its cells (i-1, j-1) =
  // (s - r, r) read 8 consecutive cells of 8 rows, and every record store of
  // the eight lanes of a step fills 8 consecutive positions of both sets
  const int S = B.sk_S;
  const int tiles_s = (S + 31) / 32;
  const int tile = blockIdx.x;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int s = (tile % tiles_s) * 32 + wave * 8 + (lane >> 3);  // step
  const int r = (tile / tiles_s) * 8 + (lane & 7);               // j - 1
  const int c = s - r;                                           // i - 1
  const double g = a.cps[B.case_id].gamma;
  const double* base = a.arena + B.base;
  double* fw = a.sk + B.sk_base;
  const int64_t R = B.sk_R;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    // -0.0 in the first trash-pad slot of the forward X pairs: the X operand
    // of the j-neighbour lane of a pair (no sweep stores there)
    const int64_t pad = (int64_t)B.sk_S * R;
#pragma unroll
    for (int d = 0; d < 2 * F::XP; d++) fw[(int64_t)(d / 2) * 2 * n + 2 * pad + (d & 1)] = -0.0;
  }
  if (s >= S || r >= nrow || c < 0 || c >= ncol) return;  // padding positions stay unwritten
  // the cell's forward record at (c, r), its backward record at (ncol-1-c, nrow-1-r)
  const int cb = ncol - 1 - c, rb = nrow - 1 - r;
  write_rec<NV, VISC, true>(fw, n, (int64_t)(c + r) * R + r, B, base, a.pm, c, r, g, a.alfa1);
  write_rec<NV, VISC, false>(fw + F::NREC * n, n, (int64_t)(cb + rb) * R + rb, B, base, a.pm, cb, rb, g, a.alfa1);
}

// The inputs of the fused launch's records in skewed order: logical plane k
// (RecIn) of cell (i, j) goes to sk + k n + (i-1 + j-1) R + j-1 (the
// forward record region, which the fused launch leaves free).  STATIC: the
// mesh planes (vol, normals, face areas), copied once per layout; else the
// planes a residual and an update change (in a main-loop pass the residual
// before the update writes those itself, ResArgs::sk_out, and this copy is
// left out).  A workgroup moves a tile of 16
// steps x 16 rows through LDS: 16 lanes read 16 consecutive cells of a row,
// and 16 lanes write 16 consecutive positions of a step (128 bytes each).
// A row's 16 cells start at column s0 - r, so they straddle two cache lines
// whose other halves the neighbouring tiles of the row read: the grid is
// one-dimensional and XCD-aware (workgroup w runs on XCD w % 8, and XCD x
// takes the contiguous run [x chunk, (x + 1) chunk) of the (block, tile)
// items), so that neighbouring tiles fetch their shared lines into one L2.
template <int NV, bool VISC>
__host__ __device__ constexpr bool skew_plane(int k, bool stat) {
  using I = RecIn;
  const bool is_static = k == I::VOL<NV> || k >= I::NI1<NV>;
  if (!VISC && (k == I::LVI<NV> || k == I::LVJ<NV>)) return false;
  return is_static == stat;
}

template <int NV, bool VISC, bool STATIC>
__global__ __launch_bounds__(256) void k_lusgs_skew(LusgsArgs a) {
  constexpr int NK = RecIn::COUNT<NV>;
  // the planes this launch copies, in order
  constexpr int NC = [] {
    int n = 0;
    for (int k = 0; k < NK; k++) n += skew_plane<NV, VISC>(k, STATIC) ? 1 : 0;
    return n;
  }();
  __shared__ double tile[NC][16][17];  // [plane][row][step], padded against bank conflicts
  const int chunk = (int)gridDim.x / 8;  // the grid is 8 chunk workgroups
  const int w = ((int)blockIdx.x % 8) * chunk + (int)blockIdx.x / 8;
  const int bi = w / a.sk_tiles, ti = w % a.sk_tiles;
  if (bi >= a.nblk) return;
  const BlockDesc B = a.blocks[a.block_list[bi]];
  if (case_frozen(B.case_id)) return;
  const int ncol = B.nx - 1, nrow = B.ny - 1;
  const int S = B.sk_S;
  const int tiles_s = (S + 15) / 16;
  if (ti >= tiles_s * ((nrow + 15) / 16)) return;  // the launch fits the level's largest block
  const int s0 = (ti % tiles_s) * 16, r0 = (ti / tiles_s) * 16;
  const int tid = threadIdx.x;
  const double* __restrict__ base = a.arena + B.base;
  // load: 16 lanes of a row read 16 consecutive cells (step s, cell c = s - r)
  {
    const int rr = tid >> 4, ss = tid & 15, r = r0 + rr, s = s0 + ss, c = s - r;
    if (s < S && r < nrow && c >= 0 && c < ncol) {
      const int64_t q = pidx(B, c + 1, r + 1);
      int j = 0;
#pragma unroll
      for (int k = 0; k < NK; k++)
        if (skew_plane<NV, VISC>(k, STATIC)) tile[j++][rr][ss] = base[(int64_t)rec_in_slot<NV>(a.pm, k) * B.plane + q];
    }
  }
  __syncthreads();
  // store: 16 lanes of a step write 16 consecutive positions s R + r
  {
    const int ss = tid >> 4, rr = tid & 15, r = r0 + rr, s = s0 + ss, c = s - r;
    if (s < S && r < nrow && c >= 0 && c < ncol) {
      double* __restrict__ sk = a.sk + B.sk_base + (int64_t)s * B.sk_R + r;
      int j = 0;
#pragma unroll
      for (int k = 0; k < NK; k++)
        if (skew_plane<NV, VISC>(k, STATIC)) sk[(int64_t)k * B.sk_n] = tile[j++][rr][ss];
    }
  }
}

// The reverse of k_lusgs_skew<..., false> for the planes a residual writes
// (Res, dt, Lci, Lcj and, viscous, Lvi, Lvj; not U, whose planes the update
// keeps): after a main-loop pass whose residual wrote them to the hyperplane
// copy alone (ResArgs::sk_only), this puts them back into their planes.  The
// same tiles, the two phases swapped: 16 lanes of a step read 16 consecutive
// positions, 16 lanes of a row write 16 consecutive cells.  Every block is
// copied, frozen cases included (their copy holds their last residual).
template <int NV, bool VISC>
__host__ __device__ constexpr bool unskew_plane(int k) {
  return skew_plane<NV, VISC>(k, false) && (k < RecIn::U || k >= RecIn::U + NV);
}

template <int NV, bool VISC>
__global__ __launch_bounds__(256) void k_lusgs_unskew(LusgsArgs a) {
  constexpr int NK = RecIn::COUNT<NV>;
  constexpr int NC = [] {
    int n = 0;
    for (int k = 0; k < NK; k++) n += unskew_plane<NV, VISC>(k) ? 1 : 0;
    return n;
  }();
  __shared__ double tile[NC][16][17];
  const int chunk = (int)gridDim.x / 8;
  const int w = ((int)blockIdx.x % 8) * chunk + (int)blockIdx.x / 8;
  const int bi = w / a.sk_tiles, ti = w % a.sk_tiles;
  if (bi >= a.nblk) return;
  const BlockDesc B = a.blocks[a.block_list[bi]];
  const int ncol = B.nx - 1, nrow = B.ny - 1;
  const int S = B.sk_S;
  const int tiles_s = (S + 15) / 16;
  if (ti >= tiles_s * ((nrow + 15) / 16)) return;
  const int s0 = (ti % tiles_s) * 16, r0 = (ti / tiles_s) * 16;
  const int tid = threadIdx.x;
  {
    const int ss = tid >> 4, rr = tid & 15, r = r0 + rr, s = s0 + ss, c = s - r;
    if (s < S && r < nrow && c >= 0 && c < ncol) {
      const double* __restrict__ sk = a.sk + B.sk_base + (int64_t)s * B.sk_R + r;
      int j = 0;
#pragma unroll
      for (int k = 0; k < NK; k++)
        if (unskew_plane<NV, VISC>(k)) tile[j++][rr][ss] = sk[(int64_t)k * B.sk_n];
    }
  }
  __syncthreads();
  {
    const int rr = tid >> 4, ss = tid & 15, r = r0 + rr, s = s0 + ss, c = s - r;
    if (s < S && r < nrow && c >= 0 && c < ncol) {
      double* __restrict__ base = a.arena + B.base;
      const int64_t q = pidx(B, c + 1, r + 1);
      int j = 0;
#pragma unroll
      for (int k = 0; k < NK; k++)
        if (unskew_plane<NV, VISC>(k)) base[(int64_t)rec_in_slot<NV>(a.pm, k) * B.plane + q] = tile[j++][rr][ss];
    }
  }
}

// ----------------------------------------------------------------- sweeps
// One side's share of a cell, dui (the i-neighbour) or duj (the
// j-neighbour), of sub_LU_SGS.f90:42-58 (forward) or :85-100 (backward);
// h selects the side held in the record.
template <int NV, bool VISC, bool FWD, int NS>
__device__ __forceinline__ void side_term(double* d, const Rec<NV, NS>& p, int h, const Out<NV>& nb, bool has,
                                          double g) {
  double DF[NV];
  dfn_from_old<NV>(DF, nb.W, p.F[h], p.n1[h], p.n2[h], g);
#pragma unroll
  for (int m = 0; m < NV; m++) {
    double v;
    if (FWD) v = 0.5 * (DF[m] * p.a[h] + 1.0 * p.Lc[h] * nb.du[m]);
    else v = -0.5 * (DF[m] * p.a[h] - 1.0 * p.Lc[h] * nb.du[m]);
    if constexpr (VISC) v = v + p.Lv[h] * nb.du[m];
    d[m] = has ? v : 0.0;
  }
}

// The waves of a block as a pipeline.  Wave w needs, at step s, the result
// of the last row of wave w-1 at step s-1 and nothing else from other waves,
// so waves do not meet at a barrier per step: wave w-1 puts its last row's
// result of step s into slot s % RING of an LDS ring and then publishes
// prog = s + 1; wave w waits for prog >= s, reads slot (s-1) % RING and
// publishes cons = s (the input of step s taken).  Before wave w-1 reuses a
// slot it waits for cons >= s - RING + 1.  Neither wait can block forever:
// wave w-1 waits only while wave w is RING steps behind, and then wave w does
// not wait (prog is ahead of it), so wave w advances; wave 0 never waits for
// input.  Every wave runs the same fixed number of steps, so all waits end.
// LDS operations of one wave execute in issue order, so a published counter
// is never seen before the ring data written ahead of it, and a slot is never
// rewritten before the read that cons acknowledges; compiler fences keep
// that order in the instruction stream.  The spins read LDS only and never
// wait for the global-load prefetch.
//
// Both sides read ahead, so that no LDS latency sits on a step's dependent
// path: during step s the consumer reads prog and then, behind it, the slot
// holding its input of step s+1; at step s+1 the data is used if that prog
// showed the producer past step s (the data read executed after the prog
// read, so it saw the slot written), and read again after a wait otherwise.
// The producer likewise reads cons one step ahead; a stale value only makes
// it wait longer.
//
// ring[w] is the input of wave w: its producer is wave w-1 (or, in the upper
// workgroup of a split block, the communication wave), prog[w] that
// producer's progress and cons[w] wave w's consumption.
constexpr int RING = 8;
template <int NV>
struct Chain {
  double ring[17][RING][2 * NV];  // input of wave w: (dU, U + dU) of the row below it
  int prog[17];                   // steps finished by the producer of ring[w]
  int cons[17];                   // steps whose input wave w has read
};

// per-wave read-ahead state: the next input (lane 0's) and the counters
// seen when it was read
template <int NV>
struct Ahead {
  double du[NV], W[NV];
  int prog, cons;
};

__device__ __forceinline__ int lds_get(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_put(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void wait_at_least(const int* p, int v) {
  while (lds_get(p) < v) __builtin_amdgcn_s_sleep(1);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// One step of a sweep: the cell of row t at column offset c = s - t.
// LPC = 2: lanes 2t (k = 0, the j-neighbour's term) and 2t+1 (k = 1, the
// i-neighbour's term) share the cell and swap their halves; LPC = 1: lane t
// computes both.
// ODD1 (the fused launch's layout): with an odd equation count the last X
// field holds single doubles, not a pair padded with zero, and lanes without
// a cell store nothing (no trash-pad slot)
template <int NV, bool VISC, bool FWD, int LPC, bool STORE = true, bool ODD1 = false>
__device__ __forceinline__ void sweep_step(const Rec<NV, 3 - LPC>& p, Out<NV>& own, __amdgpu_buffer_rsrc_t rs,
                                           uint32_t dst, uint32_t n16, int s, int S, int R, int ncol, int nrow, int t,
                                           int k, int lane, int wave, bool cin, bool cout, double g,
                                           Chain<NV>& ch, Ahead<NV>& ah) {
  // the j-neighbour's result: row t-1's, one lane down (the last lane of the
  // previous wave through the LDS ring); the i-neighbour's is the own row's
  Out<NV> up;
#pragma unroll
  for (int m = 0; m < NV; m++) {
    up.du[m] = shr1(own.du[m]);
    up.W[m] = shr1(own.W[m]);
  }
  if (cin) {
    // the row below this wave at step s-1 (an LDS broadcast), lane 0 keeps
    // it; read one step ahead unless the producer was not there yet
    if (ah.prog < s) {
      wait_at_least(&ch.prog[wave], s);
      const double* x = ch.ring[wave][(s - 1) & (RING - 1)];
#pragma unroll
      for (int m = 0; m < NV; m++) {
        ah.du[m] = x[m];
        ah.W[m] = x[NV + m];
      }
    }
#pragma unroll
    for (int m = 0; m < NV; m++) {
      up.du[m] = lane == 0 ? ah.du[m] : up.du[m];
      up.W[m] = lane == 0 ? ah.W[m] : up.W[m];
    }
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    lds_put(&ch.cons[wave], s);
    // the input of step s+1: prog first, the slot behind it
    ah.prog = lds_get(&ch.prog[wave]);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    const double* x = ch.ring[wave][s & (RING - 1)];
#pragma unroll
    for (int m = 0; m < NV; m++) {
      ah.du[m] = x[m];
      ah.W[m] = x[NV + m];
    }
  }
  const int c = s - t;
  const bool act = t < nrow && c >= 0 && c < ncol;
  // sum: (Res + dui) + duj forward, dui + duj backward
  double sum[NV];
  if constexpr (LPC == 2) {
    // the neighbour one lane down is the j-neighbour for lane 2t and the
    // i-neighbour for lane 2t+1
    double mine[NV];
    side_term<NV, VISC, FWD, 1>(mine, p, 0, up, k == 0 ? t != 0 : c != 0, g);
#pragma unroll
    for (int m = 0; m < NV; m++) {
      const double pre = FWD ? p.X[m] + mine[m] : mine[m];  // forward: X = Res (lane 2t+1), -0.0 (lane 2t)
      sum[m] = pre + swap_pair(pre);
    }
  } else {
    double dui[NV], duj[NV];
    side_term<NV, VISC, FWD, 2>(duj, p, 0, up, t != 0, g);
    side_term<NV, VISC, FWD, 2>(dui, p, 1, own, c != 0, g);
#pragma unroll
    for (int m = 0; m < NV; m++) sum[m] = FWD ? p.X[m] + dui[m] + duj[m] : dui[m] + duj[m];
  }
  Out<NV> o;
#pragma unroll
  for (int m = 0; m < NV; m++) {
    const bool sa = NV == 5 && m == 4;
    const double q = div_rcp(sum[m], sa ? p.al4 : p.al0, sa ? p.y4 : p.y0);
    const double du = FWD ? q : p.X[m] + q;
    o.du[m] = du;
    o.W[m] = p.U[m] + du;
  }
  // Every lane keeps its result, active or not: an inactive lane's value
  // never reaches an active cell.  Row t's first cell (c = 0) ignores its
  // i-neighbour; row t+1 at step s+1 sits at the same column offset as row t
  // at step s, so it is inactive whenever row t was.
  own = o;
  if constexpr (STORE) {
    // one lane of an active cell stores; every other lane writes its own
    // slot of the trash pad (positions S*R .. S*R+63)
    const uint32_t w = (act && k == 0) ? (FWD ? (uint32_t)((S - 1 - s) * R + (nrow - 1 - t)) : (uint32_t)(s * R + t))
                                       : (uint32_t)(S * R + lane);
    if (!ODD1 || (act && k == 0)) {
#pragma unroll
      for (int j = 0; j < (NV + 1) / 2; j++) {
        if (ODD1 && 2 * j + 1 == NV) bstore1(o.du[2 * j], rs, w * 8u, dst + (uint32_t)j * n16);
        else bstore2(o.du[2 * j], 2 * j + 1 < NV ? o.du[2 * j + 1] : 0.0, rs, w * 16u, dst + (uint32_t)j * n16);
      }
    }
  }
  if (cout) {
    // slot s % RING is free once the next wave has read step s - RING's result
    if (ah.cons < s - RING + 1) wait_at_least(&ch.cons[wave + 1], s - RING + 1);
    if (lane == 63) {
      double* x = ch.ring[wave + 1][s & (RING - 1)];
#pragma unroll
      for (int m = 0; m < NV; m++) { x[m] = own.du[m]; x[NV + m] = own.W[m]; }
    }
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    lds_put(&ch.prog[wave + 1], s + 1);
    ah.cons = lds_get(&ch.cons[wave + 1]);
  }
}

// One sweep.  With LPC = 2 lanes 2t and 2t+1 own row t: lane 2t computes
// the j-neighbour's term, lane 2t+1 the i-neighbour's, the pair swaps them
// (DPP) and both finish the cell, so both hold the cell's result.  With
// LPC = 1 lane t owns row t alone.  The i-neighbour of a cell is the own
// row's previous result (registers); the j-neighbour is the previous result
// of row t-1, which the next lane down holds: one DPP shift up.
// The records come in two register sets of U steps each, loaded one phase
// ahead: [load B] [steps of A] [load A] [steps of B] ...  Whenever the code
// waits for a set, the only loads still in flight are the other set's, issued
// after it, so the waits the compiler places never drain a prefetch that is
// not yet needed.  Loads past the last step are clamped to it.  The cell's dU goes to dst: for the forward sweep the X field of the
// backward record of the same cell, for the backward sweep the output fields.
template <int NV, bool VISC, bool FWD, int U, int LPC>
__device__ __forceinline__ void sweep(__amdgpu_buffer_rsrc_t rs, uint32_t rec, uint32_t dst, uint32_t n16, int S, int R,
                                      int ncol, int nrow, int t, int k, int lane, int wave, bool cin, bool cout,
                                      double g, Chain<NV>& ch) {
  Ahead<NV> ah = {};  // prog = 0: step 0 needs nothing (its input is never used)
  const int npad = (S + 2 * U - 1) / (2 * U) * (2 * U);
  // lanes past the block's last row load that row's records: the same cache
  // lines as its own lane, so padding rows cost no memory traffic (their
  // results never reach a cell)
  const int tl = min(t, nrow - 1);
  constexpr int NS = 3 - LPC;
  Out<NV> own = {};
  Rec<NV, NS> A[U], B[U];
  // forward pairs: lane 2t reads X (-0.0) from the first trash-pad slot
  const bool xpad = FWD && LPC == 2 && k == 0;
  const uint32_t pad16 = (uint32_t)(S * R) * 16u;
  auto load_set = [&](Rec<NV, NS>* set, int s_first) {
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint32_t v16 = (uint32_t)(min(s_first + u, S - 1) * R + tl) * 16u;
      load_rec<NV, VISC, NS>(set[u], rs, rec, n16, v16, xpad ? pad16 : v16, k);
    }
  };
  auto run_set = [&](const Rec<NV, NS>* set, int s_first) {
#pragma unroll
    for (int u = 0; u < U; u++)
      sweep_step<NV, VISC, FWD, LPC>(set[u], own, rs, dst, n16, s_first + u, S, R, ncol, nrow, t, k, lane, wave,
                                     cin, cout, g, ch, ah);
  };
  load_set(A, 0);
  for (int s0 = 0; s0 < npad; s0 += 2 * U) {
    load_set(B, s0 + U);
    __builtin_amdgcn_sched_barrier(0);
    run_set(A, s0);
    __builtin_amdgcn_sched_barrier(0);
    load_set(A, s0 + 2 * U);
    __builtin_amdgcn_sched_barrier(0);
    run_set(B, s0 + U);
    __builtin_amdgcn_sched_barrier(0);
  }
}

// ---- a block's rows split between workgroups (one CU each) --------------
// The last compute wave of a workgroup feeds an extra communication wave
// through the LDS ring as if it were the next wave; that wave forwards each
// step's row through global memory to a communication wave of the next
// workgroup, which feeds its first compute wave through the LDS ring.
// Compute waves never poll global memory (a wait on a global load would also
// wait for their record prefetch).  Every step of a sweep has its own global
// slot (steps 0 .. npad-2; the next workgroup never needs the last one), so
// the forwarder only stores: no flow control, no round trip.  A slot double
// holds the sentinel until the forwarder writes the value; the receiver
// polls each double until it differs from the sentinel (every double is its
// own 8-byte handshake, so no ordering between doubles is needed) and, when
// its sweep is done, sets all its slots back to the sentinel for the next
// launch.  The sentinel is a NaN payload that arithmetic never produces (it
// yields the default NaN).  Global waits are bounded: past a.spin_max polls
// the wave sets err and stops waiting for the rest of the sweep, so the
// kernel always ends within one budget (its results are then wrong, and the
// host reports err).  A budget of 0 makes every receiver give up at its
// first slot (the host's way to test the error path).
constexpr unsigned long long GSENT = 0x7FF4DEADBEEF0001ull;
constexpr int SPIN_MAX = 1 << 20;  // the default budget

__device__ __forceinline__ unsigned long long gload(unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void gstore(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// sets the global rings to the sentinel once, when the layout is made
__global__ void k_fill_u64(unsigned long long* p, int64_t n, unsigned long long v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// The communication waves move CK = 64 / (2 NV) steps at a time, one double
// per lane, so that one round trip to global memory covers CK steps.
template <int NV>
struct CommLane {
  static constexpr int CK = 64 / (2 * NV);
  int kk, f;
  bool on;  // this lane carries double f of step kk of the batch
  __device__ CommLane() : kk((int)(threadIdx.x & 63) / (2 * NV)), f((int)(threadIdx.x & 63) % (2 * NV)), on(kk < CK) {}
};

// sending side: take ring[nw] (the last compute wave's row) and store it
template <int NV>
__device__ void comm_forward(Chain<NV>& ch, int nw, unsigned long long* gr, int npad) {
  const CommLane<NV> c;
  for (int s0 = 0; s0 < npad - 1; s0 += c.CK) {
    const int nst = min(c.CK, npad - 1 - s0);
    wait_at_least(&ch.prog[nw], s0 + nst);
    const bool on = c.on && c.kk < nst;
    const double v = on ? ch.ring[nw][(s0 + c.kk) & (RING - 1)][c.f] : 0.0;
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    lds_put(&ch.cons[nw], s0 + nst);
    if (on) gstore(gr + (s0 + c.kk) * (2 * NV) + c.f, __builtin_bit_cast(unsigned long long, v));
  }
}

// receiving side: take each batch of rows and feed ring[0]; then reset
template <int NV>
__device__ void comm_receive(Chain<NV>& ch, unsigned long long* gr, int npad, int* err, int spin_max) {
  const CommLane<NV> c;
  static_assert(CommLane<NV>::CK <= RING, "a batch must fit the LDS ring");
  bool gone = false;  // the other workgroup stopped answering: wait no more
  for (int s0 = 0; s0 < npad - 1; s0 += c.CK) {
    const int nst = min(c.CK, npad - 1 - s0);
    const bool on = c.on && c.kk < nst;
    double v = 0.0;
    if (on && !gone) {
      unsigned long long* slot = gr + (s0 + c.kk) * (2 * NV) + c.f;
      unsigned long long x;
      int n = 0;
      while ((x = gload(slot)) == GSENT && ++n < spin_max) __builtin_amdgcn_s_sleep(1);
      if (n >= spin_max) {
        atomicOr(err, 1);
        gone = true;
      }
      v = __builtin_bit_cast(double, x);
    }
    // the slots of steps s0 .. s0+nst-1 are free once wave 0 read step s0+nst-1-RING
    wait_at_least(&ch.cons[0], s0 + nst - RING);
    if (on) ch.ring[0][(s0 + c.kk) & (RING - 1)][c.f] = v;
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    lds_put(&ch.prog[0], s0 + nst);
  }
  // every value of this launch has been read: back to the sentinel
  for (int i = (int)(threadIdx.x & 63); i < (npad - 1) * 2 * NV; i += 64) gstore(gr + i, GSENT);
}

// LPC lanes per cell (1 or 2), U steps per register set, MAXT threads at most.
// a.split workgroups per block (blockIdx.x = block * split + part); with a
// split each part has its compute waves, then a receiving communication wave
// (all parts but the first), then a sending one (all parts but the last), and
// the two sweeps are separate launches (a.dir), since the backward sweep of
// one part needs forward results of the others.
template <int NV, bool VISC, int U, int MAXT, int LPC>
__global__ __launch_bounds__(MAXT) void k_lusgs_sweep(LusgsArgs a) {
  using F = RecF<NV, VISC>;
  __shared__ Chain<NV> ch[2];  // one per sweep
  const int G = a.split, bi = blockIdx.x / G, half = blockIdx.x % G;  // half: the part of the block
  const BlockDesc B = a.blocks[a.block_list[bi]];
  if (case_frozen(B.case_id)) return;  // every part of the block returns: no part waits on another
  const double g = a.cps[B.case_id].gamma;
  const int ncol = B.nx - 1, nrow = B.ny - 1;
  const int S = B.sk_S, R = B.sk_R;
  const uint32_t n8 = (uint32_t)B.sk_n * 8u, n16 = 2u * n8;  // bytes of a field of doubles, of pairs
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // compute waves holding rows of this part (R is a multiple of split * 64 /
  // LPC); the launch may carry more for the level's tallest block, they sit out
  const int nw = R * LPC / (64 * G);
  const int t = half * (R / G) + (LPC == 2 ? tid >> 1 : tid), k = LPC == 2 ? tid & 1 : 0;
  const bool has_in = half > 0, has_out = half < G - 1;
  const bool cin = wave > 0 || has_in, cout = wave < nw - 1 || has_out;
  if (tid < 17) {
    ch[0].prog[tid] = ch[0].cons[tid] = ch[1].prog[tid] = ch[1].cons[tid] = 0;
  }
  __syncthreads();
  // the block's record region: forward set, backward set (its X pairs first),
  // output pairs
  const uint32_t bytes = (uint32_t)((2 * F::NREC + 2 * F::XP) * B.sk_n * 8);
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(a.sk + B.sk_base, 0, bytes, 0x00020000);
  const uint32_t fw = 0, bw = F::NREC * n8, out = 2 * F::NREC * n8;
  const int npad = (S + 2 * U - 1) / (2 * U) * (2 * U);
  // global ring of boundary b (between parts b and b+1) for sweep direction d
  auto gr = [&](int b, int d) { return a.gring + (((int64_t)bi * (G - 1) + b) * 2 + d) * a.gstride; };
  const bool rcv = has_in && wave == nw, snd = has_out && wave == nw + (has_in ? 1 : 0);
  if (a.dir != 2) {
    if (wave < nw) sweep<NV, VISC, true, U, LPC>(rs, fw, bw, n16, S, R, ncol, nrow, t, k, lane, wave, cin, cout, g, ch[0]);
    else if (rcv) comm_receive<NV>(ch[0], gr(half - 1, 0), npad, a.err, a.spin_max);
    else if (snd) comm_forward<NV>(ch[0], nw, gr(half, 0), npad);
  }
  if (a.dir == 0) {
    __threadfence();
    __syncthreads();
  }
  if (a.dir != 1) {
    if (wave < nw) sweep<NV, VISC, false, U, LPC>(rs, bw, out, n16, S, R, ncol, nrow, t, k, lane, wave, cin, cout, g, ch[1]);
    else if (rcv) comm_receive<NV>(ch[1], gr(half - 1, 1), npad, a.err, a.spin_max);
    else if (snd) comm_forward<NV>(ch[1], nw, gr(half, 1), npad);
  }
}

// ------------------------------------------------- preparation in the sweep
// For levels whose blocks have at most 64 rows (one wave per block, one lane
// per row): k_lusgs_fused does the work of k_lusgs_prep inside the sweep's
// workgroup, so the records never go through HBM.  Wave 0 runs both sweeps
// exactly as k_lusgs_sweep<..., LPC = 1> does (the same sweep_step), at the
// highest issue priority of its SIMD; waves 1..NP produce the records, each
// every NP-th step of the sequence (forward steps 0..npad-1, then backward
// steps 0..npad-1), lane t the cell of row t, from the skewed copy of the
// inputs that k_lusgs_skew made (make_rec_sk, the arithmetic of
// k_lusgs_prep), into a ring of LDS slots (6 for NP = 3, 7 for NP = 6).  The
// forward dU goes to the backward records' X pairs in global memory (the
// backward sweep reads it there, after a fence).  The sweep wave puts the
// backward U + dU (with update; dU without) of eight steps into an LDS
// buffer; wave NP + 1,
// the writer, takes each chunk and stores it to the block planes eight
// consecutive cells of a row at a time: this replaces k_lusgs_scatter
// (stores along a hyperplane, one cache line per lane, hold up whichever
// wave issues them).  NP = 3 or 6 (5 or 8 waves; with 6 the ring and the
// buffer of a six-equation record take 157 KB of the CU's 160 KB).
//
// The writer waits for ob_ready > chunk, reads the chunk, then sets
// ob_taken = chunk + 1; the sweep wave waits for ob_taken >= chunk before
// it writes the chunk's first step, and sets ob_ready after its last.
// Protocol: produced[p] counts the steps producer p has written, consumed
// the steps whose slot the sweep wave has finished with (it releases slot g
// after computing step g).  The sweep wave waits, before reading step g,
// for produced[g % NP] > g / NP; producer p waits, before writing step g
// into slot g % NRING, for consumed >= g - NRING + 1.  A producer waits only
// for the sweep wave and the sweep wave only for producers that are never
// more than NRING steps ahead, so every wait ends; waits are also bounded
// (a.spin_max polls, then err is set and the wave stops waiting).
template <int NV, bool VISC, int NP>
struct FusedLds {
  using F = RecF<NV, VISC>;
  static constexpr int NRING = NP == 3 ? 6 : 7;  // LDS record slots
  double sh[NRING][F::NSH][64];     // shared doubles of slot, lane
  double sd[NRING][2][F::NSD][64];  // side doubles
  double ob[8][NV][64];             // the sweep's backward U + dU (update) or dU of eight steps
  int produced[NP];
  int consumed;
  int ob_ready, ob_taken;           // chunks of eight backward steps put into / taken from ob
};

__device__ __forceinline__ bool wait_bounded(const int* p, int v, int* err, int spin_max) {
  int n = 0;
  while (lds_get(p) < v) {
    if (++n >= spin_max) {
      atomicOr(err, 2);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  return true;
}

template <int NV, bool VISC, int NP>
__global__ __launch_bounds__(64 * (2 + NP)) void k_lusgs_fused(LusgsArgs a) {
  using F = RecF<NV, VISC>;
  using LDS = FusedLds<NV, VISC, NP>;
  constexpr int FPROD = NP, FRING = LDS::NRING;
  __shared__ LDS L;
  const BlockDesc B = a.blocks[a.block_list[blockIdx.x]];
  if (case_frozen(B.case_id)) return;
  const double g = a.cps[B.case_id].gamma;
  const int ncol = B.nx - 1, nrow = B.ny - 1;
  const int S = B.sk_S, R = B.sk_R;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid < FPROD) L.produced[tid] = 0;
  if (tid == 0) L.consumed = L.ob_ready = L.ob_taken = 0;
  __syncthreads();
  constexpr int U = 2;  // the sweep's step padding (npad) of k_lusgs_sweep<..., 2, ..., 1>
  const int npad = (S + 2 * U - 1) / (2 * U) * (2 * U);
  const int ntot = 2 * npad;
  const uint32_t n8 = (uint32_t)B.sk_n * 8u, n16 = 2u * n8;
  const uint32_t bytes = (uint32_t)((2 * F::NREC + 2 * F::XP) * B.sk_n * 8);
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(a.sk + B.sk_base, 0, bytes, 0x00020000);
  const uint32_t bw = F::NREC * n8;
  double* planes = a.arena + B.base;
  bool ok = true;
  if (wave == FPROD + 1) {
    // ---- writer: each chunk of eight backward steps from ob to the block
    // planes, every store one aligned 64-byte run (8 cells) of each of 8
    // rows.  Row r's cells of a chunk (steps 8ck .. 8ck+7, columns c = step
    // - r, plane offset ncol + 1 - c in the row, whose start is 128-byte
    // aligned: the pitch is a multiple of 16) straddle two aligned runs.
    // Lane = 8 rows x position p of the run (offset mod 8): it reads the one
    // step whose cell sits at position p, always the same slot sub of the
    // chunk.  Positions 0..q (q = (ncol + 1 + r) mod 8) complete the run they
    // fall in and are stored at once; positions q+1..7 start the next run and
    // are held one chunk, while the value held from the previous chunk
    // completes this one.  A last turn stores the runs held after the final
    // chunk.  Every cell is stored exactly once, cells outside the block
    // (columns < 0 or >= ncol, rows >= nrow) are skipped.  With update the
    // planes are U (U + dU, the update of k_lusgs_scatter; nothing reads the
    // dU planes after a main-loop pass, so they are not rewritten), without
    // it dU.
    const int p = lane & 7, rl = lane >> 3;
    const int q = (ncol + 1 + rl) & 7;   // rows rl + 8k share it: the pitch is a multiple of 8
    const int sub = (q - p) & 7;
    const bool held = p > q;             // this lane's cell starts the next run
    const int dst = a.update ? a.pm.U : a.pm.dU;
    const int nck = (npad + 7) / 8;
    double keep[8][NV] = {};
    for (int ck = 0; ck <= nck && ok; ck++) {
      double v[8][NV];
      if (ck < nck) {
        ok = wait_bounded(&L.ob_ready, ck + 1, a.err, a.spin_max);
        if (!ok) break;
#pragma unroll
        for (int rg = 0; rg < 8; rg++)
#pragma unroll
          for (int m = 0; m < NV; m++) v[rg][m] = L.ob[sub][m][rg * 8 + rl];
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(0xc07f);  // the buffer read before it is handed back
        lds_put(&L.ob_taken, ck + 1);
      } else {
#pragma unroll
        for (int rg = 0; rg < 8; rg++)
#pragma unroll
          for (int m = 0; m < NV; m++) v[rg][m] = 0.0;
      }
#pragma unroll
      for (int rg = 0; rg < 8; rg++) {
        const int r = rg * 8 + rl;
        // the run's offset in the row: 8 * floor((ncol + 1 + r) / 8) - 8ck + p
        const int off = ((ncol + 1 + r) & ~7) - 8 * ck + p, c = ncol + 1 - off;
        if (r < nrow && c >= 0 && c < ncol) {
          const int64_t qi = (int64_t)(nrow - r + 1) * B.P + off;  // pidx(B, ncol - c, nrow - r)
#pragma unroll
          for (int m = 0; m < NV; m++) planes[(int64_t)(dst + m) * B.plane + qi] = held ? keep[rg][m] : v[rg][m];
        }
#pragma unroll
        for (int m = 0; m < NV; m++) keep[rg][m] = v[rg][m];
      }
    }
    return;
  }
  if (wave > 0) {
    // ---- producer: the inputs from their skewed copy (k_lusgs_skew), so
    // that the lanes (rows) of a step read consecutive positions
    const double* skin = a.sk + B.sk_base;
    const int p = wave - 1;
    for (int gs = p, k = 0; gs < ntot && ok; gs += FPROD, k++) {
      ok = wait_bounded(&L.consumed, gs - FRING + 1, a.err, a.spin_max);
      const int slot = gs % FRING;
      const bool fwd = gs < npad;
      const int st = fwd ? gs : gs - npad;
      auto SH = [&](int d) -> double& { return L.sh[slot][d][lane]; };
      auto SD = [&](int kk, int e) -> double& { return L.sd[slot][kk][e][lane]; };
      if (fwd) make_rec_sk<NV, VISC, true>(SH, SD, B, skin, st - lane, lane, g, a.alfa1);
      else make_rec_sk<NV, VISC, false>(SH, SD, B, skin, st - lane, lane, g, a.alfa1);
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      __builtin_amdgcn_s_waitcnt(0xc07f);  // the slot's LDS writes done (lgkmcnt 0) before publishing
      lds_put(&L.produced[p], k + 1);
    }
    return;
  }
  // ---- the sweep wave: lane t = row t; it issues first on its SIMD
  __builtin_amdgcn_s_setprio(3);
  const int t = lane;
  Chain<NV>& ch = *reinterpret_cast<Chain<NV>*>(&L);  // no other compute wave: never touched
  auto load_lds = [&](Rec<NV, 2>& r, int slot, bool fwd) {
#pragma unroll
    for (int m = 0; m < NV; m++) {
      if (fwd) r.X[m] = L.sh[slot][m][t];
      r.U[m] = L.sh[slot][F::U + m][t];
    }
    r.al0 = L.sh[slot][F::AL0][t];
    r.y0 = L.sh[slot][F::Y0][t];
    if constexpr (NV == 5) {
      r.al4 = L.sh[slot][F::AL4][t];
      r.y4 = L.sh[slot][F::Y4][t];
    } else {
      r.al4 = r.al0;
      r.y4 = r.y0;
    }
#pragma unroll
    for (int h = 0; h < 2; h++) {
#pragma unroll
      for (int m = 0; m < NV; m++) r.F[h][m] = L.sd[slot][h][F::F + m][t];
      r.n1[h] = L.sd[slot][h][F::N1][t];
      r.n2[h] = L.sd[slot][h][F::N2][t];
      r.a[h] = L.sd[slot][h][F::A][t];
      r.Lc[h] = L.sd[slot][h][F::LC][t];
      if constexpr (VISC) r.Lv[h] = L.sd[slot][h][F::LV][t];
      else r.Lv[h] = 0.0;
    }
  };
  // the backward sweep's X (the forward dU) from the backward records' X pairs
  // (pairs, and with an odd equation count the last field as single doubles).
  // A lane whose row has no cell at this step reads the nearest row that has
  // one: positions without a cell (about (nrow-1)/ncol of the field) are
  // never written, and a wave reading them would fetch lines nobody needs;
  // such a lane's value never reaches a cell.
  auto load_x = [&](double* x, int st) {
    const int sc = min(st, S - 1);
    const int tr = min(max(t, sc - (ncol - 1)), min(sc, nrow - 1));
    const uint32_t pos = (uint32_t)(sc * R + tr);
#pragma unroll
    for (int j = 0; j < F::XP; j++) {
      if (2 * j + 1 == NV) {
        x[2 * j] = bload1(rs, pos * 8u, bw + (uint32_t)j * n16);
      } else {
        const D2 v = bload2(rs, pos * 16u, bw + (uint32_t)j * n16);
        x[2 * j] = v.x;
        x[2 * j + 1] = v.y;
      }
    }
  };
  // one direction of the sweep; the direction is a compile-time constant so
  // that every register array keeps fixed indices.  The next step's record
  // is read from its slot while this step computes; in the backward sweep
  // each step's X (the forward dU) is loaded XD steps ahead, into one of XD
  // register sets (the loop runs XD steps per turn; npad is a multiple of 4).
  auto run_dir = [&](auto fwd_c) {
    constexpr bool fwd = decltype(fwd_c)::value;
    constexpr int dir = fwd ? 0 : 1;
    const int gs0 = dir * npad;
    Ahead<NV> ah = {};
    Out<NV> own = {};
    constexpr int XD = NV >= 6 ? 2 : 4;  // six-equation records leave no registers for four sets
    double xs[XD][2 * F::XP] = {};
    if (!fwd) {
      // the forward sweep's stores to the X pairs before these loads
      __threadfence();
#pragma unroll
      for (int k = 0; k < XD; k++) load_x(xs[k], k);
    }
    Rec<NV, 2> rn;
    if (ok) ok = wait_bounded(&L.produced[gs0 % FPROD], gs0 / FPROD + 1, a.err, a.spin_max);
    load_lds(rn, gs0 % FRING, fwd);
    int pv = -1;  // the count of the next step's producer, read one step earlier (-1: not read)
    auto one_step = [&](int st, double* x) {
      const int gs = gs0 + st;
      Rec<NV, 2> r = rn;
      // The next step's record.  pv was read a step ago, so the wait (an LDS
      // round trip on the chain) is taken only when that read found the
      // record missing.  Both loads below are unconditional: a load under a
      // branch makes its registers a phi, which the compiler resolves with
      // copies that wait for the load in flight and so undo the prefetch.
      // After the direction's last step the slot read is not waited for and
      // its values are never used (the next direction reloads after its own
      // wait), and load_x clamps the step to S - 1.
      if (st + 1 < npad && ok && pv <= (gs + 1) / FPROD)
        ok = wait_bounded(&L.produced[(gs + 1) % FPROD], (gs + 1) / FPROD + 1, a.err, a.spin_max);
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      load_lds(rn, (gs + 1) % FRING, fwd);
      pv = lds_get(&L.produced[(gs + 2) % FPROD]);
      if constexpr (!fwd) {
#pragma unroll
        for (int m = 0; m < NV; m++) r.X[m] = x[m];
        load_x(x, st + XD);
      }
      // forward: dU to the backward records' X pairs; backward: U + dU (the
      // record's U is the plane's value; with update) or dU into the output
      // buffer, which the writer wave stores eight steps at a time
      sweep_step<NV, VISC, fwd, 1, fwd, true>(r, own, rs, bw, n16, st, S, R, ncol, nrow, t, 0, lane, 0, false, false, g, ch,
                                        ah);
      // Slot gs is released after the step: its reads were issued a step ago
      // and the next slot's at this step's start, so the wait costs nothing
      // here, and it comes before the output buffer's writes so that it does
      // not wait for them.
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      __builtin_amdgcn_s_waitcnt(0xc07f);
      lds_put(&L.consumed, gs + 1);
      if constexpr (!fwd) {
        // chunks of eight steps; npad is a multiple of four, so the last
        // chunk may end after four
        const int b8 = st & 7, ck = st >> 3;
        if (b8 == 0 && ck > 0 && ok) ok = wait_bounded(&L.ob_taken, ck, a.err, a.spin_max);
        if (a.update) {
#pragma unroll
          for (int m = 0; m < NV; m++) L.ob[b8][m][t] = own.W[m];
        } else {
#pragma unroll
          for (int m = 0; m < NV; m++) L.ob[b8][m][t] = own.du[m];
        }
        if (b8 == 7 || st == npad - 1) {
          __atomic_signal_fence(__ATOMIC_SEQ_CST);
          __builtin_amdgcn_s_waitcnt(0xc07f);  // the chunk's writes done before it is handed over
          lds_put(&L.ob_ready, ck + 1);
        }
      }
    };
    for (int st = 0; st < npad; st += XD) {
#pragma unroll
      for (int k = 0; k < XD; k++) one_step(st + k, xs[k]);
    }
  };
  run_dir(std::true_type{});
  run_dir(std::false_type{});
}

// ----------------------------------------------------------------- scatter
template <int NV, bool VISC>
__global__ __launch_bounds__(256) void k_lusgs_scatter(LusgsArgs a) {
  using F = RecF<NV, VISC>;
  const BlockDesc B = a.blocks[a.block_list[blockIdx.y]];
  if (case_frozen(B.case_id)) return;
  const int ncol = B.nx - 1, nrow = B.ny - 1;
  // the tiles of k_lusgs_prep in the backward set's (step, row) space: eight
  // lanes read 8 consecutive output positions, a wave writes 8 cells of 8 rows
  const int S = B.sk_S;
  const int tiles_s = (S + 31) / 32;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int sb = (blockIdx.x % tiles_s) * 32 + wave * 8 + (lane >> 3);
  const int rb = (blockIdx.x / tiles_s) * 8 + (lane & 7);
  const int cb = sb - rb;
  if (sb >= S || rb >= nrow || cb < 0 || cb >= ncol) return;
  const int i = ncol - cb, j = nrow - rb;
  const int64_t n = B.sk_n;
  // the backward record of cell (i, j): position (ncol-i + nrow-j) * R + nrow-j
  const int64_t pos = (int64_t)sb * B.sk_R + rb;
  const double* out = a.sk + B.sk_base + 2 * F::NREC * n + 2 * pos;
  double* base = a.arena + B.base;
  const int64_t q = pidx(B, i, j);
#pragma unroll
  for (int m = 0; m < NV; m++) {
    const double du = out[(int64_t)(m / 2) * 2 * n + (m & 1)];
    base[(int64_t)(a.pm.dU + m) * B.plane + q] = du;
    if (a.update) {
      double* U = base + (int64_t)(a.pm.U + m) * B.plane;
      U[q] = U[q] + du;
    }
  }
}

}  // namespace ec2d
