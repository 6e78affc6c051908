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
//     moved with a cross-lane shuffle; between 64-row waves it goes through
//     LDS with one barrier per step.
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

template <int NV>
__device__ __forceinline__ void shfl_up_carry(Carry<NV>& dst, const Carry<NV>& src) {
#pragma unroll
  for (int m = 0; m < NV; m++) {
    dst.dU[m] = __shfl_up(src.dU[m], 1, 64);
    dst.U[m] = __shfl_up(src.U[m], 1, 64);
  }
  dst.s = __shfl_up(src.s, 1, 64);
  dst.Lc = __shfl_up(src.Lc, 1, 64);
  dst.Lv = __shfl_up(src.Lv, 1, 64);
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

template <int NV, bool VISC>
__global__ __launch_bounds__(1024) void k_lusgs(LusgsArgs a) {
  constexpr int CW = 2 * NV + 3;
  __shared__ double xfer[2][16][CW];
  const BlockDesc B = a.blocks[a.block_list[blockIdx.x]];
  const CaseParams cp = a.cps[B.case_id];
  const PlaneMap pm = a.pm;
  const double g = cp.gamma;
  double* base = a.arena + B.base;
  auto PL = [&](int slot) { return base + (int64_t)slot * B.plane; };
  const int ncol = B.nx - 1, nrow = B.ny - 1;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nwave = (nrow + 63) >> 6;
  const bool multi = nwave > 1;
  const int r = tid;
  const bool row_ok = r < nrow;
  const int nsteps = ncol + nrow - 1;
  double* Up[NV];
  double* dUp[NV];
  const double* Rp[NV];
#pragma unroll
  for (int m = 0; m < NV; m++) {
    Up[m] = PL(pm.U + m);
    dUp[m] = PL(pm.dU + m);
    Rp[m] = PL(pm.Res + m);
  }
  const double *vol = PL(pm.vol), *dtp = PL(pm.dt), *Lci = PL(pm.Lci), *Lcj = PL(pm.Lcj);
  const double *Lvi = PL(pm.Lvi), *Lvj = PL(pm.Lvj), *si = PL(pm.si), *sj = PL(pm.sj);
  const double *ni1 = PL(pm.ni1), *ni2 = PL(pm.ni2), *nj1 = PL(pm.nj1), *nj2 = PL(pm.nj2);

  // diagonal of the cell (sub_LU_SGS.f90:31-39, 76-83)
  auto diag = [&](int64_t q, double* al) {
    double v = vol[q];
    double alfa = v / dtp[q] + 1.0 * (Lci[q] + Lcj[q]) + v * a.alfa1;
    if constexpr (VISC) alfa = alfa + 2.0 * (Lvi[q] + Lvj[q]);
#pragma unroll
    for (int m = 0; m < NV; m++) al[m] = alfa;
    if constexpr (NV == 5) al[4] = al[4] + Lvi[q] + Lvj[q];
  };

  // ------------------------------------------------------------ forward sweep
  Carry<NV> side = {}, up = {}, low = {};
  for (int s = 0; s < nsteps; s++) {
    shfl_up_carry<NV>(low, up);
    if (multi && lane == 0 && wave > 0) get_carry<NV>(low, xfer[(s + 1) & 1][wave - 1]);
    const int i = s - r + 1, j = r + 1;
    if (row_ok && i >= 1 && i <= ncol) {
      int64_t q = pidx(B, i, j);
      double al[NV], acc[NV], Uc[NV];
      diag(q, al);
#pragma unroll
      for (int m = 0; m < NV; m++) {
        acc[m] = Rp[m][q];
        Uc[m] = Up[m][q];
      }
      double dui[NV], duj[NV];
      if (i != 1) {
        double DF[NV];
        comput_dfn<NV>(DF, side.U, side.dU, ni1[q], ni2[q], g);
        double sq = si[q];
#pragma unroll
        for (int m = 0; m < NV; m++) {
          dui[m] = 0.5 * (DF[m] * sq + 1.0 * side.Lc * side.dU[m]);
          if constexpr (VISC) dui[m] = dui[m] + side.Lv * side.dU[m];
        }
      } else {
#pragma unroll
        for (int m = 0; m < NV; m++) dui[m] = 0.0;
      }
      if (j != 1) {
        double DF[NV];
        comput_dfn<NV>(DF, low.U, low.dU, nj1[q], nj2[q], g);
        double sq = sj[q];
#pragma unroll
        for (int m = 0; m < NV; m++) {
          duj[m] = 0.5 * (DF[m] * sq + 1.0 * low.Lc * low.dU[m]);
          if constexpr (VISC) duj[m] = duj[m] + low.Lv * low.dU[m];
        }
      } else {
#pragma unroll
        for (int m = 0; m < NV; m++) duj[m] = 0.0;
      }
#pragma unroll
      for (int m = 0; m < NV; m++) acc[m] = acc[m] + dui[m] + duj[m];
#pragma unroll
      for (int m = 0; m < NV; m++) {
        double du = acc[m] / al[m];
        dUp[m][q] = du;
        side.dU[m] = du;
        side.U[m] = Uc[m];
        up.dU[m] = du;
        up.U[m] = Uc[m];
      }
      side.Lc = Lci[q];
      up.Lc = Lcj[q];
      if constexpr (VISC) {
        side.Lv = Lvi[q];
        up.Lv = Lvj[q];
      }
    }
    if (multi) {
      if (lane == 63) put_carry<NV>(xfer[s & 1][wave], up);
      __syncthreads();
    }
  }

  __threadfence();
  __syncthreads();

  // ----------------------------------------------------------- backward sweep
  // lane r owns row j = nrow - r; at step s it handles i = ncol - (s - r)
  for (int s = 0; s < nsteps; s++) {
    shfl_up_carry<NV>(low, up);
    if (multi && lane == 0 && wave > 0) get_carry<NV>(low, xfer[(s + 1) & 1][wave - 1]);
    const int i = ncol - (s - r), j = nrow - r;
    if (row_ok && i >= 1 && i <= ncol) {
      int64_t q = pidx(B, i, j);
      double al[NV], Uc[NV], du0[NV];
      diag(q, al);
#pragma unroll
      for (int m = 0; m < NV; m++) {
        Uc[m] = Up[m][q];
        du0[m] = dUp[m][q];
      }
      double dui[NV], duj[NV];
      if (i != ncol) {
        double DF[NV];
        comput_dfn<NV>(DF, side.U, side.dU, ni1[q], ni2[q], g);
#pragma unroll
        for (int m = 0; m < NV; m++) {
          dui[m] = -0.5 * (DF[m] * side.s - 1.0 * side.Lc * side.dU[m]);
          if constexpr (VISC) dui[m] = dui[m] + side.Lv * side.dU[m];
        }
      } else {
#pragma unroll
        for (int m = 0; m < NV; m++) dui[m] = 0.0;
      }
      if (j != nrow) {
        double DF[NV];
        comput_dfn<NV>(DF, low.U, low.dU, nj1[q], nj2[q], g);
#pragma unroll
        for (int m = 0; m < NV; m++) {
          duj[m] = -0.5 * (DF[m] * low.s - 1.0 * low.Lc * low.dU[m]);
          if constexpr (VISC) duj[m] = duj[m] + low.Lv * low.dU[m];
        }
      } else {
#pragma unroll
        for (int m = 0; m < NV; m++) duj[m] = 0.0;
      }
#pragma unroll
      for (int m = 0; m < NV; m++) {
        double du = du0[m] + (dui[m] + duj[m]) / al[m];
        dUp[m][q] = du;
        if (a.update) Up[m][q] = Uc[m] + du;
        side.dU[m] = du;
        side.U[m] = Uc[m];
        up.dU[m] = du;
        up.U[m] = Uc[m];
      }
      side.s = si[q];
      side.Lc = Lci[q];
      up.s = sj[q];
      up.Lc = Lcj[q];
      if constexpr (VISC) {
        side.Lv = Lvi[q];
        up.Lv = Lvj[q];
      }
    }
    if (multi) {
      if (lane == 63) put_carry<NV>(xfer[s & 1][wave], up);
      __syncthreads();
    }
  }
}

}  // namespace ec2d
