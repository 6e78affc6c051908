// Ghost-cell fill: Boundary_condition_onemesh (sub_boundary.f90:10-34),
// update_buffer_onemesh (:45-153) and the corner extrapolation
// U_average_conner (:595-612), fused per block.
//
// The reference fills ghosts sequentially: subfaces of a block in file order,
// blocks in order, then all 1:1 interfaces, then the four corner cells.  A
// subface's points are independent of each other, but a subface may read
// ghost cells that an earlier subface of the same block wrote (the point
// loops include the subface end points, :325-327, :478-479), and interfaces
// overwrite what boundary conditions wrote.  One workgroup per block keeps
// that order with a barrier before every subface that depends on the ones
// since the last barrier (SubfaceDesc::sync, from the host's analysis of the
// cells each subface reads and writes), and before the corners; between
// barriers the points of all subfaces form one index space over the
// workgroup's threads; blocks run in parallel
// because every stage reads only the own block's cells, except the interface
// copy, which reads neighbour INTERIOR cells that this kernel never writes.
#include "ec2d_device.h"
#include "ec2d_physics.h"

namespace ec2d {

struct GhostArgs {
  const BlockDesc* blocks;
  const int* block_list;
  const SubfaceDesc* subs;
  const int* units;       // per block (at bc_off): its subfaces in fill order, boundary conditions first
  const double* normals;  // per boundary point (n1, n2)
  const CaseParams* cps;
  double* arena;
  PlaneMap pm;
  int nvar;
  int do_bc, do_exchange;
};

// Every point reads only cells it does not write (interior or inner layers
// on the source side, ghost layers on the other), but the compiler cannot
// prove that the planes do not overlap, so a load placed after a store waits
// for the store's operands: a point that interleaved them paid one memory
// round trip per variable and layer.  Each branch therefore loads all it
// reads first, then computes and stores.
template <int NV>
__device__ void bc_point(const GhostArgs& a, const BlockDesc& B, const CaseParams& cp, const SubfaceDesc& sf, int i,
                         int j, int pt, double* const* U) {
  const double g = cp.gamma;
  const int face = sf.face;
  const int type = sf.neighb;
  if (type == -10 || type == -30) {  // wall (sub_boundary.f90:306-386), symmetry / slip wall (:391-451)
    int64_t q1[EC2D_LAP], q2[EC2D_LAP];
    double v[EC2D_LAP][NV], dwv[EC2D_LAP];
#pragma unroll
    for (int k = 1; k <= EC2D_LAP; k++) {
      int i1, j1, i2, j2;
      if (face == 1) { i1 = i + k - 1; j1 = j; i2 = i - k; j2 = j; }
      else if (face == 2) { i1 = i; j1 = j + k - 1; i2 = i; j2 = j - k; }
      else if (face == 3) { i1 = i - k; j1 = j; i2 = i + k - 1; j2 = j; }
      else { i1 = i; j1 = j - k; i2 = i; j2 = j + k - 1; }
      q1[k - 1] = pidx(B, i1, j1);
      q2[k - 1] = pidx(B, i2, j2);
#pragma unroll
      for (int m = 0; m < NV; m++) v[k - 1][m] = U[m][q1[k - 1]];
      if constexpr (NV == 6) {
        if (type == -10) dwv[k - 1] = (a.arena + B.base + (int64_t)a.pm.dw * B.plane)[q1[k - 1]];
      }
    }
    if (type == -10) {
#pragma unroll
      for (int k = 0; k < EC2D_LAP; k++) {
        const double* w = v[k];
        double d1 = 0.0, s0, s1, s2, s3;
        if (cp.Twall <= 0) {
          s0 = w[0];
          s1 = -w[1];
          s2 = -w[2];
          s3 = w[3];
        } else {
          d1 = w[0];
          double u1 = w[1] / d1, v1 = w[2] / d1;
          double p1 = (w[3] - 0.5 * d1 * (u1 * u1 + v1 * v1)) * (g - 1.0);
          double T1 = g * cp.Ma * cp.Ma * p1 / d1;
          double p2 = p1, T2 = 2.0 * cp.Twall - T1, u2 = -u1, v2 = -v1;
          double d2 = g * cp.Ma * cp.Ma * p2 / T2;
          s0 = d2;
          s1 = d2 * u2;
          s2 = d2 * v2;
          s3 = p2 / (g - 1.0) + 0.5 * d2 * (u2 * u2 + v2 * v2);
        }
        U[0][q2[k]] = s0;
        U[1][q2[k]] = s1;
        U[2][q2[k]] = s2;
        U[3][q2[k]] = s3;
        if constexpr (NV == 5) U[4][q2[k]] = 0.0;
        if constexpr (NV == 6) {
          // the ghost state just written, as the reference reads it back
          double d2 = s0;
          double T2 = (s3 - 0.5 * (s1 * s1 + s2 * s2) / d2) / (cp.Cv * d2);
          double Amu2 = (1.0 + cp.Tsb) * sqrt(T2 * T2 * T2) / (cp.Tsb + T2);
          double wt = 60.0 * Amu2 / (d2 * 0.075 * (dwv[k] * dwv[k]) * cp.Re * cp.Re);
          U[4][q2[k]] = -w[4];
          U[5][q2[k]] = (d1 + d2) * wt - w[5];
        }
      }
    } else {
      double n1 = a.normals[2 * (sf.norm_off + pt)], n2 = a.normals[2 * (sf.norm_off + pt) + 1];
#pragma unroll
      for (int k = 0; k < EC2D_LAP; k++) {
        const double* w = v[k];
        double Vn = w[1] * n1 + w[2] * n2;
        U[0][q2[k]] = w[0];
        U[1][q2[k]] = w[1] - 2.0 * Vn * n1;
        U[2][q2[k]] = w[2] - 2.0 * Vn * n2;
        U[3][q2[k]] = w[3];
#pragma unroll
        for (int m = 4; m < NV; m++) U[m][q2[k]] = w[m];
      }
    }
  } else {  // far field / inlet / outlet, sub_boundary.f90:457-581
    int i1 = i, j1 = j;
    if (face == 3) i1 = i - 1;
    else if (face == 4) j1 = j - 1;
    int64_t q1 = pidx(B, i1, j1);
    double w[NV];
#pragma unroll
    for (int m = 0; m < NV; m++) w[m] = U[m][q1];
    double n1 = a.normals[2 * (sf.norm_off + pt)], n2 = a.normals[2 * (sf.norm_off + pt) + 1];
    double d1 = w[0], u1 = w[1] / d1, v1 = w[2] / d1;
    double p1 = (w[3] - 0.5 * d1 * (u1 * u1 + v1 * v1)) * (g - 1.0);
    double c1 = sqrt(g * p1 / d1);
    double Ma_n = (u1 * n1 + v1 * n2) / c1;
    const double d_inf = 1.0;
    for (int k = 1; k <= EC2D_LAP; k++) {
      int i2 = i, j2 = j;
      if (face == 1) i2 = i - k;
      else if (face == 2) j2 = j - k;
      else if (face == 3) i2 = i + k - 1;
      else j2 = j + k - 1;
      int64_t q2 = pidx(B, i2, j2);
      if (type == -50 || (type == -20 && Ma_n >= 0.0)) {
        if (Ma_n > 1.0 || cp.P_outlet < 0 || type == -20) {
#pragma unroll
          for (int m = 0; m < NV; m++) U[m][q2] = w[m];
        } else {
          double pb = cp.P_outlet;
          double db = d1 + (cp.P_outlet - p1) / (c1 * c1);
          double ub = u1 + (p1 - cp.P_outlet) / (d1 * c1) * n1;
          double vb = v1 + (p1 - cp.P_outlet) / (d1 * c1) * n2;
          double p2 = 2.0 * pb - p1, d2 = 2.0 * db - d1, u2 = 2.0 * ub - u1, v2 = 2.0 * vb - v1;
          U[0][q2] = d2;
          U[1][q2] = d2 * u2;
          U[2][q2] = d2 * v2;
          U[3][q2] = p2 / (g - 1.0) + 0.5 * d2 * (u2 * u2 + v2 * v2);
#pragma unroll
          for (int m = 4; m < NV; m++) U[m][q2] = w[m];
        }
      } else {
        if (cp.P_outlet < 0 || Ma_n <= -1.0) {
          U[0][q2] = d_inf;
          U[1][q2] = d_inf * cp.u_inf;
          U[2][q2] = d_inf * cp.v_inf;
          U[3][q2] = cp.p_inf / (g - 1.0) + 0.5 * d_inf * (cp.u_inf * cp.u_inf + cp.v_inf * cp.v_inf);
        } else {
          double pb = 0.5 * (p1 + cp.p_inf - d1 * c1 * ((cp.u_inf - u1) * n1 + (cp.v_inf - v1) * n2));
          double db = d_inf + (pb - cp.p_inf) / (c1 * c1);
          double ub = cp.u_inf - (cp.p_inf - pb) / (d1 * c1) * n1;
          double vb = cp.v_inf - (cp.p_inf - pb) / (d1 * c1) * n2;
          double p2 = 2.0 * pb - p1, d2 = 2.0 * db - d1, u2 = 2.0 * ub - u1, v2 = 2.0 * vb - v1;
          U[0][q2] = d2;
          U[1][q2] = d2 * u2;
          U[2][q2] = d2 * v2;
          U[3][q2] = p2 / (g - 1.0) + 0.5 * d2 * (u2 * u2 + v2 * v2);
        }
        if constexpr (NV == 5) U[4][q2] = cp.vt_inf / cp.Re;
        if constexpr (NV == 6) {
          U[4][q2] = d_inf * cp.Kt_inf;
          U[5][q2] = d_inf * cp.Wt_inf;
        }
      }
    }
  }
}

template <int NV>
__device__ __forceinline__ void corner(double* const* U, int64_t q1, int64_t q2, int64_t q3, int64_t q4, double Cv) {
  double d1 = U[0][q1], uu1 = U[1][q1] / d1, v1 = U[2][q1] / d1, T1 = (U[3][q1] - (uu1 * U[1][q1] + v1 * U[2][q1]) * 0.5) / (d1 * Cv);
  double d2 = U[0][q2], uu2 = U[1][q2] / d2, v2 = U[2][q2] / d2, T2 = (U[3][q2] - (uu2 * U[1][q2] + v2 * U[2][q2]) * 0.5) / (d2 * Cv);
  double d3 = U[0][q3], uu3 = U[1][q3] / d3, v3 = U[2][q3] / d3, T3 = (U[3][q3] - (uu3 * U[1][q3] + v3 * U[2][q3]) * 0.5) / (d3 * Cv);
  double d4 = d1 + d3 - d2, uu4 = uu1 + uu3 - uu2, v4 = v1 + v3 - v2, T4 = T1 + T3 - T2;
  double t4[NV > 4 ? NV - 4 : 1];
#pragma unroll
  for (int m = 4; m < NV; m++) t4[m - 4] = U[m][q1] + U[m][q3] - U[m][q2];
  U[0][q4] = d4;
  U[1][q4] = d4 * uu4;
  U[2][q4] = d4 * v4;
  U[3][q4] = d4 * (Cv * T4 + (uu4 * uu4 + v4 * v4) * 0.5);
#pragma unroll
  for (int m = 4; m < NV; m++) U[m][q4] = t4[m - 4];
}

template <int NV>
__global__ __launch_bounds__(1024) void k_ghost(GhostArgs a) {
  const int bidx = a.block_list[blockIdx.x];
  const BlockDesc B = a.blocks[bidx];
  if (case_frozen(B.case_id)) return;
  const CaseParams cp = a.cps[B.case_id];
  double* U[NV];
#pragma unroll
  for (int m = 0; m < NV; m++) U[m] = a.arena + B.base + (int64_t)(a.pm.U + m) * B.plane;
  const int tid = threadIdx.x;
  // Phases: a unit with sync = 1 starts a new one.  The points of all units
  // of a phase are independent of each other, so the threads take them as
  // one index space (one memory round trip per phase, not per subface).
  auto on = [&](const SubfaceDesc& sf) { return sf.neighb < 0 ? a.do_bc != 0 : a.do_exchange != 0; };
  auto npoints = [&](const SubfaceDesc& sf) {
    if (!on(sf)) return 0;
    if (sf.neighb < 0) return (sf.iend - sf.ist + 1) * (sf.jend - sf.jst + 1);
    return (sf.iend - sf.ist + sf.jend - sf.jst) * EC2D_LAP;
  };
  const int* un = a.units + B.bc_off;
  for (int u0 = 0; u0 < B.nsub;) {
    int u1 = u0 + 1;
    while (u1 < B.nsub && !a.subs[un[u1]].sync) u1++;
    int total = 0;
    for (int u = u0; u < u1; u++) total += npoints(a.subs[un[u]]);
    for (int p = tid; p < total; p += blockDim.x) {
      int u = u0, lp = p;
      for (;;) {
        const int np = npoints(a.subs[un[u]]);
        if (lp < np) break;
        lp -= np;
        u++;
      }
      const SubfaceDesc sf = a.subs[un[u]];
      if (sf.neighb < 0) {
        const int ni = sf.iend - sf.ist + 1;
        bc_point<NV>(a, B, cp, sf, sf.ist + lp % ni, sf.jst + lp / ni, lp, U);
      } else {
        const BlockDesc B1 = a.blocks[sf.nb_block];
        const double* U1[NV];
#pragma unroll
        for (int m = 0; m < NV; m++) U1[m] = a.arena + B1.base + (int64_t)(a.pm.U + m) * B1.plane;
        const int n1 = sf.iend - sf.ist + sf.jend - sf.jst;
        const int kk = lp % n1 + 1, l = lp / n1 + 1;  // cell along the face (1..n1), layer (1..LAP)
        const int k1 = (sf.orient == 2) ? kk : n1 + 1 - kk;
        // source cell: slot (k1, l) of the neighbour's buffer (sub_boundary.f90:76-108)
        int si_, sj_;
        if (sf.nb_face == 1) { si_ = sf.nb_ist + l - 1; sj_ = sf.nb_jst + k1 - 1; }
        else if (sf.nb_face == 2) { si_ = sf.nb_ist + k1 - 1; sj_ = sf.nb_jst + l - 1; }
        else if (sf.nb_face == 3) { si_ = sf.nb_iend - EC2D_LAP - 1 + l; sj_ = sf.nb_jst + k1 - 1; }
        else { si_ = sf.nb_ist + k1 - 1; sj_ = sf.nb_jend - EC2D_LAP - 1 + l; }
        // destination ghost cell (:118-135)
        int di, dj;
        if (sf.face == 1) { di = sf.ist - EC2D_LAP - 1 + l; dj = sf.jst + kk - 1; }
        else if (sf.face == 2) { di = sf.ist + kk - 1; dj = sf.jst - EC2D_LAP - 1 + l; }
        else if (sf.face == 3) { di = sf.iend + l - 1; dj = sf.jst + kk - 1; }
        else { di = sf.ist + kk - 1; dj = sf.jend + l - 1; }
        const int64_t qs = pidx(B1, si_, sj_), qd = pidx(B, di, dj);
        double w[NV];
#pragma unroll
        for (int m = 0; m < NV; m++) w[m] = U1[m][qs];
#pragma unroll
        for (int m = 0; m < NV; m++) U[m][qd] = w[m];
      }
    }
    u0 = u1;
    if (u0 < B.nsub) __syncthreads();
  }
  if (a.do_exchange) {
    __syncthreads();
    // the four corners are independent: each writes its corner cell and
    // reads three cells next to it, none of them a corner cell (nx, ny >= 2)
    const int nx = B.nx, ny = B.ny;
    if (tid == 0) corner<NV>(U, pidx(B, 1, 0), pidx(B, 1, 1), pidx(B, 0, 1), pidx(B, 0, 0), cp.Cv);
    else if (tid == 64) corner<NV>(U, pidx(B, 1, ny), pidx(B, 1, ny - 1), pidx(B, 0, ny - 1), pidx(B, 0, ny), cp.Cv);
    else if (tid == 128) corner<NV>(U, pidx(B, nx - 1, 0), pidx(B, nx - 1, 1), pidx(B, nx, 1), pidx(B, nx, 0), cp.Cv);
    else if (tid == 192) corner<NV>(U, pidx(B, nx - 1, ny), pidx(B, nx - 1, ny - 1), pidx(B, nx, ny - 1), pidx(B, nx, ny), cp.Cv);
  }
}

}  // namespace ec2d
