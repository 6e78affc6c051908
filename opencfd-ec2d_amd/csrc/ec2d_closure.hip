// Learned turbulence closure: a small MLP that predicts nu_t / nu per cell
// from six local, dimensionless features, evaluated and trained with BF16
// MFMA (FP32 accumulation).  The reference has no counterpart; the closure
// runs inside an SA case: the SA equation keeps evolving nu~ and the MLP's
// mu_t replaces SA's mu_t in the mean-flow viscous flux and the spectral
// radii wherever the cell's features lie inside the training range.  Cells
// outside it keep SA's mu_t (per-cell fallback); disabling the closure (e.g.
// when the host sees the residual stall) falls back to SA everywhere.
//
// Features of cell (i, j) (gradients by central differences in index space,
// as the SA source term takes them, sub_turbulence_SA.f90:128-150):
//   f0 = log(1 + dw |u| / nu)             wall-distance Reynolds number
//   f1 = log(1 + |S| dw / |u|)            strain magnitude over |u| / dw
//   f2 = log(1 + |omega| dw / |u|)        vorticity magnitude over |u| / dw
//   f3 = asinh(dw (u . grad p) / (rho |u|^3))   streamwise pressure-gradient parameter
//   f4 = |u| / c                          local Mach number
//   f5 = mu Re                            molecular viscosity over its free-stream value
// with nu = mu / rho (mu by Sutherland's law including 1/Re), omega = v_x -
// u_y, |S| = sqrt(2 u_x^2 + 2 v_y^2 + (u_y + v_x)^2).  The logarithms and
// asinh only compress the heavy tails; the network sees (f - mean) * inv_std.
// Network: 6 -> 64 -> 64 -> 64 -> 1, tanh hidden layers, softplus output.
// Layout: weights transposed, Wt[n][k] (bf16), so that the MFMA B operand of
// a lane is 4 consecutive k of one column; W1 is padded to K = 16, W4 to
// N = 16 (only column 0 is nonzero).  One wave evaluates 16 cells per pass
// with v_mfma_f32_16x16x16_bf16: A[m][k] in lane (m = l%16, k = 4(l/16)+i),
// B[k][n] in lane (k = 4(l/16)+i, n = l%16), D[m][n] in lane
// (m = 4(l/16)+i, n = l%16).
#include "ec2d_device.h"

namespace ec2d {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int CL_NF = 6;      // features
constexpr int CL_H = 64;      // hidden width
// FP32 parameter vector: W1 (6x64), b1, W2 (64x64), b2, W3 (64x64), b3, W4 (64), b4
constexpr int CL_OW1 = 0, CL_OB1 = CL_OW1 + CL_NF * CL_H, CL_OW2 = CL_OB1 + CL_H, CL_OB2 = CL_OW2 + CL_H * CL_H,
              CL_OW3 = CL_OB2 + CL_H, CL_OB3 = CL_OW3 + CL_H * CL_H, CL_OW4 = CL_OB3 + CL_H, CL_OB4 = CL_OW4 + CL_H,
              CL_NPARAM = CL_OB4 + 1;
// parameters as the kernels read them: bf16 transposed weights + fp32 biases
struct ClosureNet {
  unsigned short w1t[64 * 16];  // [n][k], k < 6 used
  unsigned short w2t[64 * 64];
  unsigned short w3t[64 * 64];
  unsigned short w4t[16 * 64];  // [n][k], n = 0 used
  float b1[64], b2[64], b3[64], b4;
  float mean[CL_NF], inv_std[CL_NF], lo[CL_NF], hi[CL_NF];  // lo / hi: normalised training range
};

// round to nearest even, as the host-side reference does
__device__ __forceinline__ unsigned short f2bf(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// C[16][64] += A[16][K] * B[K][64]: A row-major bf16 in LDS (row stride lda);
// B through bfn(k, n) -> bf16 bits.  C[nb][i] = out[4q+i][16 nb + r].
template <int K, class BF>
__device__ __forceinline__ void mm_16x64(f32x4 (&C)[4], const unsigned short* A, int lda, BF bfn, int lane) {
  const int r = lane & 15, q = lane >> 4;
#pragma unroll
  for (int kb = 0; kb < K / 16; kb++) {
    bf16x4 a;
#pragma unroll
    for (int i = 0; i < 4; i++) a[i] = (short)A[r * lda + kb * 16 + 4 * q + i];
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
      bf16x4 b;
#pragma unroll
      for (int i = 0; i < 4; i++) b[i] = (short)bfn(kb * 16 + 4 * q + i, nb * 16 + r);
      C[nb] = mfma16(a, b, C[nb]);
    }
  }
}

// out[16][64] (D layout) -> LDS bf16 row-major [16][64]
__device__ __forceinline__ void store_act(unsigned short* S, const f32x4 (&H)[4], int lane) {
  const int r = lane & 15, q = lane >> 4;
#pragma unroll
  for (int nb = 0; nb < 4; nb++)
#pragma unroll
    for (int i = 0; i < 4; i++) S[(4 * q + i) * 64 + nb * 16 + r] = f2bf(H[nb][i]);
}

// features of interior cell (i, j), raw (not normalised); false if undefined
__device__ __forceinline__ bool closure_features(float* f, const BlockDesc& B, const double* base, const PlaneMap& pm,
                                                 const CaseParams& cp, int i, int j) {
  auto PL = [&](int slot) { return base + (int64_t)slot * B.plane; };
  const double *U0 = PL(pm.U), *U1 = PL(pm.U + 1), *U2 = PL(pm.U + 2), *U3 = PL(pm.U + 3), *x1 = PL(pm.x1),
               *y1 = PL(pm.y1);
  const double g = cp.gamma;
  auto uu = [&](int64_t p) { return U1[p] / U0[p]; };
  auto vv = [&](int64_t p) { return U2[p] / U0[p]; };
  auto pr = [&](int64_t p) {
    const double d = U0[p], u = U1[p] / d, v = U2[p] / d;
    return (g - 1.0) * (U3[p] - 0.5 * d * (u * u + v * v));
  };
  const int64_t q = pidx(B, i, j), qe = pidx(B, i + 1, j), qw = pidx(B, i - 1, j), qn = pidx(B, i, j + 1),
                qs = pidx(B, i, j - 1);
  const double Dix = (x1[qe] - x1[qw]) * 0.5, Diy = (y1[qe] - y1[qw]) * 0.5;
  const double Djx = (x1[qn] - x1[qs]) * 0.5, Djy = (y1[qn] - y1[qs]) * 0.5;
  const double Ds = 1.0 / (Dix * Djy - Djx * Diy);
  auto grad = [&](double ve, double vw, double vn, double vs, double& gx, double& gy) {
    const double Di = (ve - vw) * 0.5, Dj = (vn - vs) * 0.5;
    gx = (Di * Djy - Dj * Diy) * Ds;
    gy = (-Di * Djx + Dj * Dix) * Ds;
  };
  double ux, uy, vx, vy, px, py;
  grad(uu(qe), uu(qw), uu(qn), uu(qs), ux, uy);
  grad(vv(qe), vv(qw), vv(qn), vv(qs), vx, vy);
  grad(pr(qe), pr(qw), pr(qn), pr(qs), px, py);
  const double d = U0[q], u = uu(q), v = vv(q), p = pr(q);
  const double T = (U3[q] - 0.5 * d * (u * u + v * v)) / (cp.Cv * d);
  const double mu = 1.0 / cp.Re * (1.0 + cp.Tsb) * sqrt(T * T * T) / (cp.Tsb + T);
  const double nu = mu / d;
  const double dw = PL(pm.dw)[q];
  const double om = fabs(vx - uy);
  const double S = sqrt(2.0 * ux * ux + 2.0 * vy * vy + (uy + vx) * (uy + vx));
  const double um = sqrt(u * u + v * v);
  const double c = sqrt(g * p / d);
  f[0] = (float)log1p(dw * um / nu);
  f[1] = (float)log1p(S * dw / um);
  f[2] = (float)log1p(om * dw / um);
  f[3] = (float)asinh(dw * (u * px + v * py) / (d * um * um * um));
  f[4] = (float)(um / c);
  f[5] = (float)(mu * cp.Re);
  bool ok = true;
#pragma unroll
  for (int k = 0; k < CL_NF; k++) ok = ok && isfinite(f[k]);
  return ok;
}

// one 16-cell pass of the forward network for this wave: X (bf16 [16][16],
// normalised features in columns 0..5) -> z4 (pre-softplus output) of row
// 4q+i in lanes with r == 0; keeps the hidden activations H1..H3 (fp32, D
// layout) for the backward pass and their bf16 copies in S1..S3.
struct Fwd {
  f32x4 H1[4], H2[4], H3[4], Z4;
};

__device__ __forceinline__ void closure_forward(Fwd& F, const ClosureNet& N, const unsigned short* X,
                                                unsigned short* S1, unsigned short* S2, unsigned short* S3,
                                                int lane) {
  const int r = lane & 15;
  const f32x4 z = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int nb = 0; nb < 4; nb++) F.H1[nb] = F.H2[nb] = F.H3[nb] = z;
  mm_16x64<16>(F.H1, X, 16, [&](int k, int n) { return N.w1t[n * 16 + k]; }, lane);
#pragma unroll
  for (int nb = 0; nb < 4; nb++)
#pragma unroll
    for (int i = 0; i < 4; i++) F.H1[nb][i] = tanhf(F.H1[nb][i] + N.b1[nb * 16 + r]);
  store_act(S1, F.H1, lane);
  __builtin_amdgcn_wave_barrier();
  mm_16x64<64>(F.H2, S1, 64, [&](int k, int n) { return N.w2t[n * 64 + k]; }, lane);
#pragma unroll
  for (int nb = 0; nb < 4; nb++)
#pragma unroll
    for (int i = 0; i < 4; i++) F.H2[nb][i] = tanhf(F.H2[nb][i] + N.b2[nb * 16 + r]);
  store_act(S2, F.H2, lane);
  __builtin_amdgcn_wave_barrier();
  mm_16x64<64>(F.H3, S2, 64, [&](int k, int n) { return N.w3t[n * 64 + k]; }, lane);
#pragma unroll
  for (int nb = 0; nb < 4; nb++)
#pragma unroll
    for (int i = 0; i < 4; i++) F.H3[nb][i] = tanhf(F.H3[nb][i] + N.b3[nb * 16 + r]);
  store_act(S3, F.H3, lane);
  __builtin_amdgcn_wave_barrier();
  // output layer as one 16x16 tile (column 0 of W4t is the real one)
  {
    const int q = lane >> 4;
    f32x4 acc = z;
#pragma unroll
    for (int kb = 0; kb < 4; kb++) {
      bf16x4 a, b;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        a[i] = (short)S3[r * 64 + kb * 16 + 4 * q + i];
        b[i] = (short)N.w4t[r * 64 + kb * 16 + 4 * q + i];
      }
      acc = mfma16(a, b, acc);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) acc[i] = acc[i] + N.b4;
    F.Z4 = acc;
  }
}

__device__ __forceinline__ float softplus(float z) { return fmaxf(z, 0.f) + log1pf(expf(-fabsf(z))); }

// load the network into LDS (cooperatively, the whole workgroup)
__device__ __forceinline__ void load_net(ClosureNet& S, const ClosureNet* G) {
  const unsigned* src = reinterpret_cast<const unsigned*>(G);
  unsigned* dst = reinterpret_cast<unsigned*>(&S);
  for (int k = threadIdx.x; k < (int)(sizeof(ClosureNet) / 4); k += blockDim.x) dst[k] = src[k];
  __syncthreads();
}

struct ClosureArgs {
  const BlockDesc* blocks;
  const long long* cells; // interior cells of the fine level: block << 40 | j << 20 | i
  int ncell;
  const CaseParams* cps;
  double* arena;          // inference writes planes pm.clo (mu_t) and pm.clo + 1 (1: in range)
  PlaneMap pm;
  const ClosureNet* net;
  float* data;            // collection: 8 floats per listed cell: f0..f5, target nu_t/nu, valid
  int standin;            // 1: the closure is the eddy-viscosity model: every listed cell gets mu_t
                          // from its features clamped into the training range (mask: inside it)
  double relax;           // model 4: mu_t <- mu_t + relax (mu_t(network) - mu_t); 1: no relaxation
};

__device__ __forceinline__ void unpack_cell(long long c, int& b, int& i, int& j) {
  b = (int)(c >> 40);
  j = (int)((c >> 20) & 0xFFFFF);
  i = (int)(c & 0xFFFFF);
}

// Training data from the current SA state: features and the target
// nu_t / nu = fv1 * d * nu~ / mu (sub_turbulence_SA.f90:24-33).
__global__ __launch_bounds__(256) void k_closure_collect(ClosureArgs a) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= a.ncell) return;
  int b, i, j;
  unpack_cell(a.cells[c], b, i, j);
  const BlockDesc B = a.blocks[b];
  const CaseParams cp = a.cps[B.case_id];
  const double* base = a.arena + B.base;
  float f[CL_NF];
  const bool ok = closure_features(f, B, base, a.pm, cp, i, j);
  const int64_t q = pidx(B, i, j);
  auto PL = [&](int slot) { return base + (int64_t)slot * B.plane; };
  const double d = PL(a.pm.U)[q], u = PL(a.pm.U + 1)[q] / d, v = PL(a.pm.U + 2)[q] / d;
  const double T = (PL(a.pm.U + 3)[q] - 0.5 * d * (u * u + v * v)) / (cp.Cv * d);
  const double mu = 1.0 / cp.Re * (1.0 + cp.Tsb) * sqrt(T * T * T) / (cp.Tsb + T);
  const double vt = PL(a.pm.U + 4)[q];
  const double X = d * vt / mu, X3 = X * X * X;
  const double fv1 = X3 / (X3 + 7.1 * 7.1 * 7.1);
  const double target = fv1 * d * vt / mu;
  float* o = a.data + (int64_t)c * 8;
#pragma unroll
  for (int k = 0; k < CL_NF; k++) o[k] = f[k];
  o[6] = (float)target;
  o[7] = (ok && isfinite(target) && target >= 0.0) ? 1.f : 0.f;
}

// Inference over the listed cells: 64 cells per workgroup (16 per wave).
__global__ __launch_bounds__(256) void k_closure_infer(ClosureArgs a) {
  __shared__ ClosureNet N;
  __shared__ unsigned short sX[4][16 * 16], s1[4][16 * 64], s2[4][16 * 64], s3[4][16 * 64];
  __shared__ int sok[4][16];
  load_net(N, a.net);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int c0 = blockIdx.x * 64 + wave * 16;
  // features: lanes 0..15 one cell each
  if (lane < 16) {
    const int c = c0 + lane;
    float f[CL_NF] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    bool ok = false;
    if (c < a.ncell) {
      int b, i, j;
      unpack_cell(a.cells[c], b, i, j);
      const BlockDesc B = a.blocks[b];
      ok = closure_features(f, B, a.arena + B.base, a.pm, a.cps[B.case_id], i, j);
    }
    const bool finite = ok;
    float xs[CL_NF];
#pragma unroll
    for (int k = 0; k < CL_NF; k++) {
      const float x = (f[k] - N.mean[k]) * N.inv_std[k];
      ok = ok && x >= N.lo[k] && x <= N.hi[k];
      xs[k] = finite ? fminf(fmaxf(x, N.lo[k]), N.hi[k]) : 0.f;  // stand-in: clamped into the range
    }
#pragma unroll
    for (int k = 0; k < 16; k++) {
      float x = 0.f;
      if (k < CL_NF) x = a.standin ? xs[k] : (ok ? xs[k] : 0.f);
      sX[wave][lane * 16 + k] = f2bf(x);
    }
    sok[wave][lane] = ok ? 1 : 0;
  }
  __builtin_amdgcn_wave_barrier();
  Fwd F;
  closure_forward(F, N, sX[wave], s1[wave], s2[wave], s3[wave], lane);
  const int r = lane & 15, q = lane >> 4;
  if (r == 0) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int m = 4 * q + i, c = c0 + m;
      if (c >= a.ncell) continue;
      int b, ii, jj;
      unpack_cell(a.cells[c], b, ii, jj);
      const BlockDesc B = a.blocks[b];
      if (case_frozen(B.case_id)) continue;
      const CaseParams cp = a.cps[B.case_id];
      double* base = a.arena + B.base;
      const int64_t qq = pidx(B, ii, jj);
      auto PL = [&](int slot) { return base + (int64_t)slot * B.plane; };
      const double d = PL(a.pm.U)[qq], u = PL(a.pm.U + 1)[qq] / d, v = PL(a.pm.U + 2)[qq] / d;
      const double T = (PL(a.pm.U + 3)[qq] - 0.5 * d * (u * u + v * v)) / (cp.Cv * d);
      const double mu = 1.0 / cp.Re * (1.0 + cp.Tsb) * sqrt(T * T * T) / (cp.Tsb + T);
      const bool ok = sok[wave][m] != 0;
      const double mt = (ok || a.standin) ? mu * (double)softplus(F.Z4[i]) : 0.0;
      double* out = PL(a.pm.clo) + qq;
      *out = (a.standin && a.relax < 1.0) ? *out + a.relax * (mt - *out) : mt;
      PL(a.pm.clo + 1)[qq] = ok ? 1.0 : 0.0;
    }
  }
}

// Training: forward + backward over a mini-batch of samples (rows of
// data[idx[k]]), 64 per workgroup; each workgroup writes its gradient sum
// (CL_NPARAM floats) and its loss sum to part[blockIdx.x].
// Loss per sample: (log(1 + yhat) - log(1 + t))^2 / nbatch.
struct TrainArgs {
  const ClosureNet* net;
  const float* data;   // 8 floats per sample
  const int* idx;      // samples of this mini-batch
  int nbatch;
  float* part;         // [gridDim.x][CL_NPARAM + 1]
};

__global__ __launch_bounds__(256) void k_closure_train(TrainArgs a) {
  __shared__ ClosureNet N;
  __shared__ unsigned short sX[4][16 * 16], s1[4][16 * 64], s2[4][16 * 64], s3[4][16 * 64], sd[4][16 * 64];
  __shared__ float sdel4[4][16];
  __shared__ float sred[1][CL_NPARAM + 1];
  load_net(N, a.net);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
  const int c0 = blockIdx.x * 64 + wave * 16;
  const float inv_n = 1.f / (float)a.nbatch;
  // gradient accumulators of this wave (D layouts): dW3, dW2 as [k-tile][n-tile], dW1 [n-tile] (k < 16)
  f32x4 g3[4][4], g2[4][4], g1[4];
  const f32x4 z = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int x = 0; x < 4; x++) {
    g1[x] = z;
#pragma unroll
    for (int y = 0; y < 4; y++) g3[x][y] = g2[x][y] = z;
  }
  // samples of this wave
  float t_row = 0.f, w_row = 0.f;
  if (lane < 16) {
    const int c = c0 + lane;
    float f[CL_NF] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (c < a.nbatch) {
      const float* row = a.data + (int64_t)a.idx[c] * 8;
#pragma unroll
      for (int k = 0; k < CL_NF; k++) f[k] = (row[k] - N.mean[k]) * N.inv_std[k];
      t_row = row[6];
      w_row = row[7];
    }
#pragma unroll
    for (int k = 0; k < 16; k++) sX[wave][lane * 16 + k] = f2bf(k < CL_NF && w_row != 0.f ? f[k] : 0.f);
  }
  __builtin_amdgcn_wave_barrier();
  Fwd F;
  closure_forward(F, N, sX[wave], s1[wave], s2[wave], s3[wave], lane);
  // output gradient: lanes r == 0 hold z4 of rows 4q+i; targets sit in lanes 0..15
  float loss = 0.f;
  {
    float del[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int m = 4 * q + i;
      const float t = __shfl(t_row, m), w = __shfl(w_row, m);
      const float zz = F.Z4[i];
      const float y = softplus(zz);
      const float e = log1pf(y) - log1pf(fmaxf(t, 0.f));
      const float sg = 1.f / (1.f + expf(-zz));
      del[i] = w * 2.f * e / (1.f + y) * sg * inv_n;
      if (r == 0) loss += w * e * e * inv_n;
    }
    if (r == 0)
#pragma unroll
      for (int i = 0; i < 4; i++) sdel4[wave][4 * q + i] = del[i];
  }
  __builtin_amdgcn_wave_barrier();
  // delta3 = delta4 * w4 * (1 - H3^2)  (D layout), to LDS as bf16
  f32x4 D3[4];
#pragma unroll
  for (int nb = 0; nb < 4; nb++)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const float h = F.H3[nb][i];
      D3[nb][i] = sdel4[wave][4 * q + i] * bf2f(N.w4t[nb * 16 + r]) * (1.f - h * h);
    }
  store_act(sd[wave], D3, lane);
  __builtin_amdgcn_wave_barrier();
  // gradient of a layer: g[kt][nt] += H^T delta over the 16 rows
  auto wgrad = [&](f32x4 (&G)[4][4], const unsigned short* H, const unsigned short* Dl) {
#pragma unroll
    for (int kt = 0; kt < 4; kt++) {
      bf16x4 aa;
#pragma unroll
      for (int i = 0; i < 4; i++) aa[i] = (short)H[(4 * q + i) * 64 + kt * 16 + r];
#pragma unroll
      for (int nt = 0; nt < 4; nt++) {
        bf16x4 bb;
#pragma unroll
        for (int i = 0; i < 4; i++) bb[i] = (short)Dl[(4 * q + i) * 64 + nt * 16 + r];
        G[kt][nt] = mfma16(aa, bb, G[kt][nt]);
      }
    }
  };
  wgrad(g3, s2[wave], sd[wave]);
  // delta2 = (delta3 W3^T) * (1 - H2^2): B[n][k'] = W3[k'][n] = w3t[n][k']
  f32x4 D2[4] = {z, z, z, z};
  mm_16x64<64>(D2, sd[wave], 64, [&](int kk, int n) { return N.w3t[kk * 64 + n]; }, lane);
#pragma unroll
  for (int nb = 0; nb < 4; nb++)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const float h = F.H2[nb][i];
      D2[nb][i] = D2[nb][i] * (1.f - h * h);
    }
  // bias gradients need the fp32 deltas: reduce over rows into sred later
  float gb3[4], gb2[4], gb1[4], gw4[4];
#pragma unroll
  for (int nb = 0; nb < 4; nb++) {
    float s3v = 0.f, s2v = 0.f, w4v = 0.f;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      s3v += D3[nb][i];
      s2v += D2[nb][i];
      w4v += F.H3[nb][i] * sdel4[wave][4 * q + i];
    }
    gb3[nb] = s3v;
    gb2[nb] = s2v;
    gw4[nb] = w4v;
  }
  __builtin_amdgcn_wave_barrier();
  store_act(sd[wave], D2, lane);
  __builtin_amdgcn_wave_barrier();
  wgrad(g2, s1[wave], sd[wave]);
  // delta1 = (delta2 W2^T) * (1 - H1^2)
  f32x4 D1[4] = {z, z, z, z};
  mm_16x64<64>(D1, sd[wave], 64, [&](int kk, int n) { return N.w2t[kk * 64 + n]; }, lane);
#pragma unroll
  for (int nb = 0; nb < 4; nb++) {
    float s1v = 0.f;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const float h = F.H1[nb][i];
      D1[nb][i] = D1[nb][i] * (1.f - h * h);
      s1v += D1[nb][i];
    }
    gb1[nb] = s1v;
  }
  __builtin_amdgcn_wave_barrier();
  store_act(sd[wave], D1, lane);
  __builtin_amdgcn_wave_barrier();
  // dW1[k][n] = X^T delta1 (k < 16)
  {
    bf16x4 aa;
#pragma unroll
    for (int i = 0; i < 4; i++) aa[i] = (short)sX[wave][(4 * q + i) * 16 + r];
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
      bf16x4 bb;
#pragma unroll
      for (int i = 0; i < 4; i++) bb[i] = (short)sd[wave][(4 * q + i) * 64 + nt * 16 + r];
      g1[nt] = mfma16(aa, bb, g1[nt]);
    }
  }
  // ---- reduce the four waves in order into sred, then one row to part ----
  for (int w = 0; w < 4; w++) {
    if (wave == w) {
      auto put = [&](int off, float v) {
        if (w == 0) sred[0][off] = v;
        else sred[0][off] = sred[0][off] + v;
      };
#pragma unroll
      for (int kt = 0; kt < 4; kt++)
#pragma unroll
        for (int nt = 0; nt < 4; nt++)
#pragma unroll
          for (int i = 0; i < 4; i++) {
            const int k = kt * 16 + 4 * q + i, n = nt * 16 + r;
            put(CL_OW3 + k * 64 + n, g3[kt][nt][i]);
            put(CL_OW2 + k * 64 + n, g2[kt][nt][i]);
          }
#pragma unroll
      for (int nt = 0; nt < 4; nt++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int k = 4 * q + i, n = nt * 16 + r;
          if (k < CL_NF) put(CL_OW1 + k * 64 + n, g1[nt][i]);
        }
      // bias and output-layer sums: fold the four row groups q in order
      for (int qq = 0; qq < 4; qq++) {
        if (q == qq) {
#pragma unroll
          for (int nb = 0; nb < 4; nb++) {
            const int n = nb * 16 + r;
            const bool first = (w == 0 && qq == 0);
            sred[0][CL_OB3 + n] = first ? gb3[nb] : sred[0][CL_OB3 + n] + gb3[nb];
            sred[0][CL_OB2 + n] = first ? gb2[nb] : sred[0][CL_OB2 + n] + gb2[nb];
            sred[0][CL_OB1 + n] = first ? gb1[nb] : sred[0][CL_OB1 + n] + gb1[nb];
            sred[0][CL_OW4 + n] = first ? gw4[nb] : sred[0][CL_OW4 + n] + gw4[nb];
          }
        }
        __builtin_amdgcn_wave_barrier();
      }
      if (lane == 0) {
        float b4 = 0.f;
        for (int m = 0; m < 16; m++) b4 += sdel4[wave][m];
        sred[0][CL_OB4] = (w == 0) ? b4 : sred[0][CL_OB4] + b4;
      }
      // loss: lanes 0, 16, 32, 48
      const float L = __shfl(loss, 0) + __shfl(loss, 16) + __shfl(loss, 32) + __shfl(loss, 48);
      if (lane == 0) sred[0][CL_NPARAM] = (w == 0) ? L : sred[0][CL_NPARAM] + L;
    }
    __syncthreads();
  }
  float* out = a.part + (int64_t)blockIdx.x * (CL_NPARAM + 1);
  for (int k = threadIdx.x; k < CL_NPARAM + 1; k += 256) out[k] = sred[0][k];
}

// The workgroup partials of one mini-batch summed in order (the gradient
// k_closure_adam steps with), into out[CL_NPARAM + 1] (the loss last).
__global__ __launch_bounds__(256) void k_closure_gsum(const float* part, int nparts, float* out) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k > CL_NPARAM) return;
  float g = 0.f;
  for (int p = 0; p < nparts; p++) g += part[(int64_t)p * (CL_NPARAM + 1) + k];
  out[k] = g;
}

// Sum the workgroup partials in order and take one Adam step on the FP32
// master parameters; refresh the bf16 network.
struct AdamArgs {
  const float* part;
  int nparts;
  float* theta;   // master parameters
  float* m1;
  float* m2;
  int step;       // 1-based
  float lr, beta1, beta2, eps;
  ClosureNet* net;
  float* loss;    // out: loss of the mini-batch
};

__global__ __launch_bounds__(256) void k_closure_adam(AdamArgs a) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k > CL_NPARAM) return;
  float g = 0.f;
  for (int p = 0; p < a.nparts; p++) g += a.part[(int64_t)p * (CL_NPARAM + 1) + k];
  if (k == CL_NPARAM) {
    *a.loss = g;
    return;
  }
  const float m = a.beta1 * a.m1[k] + (1.f - a.beta1) * g;
  const float v = a.beta2 * a.m2[k] + (1.f - a.beta2) * g * g;
  a.m1[k] = m;
  a.m2[k] = v;
  const float mh = m / (1.f - powf(a.beta1, (float)a.step));
  const float vh = v / (1.f - powf(a.beta2, (float)a.step));
  const float th = a.theta[k] - a.lr * mh / (sqrtf(vh) + a.eps);
  a.theta[k] = th;
  // bf16 / fp32 copy for the kernels
  ClosureNet& N = *a.net;
  if (k < CL_OB1) {
    const int kk = (k - CL_OW1) / 64, n = (k - CL_OW1) % 64;
    N.w1t[n * 16 + kk] = f2bf(th);
  } else if (k < CL_OW2) {
    N.b1[k - CL_OB1] = th;
  } else if (k < CL_OB2) {
    const int kk = (k - CL_OW2) / 64, n = (k - CL_OW2) % 64;
    N.w2t[n * 64 + kk] = f2bf(th);
  } else if (k < CL_OW3) {
    N.b2[k - CL_OB2] = th;
  } else if (k < CL_OB3) {
    const int kk = (k - CL_OW3) / 64, n = (k - CL_OW3) % 64;
    N.w3t[n * 64 + kk] = f2bf(th);
  } else if (k < CL_OW4) {
    N.b3[k - CL_OB3] = th;
  } else if (k < CL_OB4) {
    N.w4t[k - CL_OW4] = f2bf(th);  // column n = 0: w4t[0][k]
  } else {
    N.b4 = th;
  }
}

}  // namespace ec2d
