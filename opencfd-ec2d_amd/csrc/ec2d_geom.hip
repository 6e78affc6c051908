// Mesh set-up on the device: ghost nodes, cell centres and face metrics of
// every block of a mesh level from its interior nodes, as the reference
// computes them once at start-up:
//   ghost nodes        Update_coordinate_buffer_onemesh (sub_boundary.f90:
//                      170-281): per subface in file order, an interface
//                      copies the neighbour's first LAP interior node lines
//                      (orientation as in the file), any other subface
//                      extrapolates 2 x(boundary) - x(interior); then the
//                      LAP x LAP corner nodes x(i1,j1) = x(i1,1) + x(1,j1) -
//                      x(1,1) (and the three other corners)
//   cell centres       (sub_boundary.f90:275-280) over 1-LAP .. n+LAP-1
//   volume, face areas and normals  Creat_Mesh (sub_init.f90:180-204)
// One workgroup per block; subfaces one after the other (a later subface
// overwrites the end points an earlier one wrote, as in the reference).
// Every expression keeps the reference's operation order, so the results are
// those of the host restatement bit for bit.
#include "ec2d_device.h"

namespace ec2d {

struct GeomArgs {
  double* x;             // nodes of all blocks, (-1:nx+2,-1:ny+2) each, column-major
  double* y;
  const int64_t* noff;   // first node of block b
  const int* dims;       // nx, ny of block b
  const int* bc;         // 9 ints per subface (bc2d_1.in), blocks one after the other
  const int* sub0;       // first subface of block b, b + 1 (block count + 1 entries)
  const int* nbblk;      // per subface: the neighbour's block index in this launch, or -1
  double* cen;           // x1, y1 of block b at coff[b]: 2 x (nx+3)(ny+3)
  const int64_t* coff;
  double* met;           // vol, si, sj, ni1, ni2, nj1, nj2 of block b at moff[b]: 7 x nx ny
  const int64_t* moff;
};

__global__ __launch_bounds__(256) void k_geometry(GeomArgs a) {
  const int b = blockIdx.x;
  const int nx = a.dims[2 * b], ny = a.dims[2 * b + 1];
  double* X = a.x + a.noff[b];
  double* Y = a.y + a.noff[b];
  auto N = [&](int i, int j) { return (int64_t)(i + 1) + (int64_t)(nx + 4) * (j + 1); };  // (-1:nx+2, -1:ny+2)
  constexpr int LAP = EC2D_LAP;
  for (int k = a.sub0[b]; k < a.sub0[b + 1]; k++) {
    const int* s = a.bc + 9 * k;
    const int face = s[1], ib2 = s[2], ie2 = s[3], jb2 = s[4], je2 = s[5], orient = s[8];
    if (s[6] > 0) {
      // interface: the neighbour's node lines next to its subface
      const int nb = a.nbblk[k];
      const int nx1 = a.dims[2 * nb];
      const double* X1 = a.x + a.noff[nb];
      const double* Y1 = a.y + a.noff[nb];
      auto N1 = [&](int i, int j) { return (int64_t)(i + 1) + (int64_t)(nx1 + 4) * (j + 1); };
      const int* s1 = a.bc + 9 * (a.sub0[nb] + s[7] - 1);
      const int face1 = s1[1], ib1 = s1[2], ie1 = s1[3], jb1 = s1[4], je1 = s1[5];
      const int n1 = ie2 - ib2 + je2 - jb2 + 1;
      for (int idx = threadIdx.x; idx < n1 * LAP; idx += blockDim.x) {
        const int kk = idx % n1 + 1, i = idx / n1 + 1;  // point k of the subface, line i
        const int k1 = orient == 2 ? kk : n1 + 1 - kk;
        int64_t src;
        if (face1 == 1) src = N1(ib1 + i, jb1 + k1 - 1);
        else if (face1 == 2) src = N1(ib1 + k1 - 1, jb1 + i);
        else if (face1 == 3) src = N1(ie1 - LAP - 1 + i, jb1 + k1 - 1);
        else src = N1(ib1 + k1 - 1, je1 - LAP - 1 + i);
        int64_t dst;
        if (face == 1) dst = N(ib2 - LAP - 1 + i, jb2 + kk - 1);
        else if (face == 2) dst = N(ib2 + kk - 1, jb2 - LAP - 1 + i);
        else if (face == 3) dst = N(ie2 + i, jb2 + kk - 1);
        else dst = N(ib2 + kk - 1, je2 + i);
        X[dst] = X1[src];
        Y[dst] = Y1[src];
      }
    } else {
      // a physical boundary: mirror the interior node about the boundary node
      const bool alj = face == 1 || face == 3;  // points along j
      const int n1 = alj ? je2 - jb2 + 1 : ie2 - ib2 + 1;
      for (int idx = threadIdx.x; idx < n1 * LAP; idx += blockDim.x) {
        const int p = idx % n1, i = idx / n1 + 1;
        int64_t g, bnd, in;
        if (face == 1) g = N(ib2 - i, jb2 + p), bnd = N(ib2, jb2 + p), in = N(ib2 + i, jb2 + p);
        else if (face == 2) g = N(ib2 + p, jb2 - i), bnd = N(ib2 + p, jb2), in = N(ib2 + p, jb2 + i);
        else if (face == 3) g = N(ie2 + i, jb2 + p), bnd = N(ie2, jb2 + p), in = N(ie2 - i, jb2 + p);
        else g = N(ib2 + p, je2 + i), bnd = N(ib2 + p, je2), in = N(ib2 + p, je2 - i);
        X[g] = 2.0 * X[bnd] - X[in];
        Y[g] = 2.0 * Y[bnd] - Y[in];
      }
    }
    __syncthreads();
  }
  // corner nodes (LAP x LAP at each corner)
  for (int idx = threadIdx.x; idx < 4 * LAP * LAP * 2; idx += blockDim.x) {
    const int q = idx % (LAP * LAP), corner = (idx / (LAP * LAP)) % 4, xy = idx / (4 * LAP * LAP);
    double* A = xy ? Y : X;
    const int i = q % LAP + 1, j = q / LAP + 1;
    const int i1 = 1 - i, j1 = 1 - j, i2 = nx + i, j2 = ny + j;
    if (corner == 0) A[N(i1, j1)] = A[N(i1, 1)] + A[N(1, j1)] - A[N(1, 1)];
    else if (corner == 1) A[N(i2, j2)] = A[N(i2, ny)] + A[N(nx, j2)] - A[N(nx, ny)];
    else if (corner == 2) A[N(i1, j2)] = A[N(i1, ny)] + A[N(1, j2)] - A[N(1, ny)];
    else A[N(i2, j1)] = A[N(i2, 1)] + A[N(nx, j1)] - A[N(nx, 1)];
  }
  __syncthreads();
  // cell centres over -1..nx+1 x -1..ny+1
  double* x1 = a.cen + a.coff[b];
  double* y1 = x1 + (int64_t)(nx + 3) * (ny + 3);
  for (int idx = threadIdx.x; idx < (nx + 3) * (ny + 3); idx += blockDim.x) {
    const int i = idx % (nx + 3) - 1, j = idx / (nx + 3) - 1;
    x1[idx] = (X[N(i, j)] + X[N(i + 1, j)] + X[N(i, j + 1)] + X[N(i + 1, j + 1)]) * 0.25;
    y1[idx] = (Y[N(i, j)] + Y[N(i + 1, j)] + Y[N(i, j + 1)] + Y[N(i + 1, j + 1)]) * 0.25;
  }
  // volume, face areas and normals over 1..nx x 1..ny (zero where undefined)
  const int64_t np = (int64_t)nx * ny;
  double* vol = a.met + a.moff[b];
  double *si = vol + np, *sj = vol + 2 * np, *ni1 = vol + 3 * np, *ni2 = vol + 4 * np, *nj1 = vol + 5 * np,
         *nj2 = vol + 6 * np;
  for (int idx = threadIdx.x; idx < nx * ny; idx += blockDim.x) {
    const int i = idx % nx + 1, j = idx / nx + 1;
    double v = 0.0, s_i = 0.0, a1 = 0.0, a2 = 0.0, s_j = 0.0, b1 = 0.0, b2 = 0.0;
    if (i <= nx - 1 && j <= ny - 1)
      v = fabs((X[N(i, j)] - X[N(i + 1, j + 1)]) * (Y[N(i + 1, j)] - Y[N(i, j + 1)]) -
               (X[N(i + 1, j)] - X[N(i, j + 1)]) * (Y[N(i, j)] - Y[N(i + 1, j + 1)])) *
          0.5;
    if (j <= ny - 1) {
      const double dx = X[N(i, j + 1)] - X[N(i, j)], dy = Y[N(i, j + 1)] - Y[N(i, j)];
      s_i = sqrt(dx * dx + dy * dy);
      a1 = dy / s_i;
      a2 = -dx / s_i;
    }
    if (i <= nx - 1) {
      const double dx = X[N(i + 1, j)] - X[N(i, j)], dy = Y[N(i + 1, j)] - Y[N(i, j)];
      s_j = sqrt(dx * dx + dy * dy);
      b1 = -dy / s_j;
      b2 = dx / s_j;
    }
    vol[idx] = v;
    si[idx] = s_i;
    sj[idx] = s_j;
    ni1[idx] = a1;
    ni2[idx] = a2;
    nj1[idx] = b1;
    nj2[idx] = b2;
  }
}

}  // namespace ec2d
