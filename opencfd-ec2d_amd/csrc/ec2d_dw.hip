// Wall distance on the device: comput_dw (sub_turbulence_SST.f90:286-337)
// for blocks whose caller gave no dw (SA, SST and the learned closure's model
// 4 need it).  Every node (1..nx, 1..ny) of a block takes the smallest
// distance to the wall nodes of its case, sqrt(dx**2 + dy**2) with strict
// less-than updates (the minimum does not depend on their order); each cell
// then takes the mean of its four nodes, ((n(i,j) + n(i+1,j)) + n(i,j+1) +
// n(i+1,j+1)) * 0.25 as the reference sums it.
#include "ec2d_device.h"

namespace ec2d {

struct DwArgs {
  const double* nx_y;     // node coordinates of the blocks: x then y, nx*ny each, per block at off[b]
  const int64_t* off;     // per block: first node in nx_y
  const int* blk;         // per block: its BlockDesc index in the level
  const int* wall0;       // per block: first wall node of its case in wx / wy, and one past the last
  const double* wx;
  const double* wy;
  const BlockDesc* blocks;
  double* arena;
  int dw;                 // plane of dw
  double* node_dw;        // scratch, one per node
};

__global__ __launch_bounds__(256) void k_dw_nodes(DwArgs a) {
  const int b = blockIdx.y;
  const BlockDesc B = a.blocks[a.blk[b]];
  const int nn = B.nx * B.ny;
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= nn) return;
  const double x = a.nx_y[a.off[b] + p], y = a.nx_y[a.off[b] + nn + p];
  double d = 1.0e10;
  for (int k = a.wall0[2 * b]; k < a.wall0[2 * b + 1]; k++) {
    const double dx = x - a.wx[k], dy = y - a.wy[k];
    const double d1 = sqrt(dx * dx + dy * dy);
    if (d1 < d) d = d1;
  }
  a.node_dw[a.off[b] / 2 + p] = d;
}

__global__ __launch_bounds__(256) void k_dw_cells(DwArgs a) {
  const int b = blockIdx.y;
  const BlockDesc B = a.blocks[a.blk[b]];
  const int nx = B.nx, ny = B.ny;
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= (nx - 1) * (ny - 1)) return;
  const int i = 1 + c % (nx - 1), j = 1 + c / (nx - 1);
  const double* n = a.node_dw + a.off[b] / 2;
  auto nd = [&](int ii, int jj) { return n[(jj - 1) * nx + (ii - 1)]; };
  double* dw = a.arena + B.base + (int64_t)a.dw * B.plane;
  dw[pidx(B, i, j)] = (nd(i, j) + nd(i + 1, j) + nd(i, j + 1) + nd(i + 1, j + 1)) * 0.25;
}

}  // namespace ec2d
