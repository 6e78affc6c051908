// Point-wise kernels of the time-advance and multigrid drivers, one launch
// per operation over all blocks of a mesh level (one workgroup per row of a
// block; rows come from a per-level row table).
//   update      U = U + dU                     sub_NS_singlegrid.f90:203-211
//   clip        nu_t, k, omega < 0 -> 1e-10    sub_NS_singlegrid.f90:30-60
//   euler1      U = U + dt*Res/vol             sub_NS_singlegrid.f90:84-93
//   copy_un     Un = U (0:nx, 0:ny)            sub_NS_singlegrid.f90:131-139
//   rk3         U = a Un + c U + dt b Res/vol  sub_NS_singlegrid.f90:124-163
//   force       QF = QF - Res                  sub_NS_multigrid.f90:156-173
//   add_force   Res = Res + QF                 sub_NS_multigrid.f90:177-195
//   delt_u      deltU = U - Un (0:nx, 0:ny)    sub_NS_multigrid.f90:199-216
//   set_un      Un = U (0:nx, 0:ny)            sub_NS_multigrid.f90:219-237
//   new_u       U = U + deltU                  sub_NS_multigrid.f90:240-260
//   restrict    fine -> coarse (U vol-weighted, Res summed into QF)
//                                              sub_NS_multigrid.f90:349-398
//   prolong     coarse -> fine (9,3,3,1)/16 of U or deltU
//                                              sub_NS_multigrid.f90:271-343
#include "ec2d_device.h"

namespace ec2d {

enum PwOp { PW_UPDATE, PW_CLIP, PW_EULER1, PW_COPY_UN, PW_RK3, PW_FORCE, PW_ADD_FORCE, PW_DELT_U, PW_SET_UN, PW_NEW_U,
            PW_DUAL_SRC, PW_DUAL_INIT, PW_DUAL_SHIFT };

struct PwArgs {
  const BlockDesc* blocks;
  const int2* rows;  // (block, j) per workgroup
  double* arena;
  PlaneMap pm;
  int nvar, op, krk;
  double ra, rb, rg;  // RK3 stage coefficients
  double sfac;        // dual-time source factor of this residual call
};

__global__ __launch_bounds__(256) void k_pointwise(PwArgs a) {
  const int2 rj = a.rows[blockIdx.x];
  const BlockDesc B = a.blocks[rj.x];
  if (case_frozen(B.case_id)) return;
  const int j = rj.y;
  double* base = a.arena + B.base;
  auto PL = [&](int slot) { return base + (int64_t)slot * B.plane; };
  const PlaneMap pm = a.pm;
  const int nx = B.nx, ny = B.ny;
  const bool interior_row = j >= 1 && j <= ny - 1;
  switch (a.op) {
    case PW_UPDATE:
      if (!interior_row) return;
      for (int i = 1 + threadIdx.x; i <= nx - 1; i += blockDim.x) {
        int64_t q = pidx(B, i, j);
        for (int m = 0; m < a.nvar; m++) PL(pm.U + m)[q] = PL(pm.U + m)[q] + PL(pm.dU + m)[q];
      }
      break;
    case PW_CLIP:
      if (!interior_row) return;
      for (int i = 1 + threadIdx.x; i <= nx - 1; i += blockDim.x) {
        int64_t q = pidx(B, i, j);
        for (int m = 4; m < a.nvar; m++)
          if (PL(pm.U + m)[q] < 0) PL(pm.U + m)[q] = 1.0e-10;
      }
      break;
    case PW_EULER1:
      if (!interior_row) return;
      for (int i = 1 + threadIdx.x; i <= nx - 1; i += blockDim.x) {
        int64_t q = pidx(B, i, j);
        for (int m = 0; m < a.nvar; m++) {
          double du = PL(pm.Res + m)[q] / PL(pm.vol)[q];
          PL(pm.U + m)[q] = PL(pm.U + m)[q] + PL(pm.dt)[q] * du;
        }
      }
      break;
    case PW_RK3:
      if (!interior_row) return;
      for (int i = 1 + threadIdx.x; i <= nx - 1; i += blockDim.x) {
        int64_t q = pidx(B, i, j);
        for (int m = 0; m < a.nvar; m++) {
          double du = PL(pm.Res + m)[q] / PL(pm.vol)[q];
          PL(pm.U + m)[q] = a.ra * PL(pm.Un + m)[q] + a.rg * PL(pm.U + m)[q] + PL(pm.dt)[q] * a.rb * du;
        }
      }
      break;
    case PW_COPY_UN:
    case PW_SET_UN: {
      if (j < 0 || j > ny) return;
      int nv = (a.op == PW_SET_UN) ? 4 : a.nvar;
      for (int i = threadIdx.x; i <= nx; i += blockDim.x) {
        int64_t q = pidx(B, i, j);
        for (int m = 0; m < nv; m++) PL(pm.Un + m)[q] = PL(pm.U + m)[q];
      }
      break;
    }
    case PW_DELT_U:
      if (j < 0 || j > ny) return;
      for (int i = threadIdx.x; i <= nx; i += blockDim.x) {
        int64_t q = pidx(B, i, j);
        for (int m = 0; m < 4; m++) PL(pm.deltU + m)[q] = PL(pm.U + m)[q] - PL(pm.Un + m)[q];
      }
      break;
    case PW_FORCE:
      if (!interior_row) return;
      for (int i = 1 + threadIdx.x; i <= nx - 1; i += blockDim.x) {
        int64_t q = pidx(B, i, j);
        for (int m = 0; m < 4; m++) PL(pm.QF + m)[q] = PL(pm.QF + m)[q] - PL(pm.Res + m)[q];
      }
      break;
    case PW_ADD_FORCE:
      if (!interior_row) return;
      for (int i = 1 + threadIdx.x; i <= nx - 1; i += blockDim.x) {
        int64_t q = pidx(B, i, j);
        for (int m = 0; m < 4; m++) PL(pm.Res + m)[q] = PL(pm.Res + m)[q] + PL(pm.QF + m)[q];
      }
      break;
    case PW_DUAL_SRC:
      // dual-time source (sub_Residual.f90:55-64): Res - (3 U - 4 Un + Un1) vol Sfac,
      // evaluated left to right
      if (!interior_row) return;
      for (int i = 1 + threadIdx.x; i <= nx - 1; i += blockDim.x) {
        int64_t q = pidx(B, i, j);
        for (int m = 0; m < a.nvar; m++) {
          double t = (3.0 * PL(pm.U + m)[q] - 4.0 * PL(pm.Un + m)[q]) + PL(pm.Un1 + m)[q];
          PL(pm.Res + m)[q] = PL(pm.Res + m)[q] - (t * PL(pm.vol)[q]) * a.sfac;
        }
      }
      break;
    case PW_DUAL_INIT:  // Init_flow: Un = Un1 = U (sub_init.f90:396-409)
    case PW_DUAL_SHIFT:  // end of a physical step: Un1 = Un, Un = U (sub_NS_singlegrid.f90:282-299)
      if (!interior_row) return;
      for (int i = 1 + threadIdx.x; i <= nx - 1; i += blockDim.x) {
        int64_t q = pidx(B, i, j);
        for (int m = 0; m < a.nvar; m++) {
          PL(pm.Un1 + m)[q] = a.op == PW_DUAL_INIT ? PL(pm.U + m)[q] : PL(pm.Un + m)[q];
          PL(pm.Un + m)[q] = PL(pm.U + m)[q];
        }
      }
      break;
    case PW_NEW_U:
      if (!interior_row) return;
      for (int i = 1 + threadIdx.x; i <= nx - 1; i += blockDim.x) {
        int64_t q = pidx(B, i, j);
        for (int m = 0; m < 4; m++) PL(pm.U + m)[q] = PL(pm.U + m)[q] + PL(pm.deltU + m)[q];
      }
      break;
  }
}

struct XferArgs {
  const BlockDesc* fine_blocks;
  const BlockDesc* coarse_blocks;
  const int2* rows;  // rows of the level being written: (block, j)
  double* fine_arena;
  double* coarse_arena;
  PlaneMap fpm, cpm;
  int flag;          // restrict: 1 U, 2 Res -> QF, 3 both; prolong: 1 U, 2 deltU,
                     // 3 the coarse-grid correction (see k_prolong)
};

// fine -> coarse (rows are coarse rows j = 1..ny2-1)
__global__ __launch_bounds__(256) void k_restrict(XferArgs a) {
  const int2 rj = a.rows[blockIdx.x];
  const BlockDesc B2 = a.coarse_blocks[rj.x];
  const BlockDesc B1 = a.fine_blocks[rj.x];
  const int j = rj.y;
  if (j < 1 || j > B2.ny - 1 || case_frozen(B1.case_id)) return;
  double* c = a.coarse_arena + B2.base;
  double* f = a.fine_arena + B1.base;
  const int j1 = 2 * j - 1, j2 = 2 * j;
  for (int i = 1 + threadIdx.x; i <= B2.nx - 1; i += blockDim.x) {
    const int i1 = 2 * i - 1, i2 = 2 * i;
    int64_t q11 = pidx(B1, i1, j1), q12 = pidx(B1, i1, j2), q21 = pidx(B1, i2, j1), q22 = pidx(B1, i2, j2);
    int64_t qc = pidx(B2, i, j);
    if (a.flag & 1) {
      const double* vol1 = f + (int64_t)a.fpm.vol * B1.plane;
      double v11 = vol1[q11], v12 = vol1[q12], v21 = vol1[q21], v22 = vol1[q22];
      double vc = c[(int64_t)a.cpm.vol * B2.plane + qc];
      for (int m = 0; m < 4; m++) {
        const double* P1 = f + (int64_t)(a.fpm.U + m) * B1.plane;
        c[(int64_t)(a.cpm.U + m) * B2.plane + qc] = (P1[q11] * v11 + P1[q12] * v12 + P1[q21] * v21 + P1[q22] * v22) / vc;
      }
    }
    if (a.flag & 2) {
      for (int m = 0; m < 4; m++) {
        const double* P1 = f + (int64_t)(a.fpm.Res + m) * B1.plane;
        c[(int64_t)(a.cpm.QF + m) * B2.plane + qc] = P1[q11] + P1[q12] + P1[q21] + P1[q22];
      }
    }
  }
}

// coarse -> fine (rows are fine rows j = 0..ny2)
__global__ __launch_bounds__(256) void k_prolong(XferArgs a) {
  const int2 rj = a.rows[blockIdx.x];
  const BlockDesc B2 = a.fine_blocks[rj.x];
  const BlockDesc B1 = a.coarse_blocks[rj.x];
  const int j = rj.y;
  if (j < 0 || j > B2.ny || case_frozen(B2.case_id)) return;
  const double a1 = 9.0 / 16.0, a2 = 3.0 / 16.0, a3 = 1.0 / 16.0;
  const int ja1 = (j % 2 == 0) ? j / 2 : j / 2 + 1, ja2 = (j % 2 == 0) ? j / 2 + 1 : j / 2;
  double* c = a.coarse_arena + B1.base;
  double* f = a.fine_arena + B2.base;
  if (a.flag == 3) {
    // coarse-grid correction, the sequence comput_delt_U (coarse), prolong_U
    // (flag 2) and comput_new_U (fine) of sub_NS_multigrid.f90:214-260 in
    // one pass: the coarse deltU = U - Un of the four parents is formed where
    // it is read; the fine row stores the prolonged deltU (0..nx) and adds it
    // to U on interior cells.  Every coarse cell's deltU is also stored, once:
    // by the fine cell of even index, or of the last index for the last
    // coarse column / row.
    const bool jw = (j % 2 == 0) || j == B2.ny;
    const int jc = (j % 2 == 0) ? j / 2 : j / 2 + 1;
    for (int i = threadIdx.x; i <= B2.nx; i += blockDim.x) {
      const int ia1 = (i % 2 == 0) ? i / 2 : i / 2 + 1, ia2 = (i % 2 == 0) ? i / 2 + 1 : i / 2;
      int64_t q11 = pidx(B1, ia1, ja1), q21 = pidx(B1, ia2, ja1), q12 = pidx(B1, ia1, ja2), q22 = pidx(B1, ia2, ja2);
      int64_t qf = pidx(B2, i, j);
      const bool interior = i >= 1 && i <= B2.nx - 1 && j >= 1 && j <= B2.ny - 1;
      const bool cw = jw && ((i % 2 == 0) || i == B2.nx);
      const int64_t qc = pidx(B1, (i % 2 == 0) ? i / 2 : i / 2 + 1, jc);
      for (int m = 0; m < 4; m++) {
        const double* Uc = c + (int64_t)(a.cpm.U + m) * B1.plane;
        const double* Unc = c + (int64_t)(a.cpm.Un + m) * B1.plane;
        const double d11 = Uc[q11] - Unc[q11], d21 = Uc[q21] - Unc[q21];
        const double d12 = Uc[q12] - Unc[q12], d22 = Uc[q22] - Unc[q22];
        const double v = a1 * d11 + a2 * (d21 + d12) + a3 * d22;
        f[(int64_t)(a.fpm.deltU + m) * B2.plane + qf] = v;
        if (interior) {
          double* Uf = f + (int64_t)(a.fpm.U + m) * B2.plane;
          Uf[qf] = Uf[qf] + v;
        }
        if (cw) c[(int64_t)(a.cpm.deltU + m) * B1.plane + qc] = Uc[qc] - Unc[qc];
      }
    }
    return;
  }
  const int slot_c = (a.flag == 1) ? a.cpm.U : a.cpm.deltU;
  const int slot_f = (a.flag == 1) ? a.fpm.U : a.fpm.deltU;
  for (int i = threadIdx.x; i <= B2.nx; i += blockDim.x) {
    const int ia1 = (i % 2 == 0) ? i / 2 : i / 2 + 1, ia2 = (i % 2 == 0) ? i / 2 + 1 : i / 2;
    int64_t q11 = pidx(B1, ia1, ja1), q21 = pidx(B1, ia2, ja1), q12 = pidx(B1, ia1, ja2), q22 = pidx(B1, ia2, ja2);
    int64_t qf = pidx(B2, i, j);
    for (int m = 0; m < 4; m++) {
      const double* P1 = c + (int64_t)(slot_c + m) * B1.plane;
      f[(int64_t)(slot_f + m) * B2.plane + qf] = a1 * P1[q11] + a2 * (P1[q21] + P1[q12]) + a3 * P1[q22];
    }
  }
}

}  // namespace ec2d
