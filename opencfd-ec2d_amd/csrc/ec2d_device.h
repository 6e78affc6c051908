// Data structures shared by the host runtime and the HIP kernels.
//
// Device layout (one arena per mesh level of a batch):
//   every block owns a set of planes; a plane is a (ny+3) x P array of FP64,
//   row r = j+1, column c = i+1, for the reference's ghosted index range
//   i in [-1, nx+1], j in [-1, ny+1] (LAP = 2 ghost layers).  P (the row
//   pitch) is nx+3 rounded up to 16 doubles (128 B), so each row starts on a
//   cache line.  All per-cell and per-face arrays of the reference (U, Res,
//   dU, dt, Lci..Lvj, vol, si, sj, ni*, nj*, x1, y1, Amu, Amu_t, dw, Un, QF,
//   deltU and the finite-difference metrics) live in planes of this shape, so
//   one index expression serves every array of a block.
#pragma once
#include <stdint.h>

#define EC2D_LAP 2
#define EC2D_MAXSUB 64

// plane slots (order inside a block's plane set)
enum {
  PL_U = 0,          // NV planes
};

struct CaseParams {
  double gamma, Ma, Cv, Cp, Re, Pr, Tsb, CFL, dtmax, dtmin, dt_global, P_outlet;
  double u_inf, v_inf, p_inf, Twall, vt_inf, Kt_inf, Wt_inf, p00, inv_Re;
  int If_viscous, Iflag_local_dt, Time_Method, pad;
};

// Per-mesh-level numerical options; identical for all blocks in one launch.
struct MeshOpts {
  int nvar, scheme, flux, recon, turb, global_scheme, nvar_un, pad;
};

struct BlockDesc {
  int nx, ny, P, case_id;
  int64_t base;        // first double of this block's planes in the level arena
  int64_t plane;       // doubles per plane = (ny+3)*P
  int fdm;             // 1: finite-difference interior (LFDM = 4 edge layers stay FVM)
  int bc_off, nsub;    // subfaces in the level's subface table
  int sk_S, sk_R;      // LU-SGS records: steps (ncol + nrow - 1), rows (whole sweep waves; the fused launch: nrow)
  int64_t sk_base;     // first double of the block's LU-SGS records in the level's record arena
  int64_t sk_n;        // stride of a record field in doubles: sk_S * sk_R positions + 64 trash slots
};

// plane indices within a block (computed per level from nvar)
struct PlaneMap {
  int U, Res, dU, dt, Lci, Lcj, Lvi, Lvj, vol, si, sj, ni1, ni2, nj1, nj2, x1, y1, Amu, Amu_t, dw,
      Un, QF, deltU, Akx, Aky, Aix, Aiy, Ajac, flag, fx, fy, ev, sst, clo,
      Un1,  // U(n-1) of dual-time stepping (fine level, Time_Method -1), else -1
      bl,   // Baldwin-Lomax scratch (fine level, model 1): vorticity, first-write flag; else -1
      nplanes;
};

// One subface of bc2d_1.in, resolved for the device.
struct SubfaceDesc {
  int face, ist, iend, jst, jend, neighb, nsub_nb, orient;  // as in the file
  int nb_block;        // neighbour block index in the level (0-based), or -1
  int nb_face, nb_ist, nb_iend, nb_jst, nb_jend;
  int norm_off;        // offset of per-point outward normals (farfield / symmetry), -1 if none
  int sync;            // ghost fill: 1 = a barrier before this subface (it depends on the ones since the last)
};

struct TileDesc {
  int block, i0, j0, pad;  // first interior cell (1-based) of the tile
};

// Per-case run mask of a run to a residual drop (ec2d_run_to_drop): null
// outside such a run; case c is frozen (every kernel leaves its blocks
// untouched) while g_case_active[c] == 0.  Kernels test it once per
// workgroup, before any barrier: a workgroup only ever holds cells of one
// block, hence of one case.
__device__ const int* g_case_active = nullptr;

__device__ inline bool case_frozen(int case_id) {
  const int* p = g_case_active;
  return p != nullptr && p[case_id] == 0;
}

__host__ __device__ inline int64_t pidx(const BlockDesc& b, int i, int j) {
  return (int64_t)(j + 1) * b.P + (i + 1);
}
