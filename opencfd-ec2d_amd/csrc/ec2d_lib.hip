// Runtime behind include/ec2d.h: batch bookkeeping, device arenas, launches,
// HIP graph capture of the main-loop body.  One translation unit so that the
// kernels (defined in the included files) can be launched from here.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstdlib>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ec2d.h"
#include "ec2d_device.h"
#include "ec2d_fdm.hip"
#include "ec2d_ghost.hip"
#include "ec2d_lusgs.hip"
#include "ec2d_sst.hip"
#include "ec2d_forces.hip"
#include "ec2d_closure.hip"
#include "ec2d_pointwise.hip"
#include "ec2d_residual.hip"
#include "ec2d_bl.hip"
#include "ec2d_dw.hip"
#include "ec2d_geom.hip"

using namespace ec2d;

namespace {

struct HBlock {
  int nx = 0, ny = 0, nsub = 0, fdm = 0;
  bool defined = false;
  bool geom_dev = false;  // ghost nodes, centres and metrics made on the device at commit
  std::vector<int> bc;
  std::vector<double> x, y, x1, y1, vol, si, sj, ni1, ni2, nj1, nj2, dw, fd;
};

struct HCase {
  ec2d_control ctl;
  CaseParams cp;
  int num_mesh = 1;
  std::vector<std::vector<HBlock>> levels;  // [level][block]
  int kstep[3] = {0, 0, 0};
  double tt[3] = {0, 0, 0};
};

struct DLevel {
  MeshOpts o{};
  PlaneMap pm{};
  double* arena = nullptr;
  int64_t arena_n = 0;
  std::vector<BlockDesc> hb;
  std::vector<int> blk_case, blk_local;
  BlockDesc* db = nullptr;
  SubfaceDesc* ds = nullptr;
  int* dunits = nullptr;  // ghost fill order per block: subface indices, boundary conditions first
  int bl_walls = 0, bl_lines = 0;  // Baldwin-Lomax: most wall subfaces of a block, most lines of one
  int64_t max_plane = 0;           // largest block plane (doubles)
  double* dnorm = nullptr;
  std::vector<TileDesc> ht;
  TileDesc* dt = nullptr;
  int* dperm = nullptr;
  int ntiles = 0, nlaunch = 0;
  int tile_rows = 4;          // residual tile: tile_cols x tile_rows cells
  int tile_cols = 64;         // 64, or 32 with 8 rows
  std::vector<int> case_tile0;
  int* dcase_tile0 = nullptr;
  double* dncell = nullptr;
  double* dpartials = nullptr;
  double* dnorms = nullptr;
  bool residual_done = false;  // a residual launch has written dpartials (reduced on demand)
  int2* drows = nullptr;
  int nrows = 0;
  int* dblocks_all = nullptr;
  std::vector<int> case_first_block;
  int lusgs_threads = 64;
  int lusgs_lpc = 1;          // LU-SGS lanes per cell: 2 when every block has at most 32 rows
  int lusgs_split = 1;        // LU-SGS sweep workgroups per block (2 or 4: a block's rows split between CUs)
  int lusgs_u = 2;            // LU-SGS sweep steps per register set (record prefetch depth), up to 4 waves
  bool lusgs_fused = false;   // LU-SGS records made inside the sweep workgroup (k_lusgs_fused)
  bool planar_stale = false;  // Res, dt, Lc, Lv are current only in the hyperplane copy (planar_refresh)
  bool skew_static = false;   // the record arena holds the skewed copy of the mesh planes
  int fused_np = 6;           // its producer waves (3 or 6)
  unsigned long long* dgring = nullptr;  // split sweeps: one ring per boundary between workgroups and direction
  int64_t gstride = 0;                   // slots of one ring
  int4* dwall = nullptr;      // wall faces (block, face, i, j), cases in order
  int* dcase_wall0 = nullptr;
  double* dforces = nullptr;  // 4 per case
  double* sk = nullptr;       // LU-SGS records (skewed, see ec2d_lusgs.hip)
  int64_t sk_max_n = 0;       // largest record field of a block (doubles)
  int64_t sk_total = 0;       // doubles of the level's record arena
  int max_cells = 0;          // largest block (cells)
  int prep_tiles = 0;         // LU-SGS prep workgroups per block: 32 x 8 cell tiles of the largest block
  int skew_tiles = 0;         // k_lusgs_skew workgroups per block: 16 x 16 tiles of the largest block's skewed space
  std::vector<long long> case_cells;
  bool fdm = false;
  int2 *d_ilines = nullptr, *d_frows = nullptr;  // d_ilines: i-lines then j-lines
  int n_ilines = 0, n_jlines = 0, n_frows = 0;
};

struct Runtime {
  bool inited = false, committed = false;
  std::vector<HCase> cases;
  std::vector<DLevel> lv;
  CaseParams* dcps = nullptr;
  unsigned long long* derr = nullptr;
  int* dsweep_err = nullptr;  // a split LU-SGS sweep ran out of its wait budget
  int wait_budget = SPIN_MAX;  // polls of a bounded LU-SGS wait (ec2d_set_wait_budget)
  long long sko_passes = 0;    // passes whose residual wrote the hyperplane copy alone (advance_lusgs)
  bool graph_sko = false;      // the captured loop body is such a pass
  hipStream_t s = nullptr;
  std::string err;
  int num_mesh = 1;
  int tile_rows = 4;
  int tile_cols = 64;
  bool clip_pending[3] = {false, false, false};
  // dual-time stepping: the reference's first-call flags (integer, save) of
  // Comput_Residual_one_mesh (Sfac, sub_Residual.f90:18,28-34) and of
  // Dual_time_LU_SGS (alfa1, sub_NS_singlegrid.f90:241-247), counted from
  // ec2d_init_flow; the last step's inner-loop index as the reference prints it
  long long residual_calls = 0, dual_steps = 0;
  int dual_kt = 0;
  hipGraphExec_t graph = nullptr;
  bool graph_ok = false;
  bool graph_pend[3] = {false, false, false};  // clip flags the cached graph was captured in (and leaves)
  int graph_ticks[3] = {0, 0, 0};              // tick() calls of one captured pass, per level
  // run to a residual drop (ec2d_run_to_drop): per-case mask and state
  int* dactive = nullptr;
  ConvState* dconv = nullptr;
  int* dcold = nullptr;  // per case: a cell fell below T = 1e-5 in a per-case run
  bool conv_on = false, conv_pending = false;
  double conv_drop = 0.0;
  int conv_max = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_run_ms = 0, res_ms = 0, lusgs_ms = 0, ghost_ms = 0;
  int profile = 0;
  std::vector<hipEvent_t> pev;
  std::vector<int> pev_kind;
  double prof_mean[40] = {0};
  // learned closure (ec2d_closure.hip)
  struct {
    bool ready = false, enabled = false, stalled = false;  // stalled: a model-4 batch whose residual stalled
    double relax = 1.0;                                    // model 4: under-relaxation of mu_t
    ClosureNet* dnet = nullptr;
    float *theta = nullptr, *m1 = nullptr, *m2 = nullptr, *part = nullptr, *losses = nullptr, *data = nullptr;
    long long* cells = nullptr;
    int ncell = 0, step = 0, npart = 0, nlosses = 0;
    int* idx = nullptr;
    float norm[4 * CL_NF] = {0};
    double best_rms = 0.0;
  } clo;
  int prof_count[40] = {0};
};

Runtime R;

// Scratch of ec2d_du_lu_sgs_2d_noalloc: the LU-SGS records (and split-sweep
// rings) of one level in a device array that exists as soon as the library
// is loaded on a device, so that the entry point allocates nothing.  Each GPU
// holds its own copy until the process ends; host code uses only the address
// hipGetSymbolAddress gave for the device (scratch_addr), never the array's
// name, which in host code is a stand-in and not a device address.
constexpr size_t SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) unsigned long long scratch[SCRATCH_BYTES / 8];
constexpr int SCRATCH_DEVICES = 64;
void* scratch_addr[SCRATCH_DEVICES] = {nullptr};  // per device, set once by ec2d_init

int fail(const std::string& m) {
  R.err = m;
  return 1;
}

#define HIPCHK(x)                                                                             \
  do {                                                                                        \
    hipError_t e_ = (x);                                                                      \
    if (e_ != hipSuccess) return fail(std::string(#x) + ": " + hipGetErrorString(e_));       \
  } while (0)

// Ghost-fill phases of one block (k_ghost): the fill runs the boundary
// subfaces in file order, then the interfaces, then the corners; units gets
// that order (level subface indices).  A subface
// joins the running phase (no barrier) when it writes no cell that the phase
// read or wrote and reads no cell that the phase wrote; otherwise it starts a
// new phase (sync = 1).  The cell sets follow bc_point and the interface copy
// in ec2d_ghost.hip exactly; interfaces read neighbour interior cells, which
// no fill writes, so only their writes count.  The corners always follow a
// barrier.
void ghost_phases(const BlockDesc& B, SubfaceDesc* s, int* units) {
  std::vector<int> wph((size_t)B.plane, 0), rph((size_t)B.plane, 0);
  int ph = 1;
  std::vector<int64_t> rd, wr;
  auto q = [&](int i, int j) { return pidx(B, i, j); };
  int nu = 0;
  auto join = [&](SubfaceDesc& sf) {
    units[nu++] = B.bc_off + (int)(&sf - s);
    bool dep = false;
    for (int64_t c : wr) dep = dep || c < 0 || c >= B.plane || wph[c] == ph || rph[c] == ph;
    for (int64_t c : rd) dep = dep || c < 0 || c >= B.plane || wph[c] == ph;
    sf.sync = dep ? 1 : 0;
    if (dep) ph++;
    for (int64_t c : wr) if (c >= 0 && c < B.plane) wph[c] = ph;
    for (int64_t c : rd) if (c >= 0 && c < B.plane) rph[c] = ph;
  };
  for (int k = 0; k < B.nsub; k++) {
    SubfaceDesc& sf = s[k];
    if (sf.neighb >= 0) continue;
    rd.clear();
    wr.clear();
    for (int j = sf.jst; j <= sf.jend; j++)
      for (int i = sf.ist; i <= sf.iend; i++) {
        const int f = sf.face;
        if (sf.neighb == -10 || sf.neighb == -30) {
          for (int l = 1; l <= EC2D_LAP; l++) {
            if (f == 1) { rd.push_back(q(i + l - 1, j)); wr.push_back(q(i - l, j)); }
            else if (f == 2) { rd.push_back(q(i, j + l - 1)); wr.push_back(q(i, j - l)); }
            else if (f == 3) { rd.push_back(q(i - l, j)); wr.push_back(q(i + l - 1, j)); }
            else { rd.push_back(q(i, j - l)); wr.push_back(q(i, j + l - 1)); }
          }
        } else {
          rd.push_back(q(f == 3 ? i - 1 : i, f == 4 ? j - 1 : j));
          for (int l = 1; l <= EC2D_LAP; l++) {
            if (f == 1) wr.push_back(q(i - l, j));
            else if (f == 2) wr.push_back(q(i, j - l));
            else if (f == 3) wr.push_back(q(i + l - 1, j));
            else wr.push_back(q(i, j + l - 1));
          }
        }
      }
    join(sf);
  }
  for (int k = 0; k < B.nsub; k++) {
    SubfaceDesc& sf = s[k];
    if (sf.neighb < 0) continue;
    rd.clear();
    wr.clear();
    const int n1 = sf.iend - sf.ist + sf.jend - sf.jst;
    for (int l = 1; l <= EC2D_LAP; l++)
      for (int kk = 1; kk <= n1; kk++) {
        if (sf.face == 1) wr.push_back(q(sf.ist - EC2D_LAP - 1 + l, sf.jst + kk - 1));
        else if (sf.face == 2) wr.push_back(q(sf.ist + kk - 1, sf.jst - EC2D_LAP - 1 + l));
        else if (sf.face == 3) wr.push_back(q(sf.iend + l - 1, sf.jst + kk - 1));
        else wr.push_back(q(sf.ist + kk - 1, sf.jend + l - 1));
      }
    join(sf);
  }
}

// LU-SGS record layout of a level for lpc lanes per cell (see ec2d_lusgs.hip);
// upload: also refresh the device block table (after ec2d_commit).  fused:
// the layout of k_lusgs_fused, whose single sweep wave needs no whole waves
// of rows, so a step holds exactly the block's rows (R = nrow): the rows of
// consecutive steps share cache lines instead of padding each step to 64.
int lusgs_layout(DLevel& L, int lpc, int split, bool upload, bool fused = false) {
  L.skew_static = false;  // the arena is cleared below
  int max_rows = 1;
  for (auto& d : L.hb) max_rows = std::max(max_rows, d.ny - 1);
  const int rq = 64 / lpc;  // rows per wave
  if (split != 1 && split != 2 && split != 4) return fail("LU-SGS split must be 1, 2 or 4");
  // every workgroup of a split launch must be resident at once: one per CU at most
  if (split > 1 && (max_rows <= rq * (split / 2) || split * (int)L.hb.size() > 256))
    return fail("a split LU-SGS sweep needs blocks taller than split/2 waves and at most 256/split blocks");
  // rows rounded to whole waves per workgroup; split: one communication wave
  // more, two for the middle parts of a split into four
  const int waves = (max_rows + rq * split - 1) / (rq * split) + (split > 2 ? 2 : split - 1);
  const int threads = 64 * waves;
  if (threads > 1024) return fail("blocks with more than 1024 rows of cells are not supported by the LU-SGS kernel");
  const int nv = L.o.nvar;
  const bool visc = R.cases[0].cp.If_viscous == 1;
  const int64_t per = 2 * lusgs_rec_doubles(nv, visc) + lusgs_out_doubles(nv);
  int64_t sk_off = 0;
  L.sk_max_n = 0;
  L.max_cells = 0;
  L.prep_tiles = 0;
  L.skew_tiles = 0;
  for (auto& d : L.hb) {
    const int ncol = d.nx - 1, nrow = d.ny - 1;
    d.sk_S = ncol + nrow - 1;
    d.sk_R = fused ? nrow : rq * split * ((nrow + rq * split - 1) / (rq * split));
    d.sk_n = (int64_t)d.sk_S * d.sk_R + 64;
    d.sk_base = sk_off;
    if (per * d.sk_n * 8 >= (int64_t)1 << 31) return fail("block too large for the LU-SGS record buffer (2 GiB)");
    sk_off += (per * d.sk_n + 31) / 32 * 32;
    L.sk_max_n = std::max(L.sk_max_n, d.sk_n);
    L.max_cells = std::max(L.max_cells, ncol * nrow);
    L.prep_tiles = std::max(L.prep_tiles, ((d.sk_S + 31) / 32) * ((nrow + 7) / 8));
    L.skew_tiles = std::max(L.skew_tiles, ((d.sk_S + 15) / 16) * ((nrow + 15) / 16));
  }
  if (L.sk) HIPCHK(hipFree(L.sk));
  L.sk = nullptr;
  L.sk_total = std::max<int64_t>(sk_off, 1);
  HIPCHK(hipMalloc(&L.sk, sizeof(double) * (size_t)std::max<int64_t>(sk_off, 1)));
  HIPCHK(hipMemset(L.sk, 0, sizeof(double) * (size_t)std::max<int64_t>(sk_off, 1)));
  L.lusgs_lpc = lpc;
  L.lusgs_split = split;
  L.lusgs_threads = threads;
  if (L.dgring) HIPCHK(hipFree(L.dgring));
  L.dgring = nullptr;
  L.gstride = 0;
  if (split > 1) {
    // one ring per boundary between parts of a block and sweep direction, a
    // slot per step (npad <= S + 7) and double, all sentinel
    int max_S = 1;
    for (auto& d : L.hb) max_S = std::max(max_S, d.sk_S);
    L.gstride = ((int64_t)(max_S + 8) * 2 * L.o.nvar + 31) / 32 * 32;
    const int64_t n = (int64_t)L.hb.size() * (split - 1) * 2 * L.gstride;
    HIPCHK(hipMalloc(&L.dgring, sizeof(unsigned long long) * n));
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, R.s, L.dgring, n, GSENT);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(R.s));
  }
  if (upload) HIPCHK(hipMemcpy(L.db, L.hb.data(), sizeof(BlockDesc) * L.hb.size(), hipMemcpyHostToDevice));
  return 0;
}

int reduce_norms(DLevel& L);

// Residual tiles of a level (64 x L.tile_rows cells), their XCD-aware launch
// order and the per-tile norm partials.
int build_tiles(DLevel& L) {
  const int nv = L.o.nvar;
  L.ht.clear();
  L.case_tile0.clear();
  // tiles (logical order: case, block, row of tiles, column of tiles)
  int TX = L.tile_cols, TY = L.tile_rows;
  for (size_t ci = 0; ci < R.cases.size(); ci++) {
    L.case_tile0.push_back((int)L.ht.size());
    for (int b = L.case_first_block[ci]; b < L.case_first_block[ci + 1]; b++) {
      const BlockDesc& bd = L.hb[b];
      for (int j0 = 1; j0 <= bd.ny - 1; j0 += TY)
        for (int i0 = 1; i0 <= bd.nx - 1; i0 += TX) L.ht.push_back({b, i0, j0, 0});
    }
  }
  L.case_tile0.push_back((int)L.ht.size());
  L.ntiles = (int)L.ht.size();
  // XCD-aware launch order: workgroup k runs on XCD k % 8; give each XCD a
  // contiguous run of logical tiles so that halo rows shared by neighbouring
  // tiles are fetched into one L2.  Padding slots are compacted away.
  const int NX = 8;
  int chunk = (L.ntiles + NX - 1) / NX;
  std::vector<int> perm(chunk * NX, -1);
  for (int x = 0; x < NX; x++)
    for (int k = 0; k < chunk; k++) {
      int lt = x * chunk + k;
      if (lt < L.ntiles) perm[k * NX + x] = lt;
    }
  std::vector<int> perm2;
  for (int v : perm)
    if (v >= 0) perm2.push_back(v);
  perm = perm2;
  L.nlaunch = (int)perm.size();
  if (L.residual_done) {
    // keep the last residual's norms: reduce them before the partials go
    if (reduce_norms(L)) return 1;
    HIPCHK(hipStreamSynchronize(R.s));
    L.residual_done = false;
  }
  hipFree(L.dt);
  hipFree(L.dperm);
  hipFree(L.dcase_tile0);
  hipFree(L.dpartials);
  L.dt = nullptr; L.dperm = nullptr; L.dcase_tile0 = nullptr; L.dpartials = nullptr;
  HIPCHK(hipMalloc(&L.dt, sizeof(TileDesc) * L.ht.size()));
  HIPCHK(hipMemcpy(L.dt, L.ht.data(), sizeof(TileDesc) * L.ht.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&L.dperm, sizeof(int) * perm.size()));
  HIPCHK(hipMemcpy(L.dperm, perm.data(), sizeof(int) * perm.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&L.dcase_tile0, sizeof(int) * L.case_tile0.size()));
  HIPCHK(hipMemcpy(L.dcase_tile0, L.case_tile0.data(), sizeof(int) * L.case_tile0.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&L.dpartials, sizeof(double) * (size_t)L.ntiles * 2 * nv));
  return 0;
}

// clo: the learned closure's planes (SA cases, and cases whose eddy viscosity
// the closure supplies)
PlaneMap make_planes(int nv, int nv_un, bool fdm_any, bool clo, bool dual, bool bl) {
  PlaneMap p{};
  int k = 0;
  p.U = k; k += nv;
  p.Res = k; k += nv;
  p.dU = k; k += nv;
  p.dt = k++; p.Lci = k++; p.Lcj = k++; p.Lvi = k++; p.Lvj = k++;
  p.vol = k++; p.si = k++; p.sj = k++; p.ni1 = k++; p.ni2 = k++; p.nj1 = k++; p.nj2 = k++;
  p.x1 = k++; p.y1 = k++; p.Amu = k++; p.Amu_t = k++; p.dw = k++; p.flag = k++;
  p.Un = k; k += nv_un;
  p.QF = k; k += 4;
  p.deltU = k; k += 4;
  if (fdm_any) {
    p.Akx = k++; p.Aky = k++; p.Aix = k++; p.Aiy = k++; p.Ajac = k++;
    p.fx = k; k += 4;
    p.fy = k; k += 4;
    p.ev = k; k += 6;
  } else {
    p.Akx = p.Aky = p.Aix = p.Aiy = p.Ajac = -1;
    p.fx = p.fy = p.ev = -1;
  }
  if (nv == 6) {
    p.sst = k;  // scratch of the SST model
    k += SST_NPLANES;
  } else {
    p.sst = -1;
  }
  if (clo) {
    p.clo = k;  // learned closure: mu_t and its in-range mask
    k += 2;
  } else {
    p.clo = -1;
  }
  if (dual) {
    p.Un1 = k;  // dual-time stepping: U(n-1)
    k += nv;
  } else {
    p.Un1 = -1;
  }
  if (bl) {
    p.bl = k;  // Baldwin-Lomax: vorticity and first-write flag
    k += 2;
  } else {
    p.bl = -1;
  }
  p.nplanes = k;
  return p;
}

int pitch_of(int nx) { return ((nx + 3 + 15) / 16) * 16; }

// reference-layout index helpers
inline int64_t gi(int nx, int i, int j) { return (int64_t)(i + 1) + (int64_t)(nx + 3) * (j + 1); }  // (-1:nx+1,-1:ny+1)
inline int64_t ni_(int nx, int i, int j) { return (int64_t)(i + 1) + (int64_t)(nx + 4) * (j + 1); } // (-1:nx+2,...)
inline int64_t pi_(int nx, int i, int j) { return (int64_t)(i - 1) + (int64_t)nx * (j - 1); }       // (1:nx,1:ny)
inline int64_t ci_(int nx, int i, int j) { return (int64_t)(i - 1) + (int64_t)(nx - 1) * (j - 1); } // (1:nx-1,...)
inline int64_t wi_(int nx, int i, int j) { return (int64_t)i + (int64_t)(nx + 1) * j; }             // (0:nx,0:ny)

int planar_refresh();

// refresh: restore the planes a main-loop pass left in the hyperplane copy
// alone (planar_refresh); the main loop's own entry points and those that
// read none of those planes pass false
int check_ready(bool refresh = true) {
  if (!R.committed) return fail("ec2d_commit has not been called");
  return refresh ? planar_refresh() : 0;
}

// The device's two error words, read once the stream has drained.
//   dsweep_err: an LU-SGS wait between waves or workgroups ran out of its
//     budget; the results since are invalid.  The split-sweep rings are set
//     back to the sentinel here (a receiver that gave up may have left values
//     in them) and the word stays set until ec2d_init_flow: every later check
//     fails.
//   derr: the first cell with T < 1e-5 (sub_Residual.f90:143-146), status 2
//     with the reference's message; also kept until ec2d_init_flow.
// cold: also report derr (the caller may read the arrays of a stopped run).
int device_errors(bool cold) {
  unsigned long long e = ~0ull;
  int se = 0;
  HIPCHK(hipMemcpyAsync(&se, R.dsweep_err, sizeof(se), hipMemcpyDeviceToHost, R.s));
  if (cold) HIPCHK(hipMemcpyAsync(&e, R.derr, sizeof(e), hipMemcpyDeviceToHost, R.s));
  HIPCHK(hipStreamSynchronize(R.s));
  if (se) {
    for (auto& L : R.lv)
      if (L.dgring) {
        const int64_t n = (int64_t)L.hb.size() * (L.lusgs_split - 1) * 2 * L.gstride;
        hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, R.s, L.dgring, n, GSENT);
      }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(R.s));
    return fail("an LU-SGS sweep timed out waiting for another wave or workgroup (results invalid until "
                "ec2d_init_flow)");
  }
  if (e != ~0ull) {
    int blk = (int)(e >> 40), j = (int)((e >> 20) & 0xFFFFF), i = (int)(e & 0xFFFFF);
    char msg[160];
    // the level is not encoded in the key: report the block of the finest level
    int cs = (blk < (int)R.lv[0].blk_case.size()) ? R.lv[0].blk_case[blk] : -1;
    int lb = (blk < (int)R.lv[0].blk_local.size()) ? R.lv[0].blk_local[blk] + 1 : -1;
    std::snprintf(msg, sizeof(msg), "T <1.d-5 ! %d %d %d (case %d)", i, j, lb, cs);
    R.err = msg;
    return 2;
  }
  return 0;
}

// both error words back to "none" (a fresh start of the flow)
int reset_device_errors() {
  const int zero = 0;
  const unsigned long long none = ~0ull;
  HIPCHK(hipMemcpy(R.dsweep_err, &zero, sizeof(zero), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(R.derr, &none, sizeof(none), hipMemcpyHostToDevice));
  return 0;
}

// ---------------------------------------------------------------- launches
int launch_pw(int level, int op, int krk = 0);

// cells of the finest level, all cases, for the closure kernels: the
// interior of SA cases; interior and ghost ring (0:nx, 0:ny, whose mu_t the
// viscous fluxes read) when the closure supplies the eddy viscosity
int closure_cells() {
  if (R.clo.cells) return 0;
  DLevel& L = R.lv[0];
  const int g = L.o.turb == 4 ? 1 : 0;
  std::vector<long long> c;
  for (size_t b = 0; b < L.hb.size(); b++)
    for (int j = 1 - g; j <= L.hb[b].ny - 1 + g; j++)
      for (int i = 1 - g; i <= L.hb[b].nx - 1 + g; i++)
        c.push_back(((long long)b << 40) | ((long long)j << 20) | i);
  R.clo.ncell = (int)c.size();
  HIPCHK(hipMalloc(&R.clo.cells, sizeof(long long) * c.size()));
  HIPCHK(hipMemcpyAsync(R.clo.cells, c.data(), sizeof(long long) * c.size(), hipMemcpyHostToDevice, R.s));
  HIPCHK(hipStreamSynchronize(R.s));
  return 0;
}

ClosureArgs closure_args() {
  DLevel& L = R.lv[0];
  ClosureArgs a;
  a.blocks = L.db;
  a.cells = R.clo.cells;
  a.ncell = R.clo.ncell;
  a.cps = R.dcps;
  a.arena = L.arena;
  a.pm = L.pm;
  a.net = R.clo.dnet;
  a.data = R.clo.data;
  a.standin = L.o.turb == 4 ? 1 : 0;
  a.relax = R.clo.relax;
  return a;
}

int closure_infer() {
  if (!R.clo.ready) return fail("closure has no parameters (train or set them first)");
  if (closure_cells()) return 1;
  ClosureArgs a = closure_args();
  hipLaunchKernelGGL(k_closure_infer, dim3((unsigned)((R.clo.ncell + 63) / 64)), dim3(256), 0, R.s, a);
  HIPCHK(hipGetLastError());
  return 0;
}

int launch_residual(int level, int clip, int force = 0, int set_un = 0, double* sk_out = nullptr,
                    bool sk_only = false) {
  DLevel& L = R.lv[level];
  ResArgs a;
  a.tiles = L.dt;
  a.tile_perm = L.dperm;
  a.blocks = L.db;
  a.cps = R.dcps;
  a.arena = L.arena;
  a.partials = L.dpartials;
  a.err = R.derr;
  a.cold = R.conv_on ? R.dcold : nullptr;
  a.o = L.o;
  a.pm = L.pm;
  a.clip = clip;
  a.sk_out = sk_out;
  a.sk_only = sk_out != nullptr && sk_only ? 1 : 0;
  a.write_visc = 1;
  a.force = force;
  a.set_un = set_un;
  // learned closure: 1 = SA case with mu_t overridden where in range,
  // 2 = the closure supplies mu_t of every cell (turbulence model 4)
  a.closure = (level == 0 && L.o.turb == 4) ? 2 : (level == 0 && R.clo.enabled && L.pm.clo >= 0) ? 1 : 0;
  if (L.pm.bl >= 0) {
    // Baldwin-Lomax: mu_t of every cell of the level before the mean-flow
    // residual, which takes it as given (turbulence_model_BL, called from
    // Residual, sub_Residual.f90:168-171)
    a.closure = 3;
    BlArgs ba;
    ba.blocks = L.db;
    ba.subs = L.ds;
    ba.cps = R.dcps;
    ba.arena = L.arena;
    ba.pm = L.pm;
    const unsigned nb = (unsigned)L.hb.size();
    hipLaunchKernelGGL(k_bl_init, dim3((unsigned)((L.max_plane + 255) / 256), nb), dim3(256), 0, R.s, ba);
    for (int w = 0; w < L.bl_walls; w++) {
      hipLaunchKernelGGL(k_bl_lines, dim3((unsigned)std::max(1, (L.bl_lines + 3) / 4), nb), dim3(256), 0, R.s, ba, w);
      hipLaunchKernelGGL(k_bl_ghost, dim3(nb), dim3(256), 0, R.s, ba, w);
    }
    HIPCHK(hipGetLastError());
  }
  // model 4 with relaxation 0: the closure's mu_t stays frozen, the network is not evaluated
  if ((a.closure == 1 || (a.closure == 2 && R.clo.relax > 0.0)) && closure_infer()) return 1;
  bool visc = R.cases[0].cp.If_viscous == 1;
  int nv = L.o.nvar;
  int turb = L.o.turb;
  dim3 grid(L.nlaunch);
  if (nv == 6 && turb == 3) {
    // SST_kw runs before the mean-flow residual (sub_Residual.f90:158-191);
    // the nu_t / k / omega clip of the previous step comes first
    if (clip) {
      if (launch_pw(level, PW_CLIP)) return 1;
      a.clip = 0;
    }
    SstArgs sa;
    sa.blocks = L.db;
    sa.rows = L.drows;
    sa.cps = R.dcps;
    sa.arena = L.arena;
    sa.pm = L.pm;
    hipLaunchKernelGGL(k_sst_cell, dim3(L.nrows), dim3(256), 0, R.s, sa);
    hipLaunchKernelGGL(k_sst_res, dim3(L.nrows), dim3(256), 0, R.s, sa);
  }
  // the benched options (MUSCL3 / van Albada on both face directions, Roe,
  // original reconstruction) as compile-time constants, for the Euler and SA
  // equation sets
  const bool fix = L.o.scheme == 5 && L.o.global_scheme == 5 && L.o.flux == 4 && L.o.recon == 0 &&
                   ((nv == 4 && !visc) || (nv == 5 && turb == 2));
#define RLAUNCH(TX, TY, NV, V, S)                                                                  \
  do {                                                                                            \
    if (fix && ((NV == 4 && !V) || (NV == 5 && S)))                                               \
      hipLaunchKernelGGL((k_residual<TX, TY, NV, V, S, (NV == 4 && !V) || (NV == 5 && S)>), grid, \
                         dim3(TX * TY), 0, R.s, a);                                               \
    else                                                                                          \
      hipLaunchKernelGGL((k_residual<TX, TY, NV, V, S, false>), grid, dim3(TX * TY), 0, R.s, a);  \
  } while (0)
#define RBY_TY(NV, V, S)                        \
  if (L.tile_cols == 32) {                      \
    if (L.tile_rows == 16)                      \
      RLAUNCH(32, 16, NV, V, S);                \
    else                                        \
      RLAUNCH(32, 8, NV, V, S);                 \
  } else {                                      \
    switch (L.tile_rows) {                      \
      case 1: RLAUNCH(64, 1, NV, V, S); break;  \
      case 2: RLAUNCH(64, 2, NV, V, S); break;  \
      case 8: RLAUNCH(64, 8, NV, V, S); break;  \
      default: RLAUNCH(64, 4, NV, V, S); break; \
    }                                           \
  }
  if (nv == 4 && !visc) {
    RBY_TY(4, false, false)
  } else if (nv == 4 && visc) {
    RBY_TY(4, true, false)
  } else if (nv == 5 && turb == 2) {
    RBY_TY(5, true, true)
  } else if (nv == 6 && turb == 3) {
    RBY_TY(6, true, false)
  } else {
    return fail("turbulence model / equation count not supported by the residual kernel");
  }
#undef RBY_TY
#undef RLAUNCH
  HIPCHK(hipGetLastError());
  if (L.fdm) {
    // finite-difference interior of FDM.in blocks, then norms over the final Res
    FdmArgs f;
    f.blocks = L.db;
    f.cps = R.dcps;
    f.arena = L.arena;
    f.pm = L.pm;
    f.fx = L.pm.fx;
    f.fy = L.pm.fy;
    f.ev = L.pm.ev;
    f.viscous = visc ? 1 : 0;
    f.n_i = L.n_ilines;
    f.n_lines = L.n_ilines + L.n_jlines;
    f.lines = L.d_ilines;
    f.rows = L.d_frows;
    // line fluxes and, viscous, the cell-centre viscous fluxes (which read the Amu / Amu_t planes
    // k_residual has just written) in one launch
    hipLaunchKernelGGL(k_fdm_lines, dim3(f.n_lines + (visc ? L.n_frows : 0)), dim3(256), 0, R.s, f);
    hipLaunchKernelGGL(k_fdm_res, dim3(L.n_frows), dim3(256), 0, R.s, f);
    if (nv == 4) hipLaunchKernelGGL(k_tile_norms<4>, dim3(L.ntiles), dim3(256), 0, R.s, L.dt, L.db, L.arena, L.pm, L.dpartials, L.tile_cols, L.tile_rows);
    else if (nv == 5) hipLaunchKernelGGL(k_tile_norms<5>, dim3(L.ntiles), dim3(256), 0, R.s, L.dt, L.db, L.arena, L.pm, L.dpartials, L.tile_cols, L.tile_rows);
    else hipLaunchKernelGGL(k_tile_norms<6>, dim3(L.ntiles), dim3(256), 0, R.s, L.dt, L.db, L.arena, L.pm, L.dpartials, L.tile_cols, L.tile_rows);
    HIPCHK(hipGetLastError());
  }
  // the per-tile partials hold this residual's norms; they are reduced when
  // the host asks for them (reduce_norms), since nothing on the device reads
  // them and every residual launch of the level rewrites all of them
  L.residual_done = true;
  return 0;
}

// Res_max / Res_rms of every case of a level from the per-tile partials of
// the level's last residual launch.
int reduce_norms(DLevel& L) {
  const int nv = L.o.nvar, ncase = (int)R.cases.size();
  if (nv == 4) hipLaunchKernelGGL(k_norms<4>, dim3(ncase), dim3(256), 0, R.s, L.dpartials, L.dcase_tile0, L.dncell, L.dnorms);
  else if (nv == 5) hipLaunchKernelGGL(k_norms<5>, dim3(ncase), dim3(256), 0, R.s, L.dpartials, L.dcase_tile0, L.dncell, L.dnorms);
  else hipLaunchKernelGGL(k_norms<6>, dim3(ncase), dim3(256), 0, R.s, L.dpartials, L.dcase_tile0, L.dncell, L.dnorms);
  HIPCHK(hipGetLastError());
  return 0;
}

void prof_mark(int kind);

// sk, gring: the record arena and split-sweep rings to use (null: the level's own)
// skew_done: the residual before it wrote the changing inputs in hyperplane
// order (fused path, the level's own record arena)
int launch_lusgs(int level, const int* dlist, int nblk, double alfa1, int update, double* sk = nullptr,
                 unsigned long long* gring = nullptr, bool skew_done = false) {
  DLevel& L = R.lv[level];
  LusgsArgs a;
  a.blocks = L.db;
  a.block_list = dlist;
  a.cps = R.dcps;
  a.arena = L.arena;
  a.sk = sk ? sk : L.sk;
  a.pm = L.pm;
  a.alfa1 = alfa1;
  a.update = update;
  a.split = L.lusgs_split;
  a.dir = 0;
  a.gring = gring ? gring : L.dgring;
  a.gstride = L.gstride;
  a.err = R.dsweep_err;
  a.spin_max = R.wait_budget;
  a.nblk = nblk;
  a.sk_tiles = std::max(L.skew_tiles, 1);
  bool visc = R.cases[0].cp.If_viscous == 1;
  // prep and scatter grid: tiles of 32 steps x 8 rows of the largest block's skewed space
  dim3 gprep((unsigned)L.prep_tiles, nblk), gsweep(nblk * L.lusgs_split), blk(L.lusgs_threads);
  dim3 gscat((unsigned)L.prep_tiles, nblk);
  // k_lusgs_skew: a one-dimensional grid of 8 equal XCD chunks of the
  // (block, tile) items
  dim3 gskew((unsigned)(8 * (((int64_t)a.sk_tiles * nblk + 7) / 8)));
  int nv = L.o.nvar;
  // the instantiation is chosen by thread count (its launch bound): up to 4
  // waves one wave per SIMD and two steps per register set; more waves (up
  // to 16) one step per set
  const int nth = L.lusgs_threads;
  const bool pair = L.lusgs_lpc == 2;
  // every instantiation below is compiled for at most its MAXT threads
  if (nth < 64 || nth > 1024 || nth % 64 != 0) return fail("LU-SGS launch shape out of range");
  const int pk = 10 * level;
  // the fused launch's skewed copy of the mesh planes: once per layout for
  // the level's own record arena (when a launch covers every block), on
  // every launch into a caller's arena
  const bool static_skew = L.lusgs_fused && (sk != nullptr || !L.skew_static);
  if (L.lusgs_fused && sk == nullptr && dlist == L.dblocks_all && nblk == (int)L.hb.size()) L.skew_static = true;
#define LLAUNCH(NV, V)                                                                      \
  if (L.lusgs_fused) {                                                                      \
    prof_mark(4 + pk);                                                                      \
    if (static_skew)                                                                        \
      hipLaunchKernelGGL((k_lusgs_skew<NV, V, true>), gskew, dim3(256), 0, R.s, a);         \
    if (!skew_done)                                                                         \
      hipLaunchKernelGGL((k_lusgs_skew<NV, V, false>), gskew, dim3(256), 0, R.s, a);        \
    prof_mark(-(4 + pk));                                                                   \
    prof_mark(5 + pk);                                                                      \
    if (L.fused_np == 3)                                                                    \
      hipLaunchKernelGGL((k_lusgs_fused<NV, V, 3>), dim3(nblk), dim3(320), 0, R.s, a);      \
    else                                                                                    \
      hipLaunchKernelGGL((k_lusgs_fused<NV, V, 6>), dim3(nblk), dim3(512), 0, R.s, a);      \
    prof_mark(-(5 + pk));                                                                   \
  } else {                                                                                  \
  prof_mark(4 + pk);                                                                        \
  hipLaunchKernelGGL((k_lusgs_prep<NV, V>), gprep, dim3(256), 0, R.s, a);                   \
  prof_mark(-(4 + pk));                                                                     \
  prof_mark(5 + pk);                                                                        \
  for (int d = (a.split > 1 ? 1 : 0); d <= (a.split > 1 ? 2 : 0); d++) {                  \
  a.dir = d;                                                                                \
  if (pair) {                                                                                \
    if (nth <= 64 && L.lusgs_u == 4) hipLaunchKernelGGL((k_lusgs_sweep<NV, V, 4, 64, 2>), gsweep, blk, 0, R.s, a); \
    else if (nth <= 64) hipLaunchKernelGGL((k_lusgs_sweep<NV, V, 2, 64, 2>), gsweep, blk, 0, R.s, a); \
    else if (nth <= 256 && L.lusgs_u == 4) hipLaunchKernelGGL((k_lusgs_sweep<NV, V, 4, 256, 2>), gsweep, blk, 0, R.s, a); \
    else if (nth <= 256 && L.lusgs_u == 3) hipLaunchKernelGGL((k_lusgs_sweep<NV, V, 3, 256, 2>), gsweep, blk, 0, R.s, a); \
    else if (nth <= 256) hipLaunchKernelGGL((k_lusgs_sweep<NV, V, 2, 256, 2>), gsweep, blk, 0, R.s, a); \
    else if (nth <= 512) hipLaunchKernelGGL((k_lusgs_sweep<NV, V, 1, 512, 2>), gsweep, blk, 0, R.s, a); \
    else hipLaunchKernelGGL((k_lusgs_sweep<NV, V, 1, 1024, 2>), gsweep, blk, 0, R.s, a);       \
  } else {                                                                                  \
    if (nth <= 256) hipLaunchKernelGGL((k_lusgs_sweep<NV, V, 2, 256, 1>), gsweep, blk, 0, R.s, a); \
    else if (nth <= 512) hipLaunchKernelGGL((k_lusgs_sweep<NV, V, 1, 512, 1>), gsweep, blk, 0, R.s, a); \
    else hipLaunchKernelGGL((k_lusgs_sweep<NV, V, 1, 1024, 1>), gsweep, blk, 0, R.s, a);       \
  }                                                                                         \
  }                                                                                         \
  prof_mark(-(5 + pk));                                                                     \
  prof_mark(6 + pk);                                                                        \
  hipLaunchKernelGGL((k_lusgs_scatter<NV, V>), gscat, dim3(256), 0, R.s, a);                \
  prof_mark(-(6 + pk));                                                                     \
  }
  if (nv == 4 && !visc) { LLAUNCH(4, false) }
  else if (nv == 4) { LLAUNCH(4, true) }
  else if (nv == 5) { LLAUNCH(5, true) }
  else if (nv == 6) { LLAUNCH(6, true) }
#undef LLAUNCH
  else return fail("bad equation count");
  HIPCHK(hipGetLastError());
  return 0;
}

// Res, dt, Lci, Lcj, Lvi and Lvj of a level back into their planes from the
// fused LU-SGS launch's hyperplane copy, after passes whose residual wrote
// them there alone (advance_lusgs)
int planar_refresh() {
  for (int l = 0; l < R.num_mesh; l++) {
    DLevel& L = R.lv[l];
    if (!L.planar_stale) continue;
    L.planar_stale = false;
    LusgsArgs a{};
    a.blocks = L.db;
    a.block_list = L.dblocks_all;
    a.arena = L.arena;
    a.sk = L.sk;
    a.pm = L.pm;
    a.nblk = (int)L.hb.size();
    a.sk_tiles = std::max(L.skew_tiles, 1);
    const dim3 g((unsigned)(8 * (((int64_t)a.sk_tiles * a.nblk + 7) / 8)));
    const int nv = L.o.nvar;
    const bool visc = R.cases[0].cp.If_viscous == 1;
    if (nv == 4 && !visc) hipLaunchKernelGGL((k_lusgs_unskew<4, false>), g, dim3(256), 0, R.s, a);
    else if (nv == 4) hipLaunchKernelGGL((k_lusgs_unskew<4, true>), g, dim3(256), 0, R.s, a);
    else if (nv == 5) hipLaunchKernelGGL((k_lusgs_unskew<5, true>), g, dim3(256), 0, R.s, a);
    else if (nv == 6) hipLaunchKernelGGL((k_lusgs_unskew<6, true>), g, dim3(256), 0, R.s, a);
    else return fail("bad equation count");
    HIPCHK(hipGetLastError());
  }
  return 0;
}

int launch_ghost(int level, int do_bc, int do_exchange) {
  DLevel& L = R.lv[level];
  GhostArgs a;
  a.blocks = L.db;
  a.block_list = L.dblocks_all;
  a.subs = L.ds;
  a.units = L.dunits;
  a.normals = L.dnorm;
  a.cps = R.dcps;
  a.arena = L.arena;
  a.pm = L.pm;
  a.nvar = L.o.nvar;
  a.do_bc = do_bc;
  a.do_exchange = do_exchange;
  // 1024 threads: a subface of up to 1024 points in one pass per ordered phase
  dim3 grid((int)L.hb.size()), blk(1024);
  switch (L.o.nvar) {
    case 4: hipLaunchKernelGGL(k_ghost<4>, grid, blk, 0, R.s, a); break;
    case 5: hipLaunchKernelGGL(k_ghost<5>, grid, blk, 0, R.s, a); break;
    case 6: hipLaunchKernelGGL(k_ghost<6>, grid, blk, 0, R.s, a); break;
  }
  HIPCHK(hipGetLastError());
  return 0;
}

int launch_pw(int level, int op, int krk) {
  DLevel& L = R.lv[level];
  PwArgs a;
  a.blocks = L.db;
  a.rows = L.drows;
  a.arena = L.arena;
  a.pm = L.pm;
  a.nvar = L.o.nvar;
  a.op = op;
  a.krk = krk;
  a.sfac = 0.0;
  const double Ra[3] = {1.0, 3.0 / 4.0, 1.0 / 3.0}, Rb[3] = {1.0, 1.0 / 4.0, 2.0 / 3.0}, Rg[3] = {0.0, 1.0 / 4.0, 2.0 / 3.0};
  if (krk >= 1) {
    a.ra = Ra[krk - 1];
    a.rb = Rb[krk - 1];
    a.rg = Rg[krk - 1];
  } else {
    a.ra = a.rb = a.rg = 0.0;
  }
  hipLaunchKernelGGL(k_pointwise, dim3(L.nrows), dim3(256), 0, R.s, a);
  HIPCHK(hipGetLastError());
  return 0;
}

int launch_xfer(int fine, int coarse, int flag, bool restrict_) {
  DLevel& F = R.lv[fine];
  DLevel& C = R.lv[coarse];
  XferArgs a;
  a.fine_blocks = F.db;
  a.coarse_blocks = C.db;
  a.fine_arena = F.arena;
  a.coarse_arena = C.arena;
  a.fpm = F.pm;
  a.cpm = C.pm;
  a.flag = flag;
  if (restrict_) {
    a.rows = C.drows;
    hipLaunchKernelGGL(k_restrict, dim3(C.nrows), dim3(256), 0, R.s, a);
  } else {
    a.rows = F.drows;
    hipLaunchKernelGGL(k_prolong, dim3(F.nrows), dim3(256), 0, R.s, a);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

void prof_mark(int kind) {
  if (!R.profile) return;
  hipEvent_t e;
  hipEventCreate(&e);
  hipEventRecord(e, R.s);
  R.pev.push_back(e);
  R.pev_kind.push_back(kind);
}

int flush_clip(int level) {
  if (R.clip_pending[level]) {
    R.clip_pending[level] = false;
    return launch_pw(level, PW_CLIP);
  }
  return 0;
}

// ------------------------------------------------------------ composite steps
// force: 1 = comput_force_function, 2 = Add_force_function fused into the
// residual's store (coarse levels; see ResArgs::force); set_un: Set_Un too
// sk_out: see ResArgs (the fused LU-SGS that follows then skips its copy)
int residual(int level, int force = 0, int set_un = 0, double* sk_out = nullptr, bool sk_only = false) {
  int clip = R.clip_pending[level] ? 1 : 0;
  R.clip_pending[level] = false;
  prof_mark(1 + 10 * level);
  int rc = launch_residual(level, clip, force, set_un, sk_out, sk_only);
  if (!(sk_out && sk_only)) R.lv[level].planar_stale = false;  // the planes hold this residual's values
  if (!rc && R.cases[0].ctl.Time_Method == -1) {
    // dual-time source after the whole residual and before the norms:
    // Sfac = 1/(3 dt) on the first residual call, 1/(2 dt) afterwards
    DLevel& L = R.lv[level];
    const double dtg = R.cases[0].ctl.dt_global;
    PwArgs a;
    a.blocks = L.db;
    a.rows = L.drows;
    a.arena = L.arena;
    a.pm = L.pm;
    a.nvar = L.o.nvar;
    a.op = PW_DUAL_SRC;
    a.krk = 0;
    a.ra = a.rb = a.rg = 0.0;
    a.sfac = R.residual_calls == 0 ? 1.0 / (3.0 * dtg) : 1.0 / (2.0 * dtg);
    hipLaunchKernelGGL(k_pointwise, dim3(L.nrows), dim3(256), 0, R.s, a);
    const int nv = L.o.nvar;
    if (nv == 4) hipLaunchKernelGGL(k_tile_norms<4>, dim3(L.ntiles), dim3(256), 0, R.s, L.dt, L.db, L.arena, L.pm, L.dpartials, L.tile_cols, L.tile_rows);
    else if (nv == 5) hipLaunchKernelGGL(k_tile_norms<5>, dim3(L.ntiles), dim3(256), 0, R.s, L.dt, L.db, L.arena, L.pm, L.dpartials, L.tile_cols, L.tile_rows);
    else hipLaunchKernelGGL(k_tile_norms<6>, dim3(L.ntiles), dim3(256), 0, R.s, L.dt, L.db, L.arena, L.pm, L.dpartials, L.tile_cols, L.tile_rows);
    rc = hipGetLastError() == hipSuccess ? 0 : fail("dual-time source launch failed");
  }
  R.residual_calls++;
  if (!rc && level == 0 && R.conv_pending) {
    // run to a residual drop: the first fine residual of the pass decides
    // which cases go on (their LU-SGS / update / ghost launches follow)
    R.conv_pending = false;
    DLevel& L = R.lv[0];
    hipLaunchKernelGGL(k_converge, dim3((int)R.cases.size()), dim3(256), 0, R.s, L.dpartials, L.dcase_tile0,
                       L.dncell, L.o.nvar, R.dactive, R.dconv, R.conv_drop, R.conv_max, R.dcold);
    rc = hipGetLastError() == hipSuccess ? 0 : fail("convergence check launch failed");
  }
  prof_mark(-(1 + 10 * level));
  return rc;
}

int tick_count[3] = {0, 0, 0};

void tick(int level) {
  tick_count[level]++;
  for (auto& c : R.cases) {
    c.tt[level] = c.tt[level] + c.ctl.dt_global;
    c.kstep[level] += 1;
  }
}

int advance_lusgs(int level, bool add_force) {
  DLevel& L = R.lv[level];
  // the fused launch's changing inputs in hyperplane order straight from the
  // residual's epilogue (not on FD-WENO5 levels, whose interior Res a later
  // launch writes), in place of its copy k_lusgs_skew
  const bool rskew = L.lusgs_fused && !L.fdm;
  // On a single level (steady LU-SGS) nothing reads Res, dt, Lc or Lv of
  // the planes before the next pass's residual rewrites them, so they go to
  // the hyperplane copy alone; an entry point that may read them restores
  // the planes first (planar_refresh, from check_ready).  A V-cycle's
  // restriction reads Res: there every residual writes the planes.
  const bool sko = rskew && R.num_mesh == 1;
  if (residual(level, add_force ? 2 : 0, 0, rskew ? L.sk : nullptr, sko)) return 1;
  if (sko) {
    L.planar_stale = true;
    R.sko_passes++;
  }
  prof_mark(2 + 10 * level);
  if (launch_lusgs(level, L.dblocks_all, (int)L.hb.size(), 0.0, 1, nullptr, nullptr, rskew)) return 1;
  prof_mark(-(2 + 10 * level));
  prof_mark(3 + 10 * level);
  if (launch_ghost(level, 1, 1)) return 1;
  prof_mark(-(3 + 10 * level));
  tick(level);
  return 0;
}

int advance_euler1(int level) {
  if (residual(level, level != 0 ? 2 : 0)) return 1;
  if (launch_pw(level, PW_EULER1)) return 1;
  if (launch_ghost(level, 1, 1)) return 1;
  tick(level);
  return 0;
}

int advance_rk3(int level) {
  if (flush_clip(level)) return 1;
  if (launch_pw(level, PW_COPY_UN)) return 1;
  for (int k = 1; k <= 3; k++) {
    if (residual(level, level != 0 ? 2 : 0)) return 1;
    if (launch_pw(level, PW_RK3, k)) return 1;
    if (launch_ghost(level, 1, 1)) return 1;
  }
  tick(level);
  return 0;
}

// Dual_time_LU_SGS (sub_NS_singlegrid.f90:231-310): inner LU-SGS iterations
// with alfa1 = 1/dt (first physical step) or 3/(2 dt), until max_m Res_rms(m)
// <= Res_Inner_Limit or Nstep_Inner_Limit iterations; then Un1 = Un, Un = U.
// The exit test reads the norms on the host after every inner iteration.
int dual_time_lusgs() {
  DLevel& L = R.lv[0];
  const ec2d_control& k = R.cases[0].ctl;
  const double dtg = k.dt_global;
  const double alfa1 = R.dual_steps == 0 ? 1.0 / dtg : 3.0 / (2.0 * dtg);
  R.dual_steps++;
  int kt = 1;
  for (; kt <= k.Nstep_Inner_Limit; kt++) {
    if (residual(0)) return 1;
    prof_mark(2);
    if (launch_lusgs(0, L.dblocks_all, (int)L.hb.size(), alfa1, 1)) return 1;
    prof_mark(-2);
    prof_mark(3);
    if (launch_ghost(0, 1, 1)) return 1;
    prof_mark(-3);
    double mx[6], rms[6];
    const int rc = ec2d_res_norms(0, 1, mx, rms);
    if (rc) return rc;
    double mr = rms[0];
    for (int m = 0; m < L.o.nvar; m++) mr = std::max(mr, rms[m]);
    if (mr <= k.Res_Inner_Limit) break;
  }
  R.dual_kt = kt;  // Nstep_Inner_Limit + 1 when the loop ran out, as the reference prints it
  if (launch_pw(0, PW_DUAL_SHIFT)) return 1;
  tick(0);
  return 0;
}

int time_advance(int level) {
  int tm = R.cases[0].ctl.Time_Method;
  int rc;
  if (tm == 0) rc = advance_lusgs(level, false);
  else if (tm == 1) rc = advance_euler1(level);
  else if (tm == 3) rc = advance_rk3(level);
  else if (tm == -1 && level == 0) rc = dual_time_lusgs();
  else return fail("time method not on this path");
  if (rc) return rc;
  if (R.lv[level].o.nvar >= 5) R.clip_pending[level] = true;
  return 0;
}

// un_set: Set_Un already ran (fused into the residual that formed the
// forcing function; the state has not changed since)
int smooth(int level, bool un_set) {
  int tm = R.cases[0].ctl.Time_Method;
  if (tm == 1) {
    if (!un_set && launch_pw(level, PW_SET_UN)) return 1;
    for (int m = 0; m < 3; m++)
      if (advance_euler1(level)) return 1;
    return 0;
  }
  if (tm == 3) return advance_rk3(level);
  // LU-SGS smoother with the forcing function added to the residual: an
  // extension (the reference rejects LU-SGS with multigrid, :216-221)
  if (!un_set && launch_pw(level, PW_SET_UN)) return 1;
  return advance_lusgs(level, level != 0);
}

int multigrid() {
  int nm = R.num_mesh;
  int tm = R.cases[0].ctl.Time_Method;
  if (tm == 1) {
    if (advance_euler1(0)) return 1;
  } else if (tm == 3) {
    if (advance_rk3(0)) return 1;
  } else {
    // LU-SGS-smoothed cycle (extension): the fine step clips nu_t like
    // NS_Time_advance does; the clip is applied by the next residual
    if (advance_lusgs(0, false)) return 1;
    if (R.lv[0].o.nvar >= 5) R.clip_pending[0] = true;
  }
  if (residual(0)) return 1;
  if (launch_xfer(0, 1, 3, true)) return 1;  // Res -> QF and U, one launch
  if (launch_ghost(1, 1, 1)) return 1;
  // Set_Un of the Euler and LU-SGS smoothers rides on the residual that
  // forms the forcing function (RK3 copies Un itself)
  const int un = (tm == 1 || tm == 0) ? 1 : 0;
  if (residual(1, 1, un)) return 1;
  if (smooth(1, un != 0)) return 1;
  if (nm >= 3) {
    if (residual(1, 2)) return 1;
    if (launch_xfer(1, 2, 3, true)) return 1;
    if (launch_ghost(2, 1, 1)) return 1;
    if (residual(2, 1, un)) return 1;
    if (smooth(2, un != 0)) return 1;
    if (launch_xfer(1, 2, 3, false)) return 1;  // comput_delt_U, prolong_U, comput_new_U
  }
  if (launch_xfer(0, 1, 3, false)) return 1;
  if (launch_ghost(0, 1, 1)) return 1;
  return 0;
}

int main_step() {
  R.conv_pending = R.conv_on;
  if (R.num_mesh == 1) return time_advance(0);
  return multigrid();
}

// A captured run graph bakes in launch shapes and kernel arguments: any call
// that changes them drops it (the next ec2d_run captures a new one).
int drop_graph() {
  if (R.graph) {
    HIPCHK(hipGraphExecDestroy(R.graph));
    R.graph = nullptr;
  }
  return 0;
}

// -------------------------------------------------- host <-> device layouts
void to_plane(const BlockDesc& b, double* plane, const double* src, int kind, int nv, int m) {
  // kind: 0 ghosted U-like (nv comps), 1 (1:nx,1:ny), 2 (1:nx-1,1:ny-1) with nv comps, 3 (0:nx,0:ny) nv comps,
  //       4 (1:nx,1:ny) with nv comps, 5 ghosted scalar
  int nx = b.nx, ny = b.ny;
  switch (kind) {
    case 0:
      for (int j = -1; j <= ny + 1; j++)
        for (int i = -1; i <= nx + 1; i++) plane[pidx(b, i, j)] = src[m + nv * gi(nx, i, j)];
      break;
    case 5:
      for (int j = -1; j <= ny + 1; j++)
        for (int i = -1; i <= nx + 1; i++) plane[pidx(b, i, j)] = src[gi(nx, i, j)];
      break;
    case 1:
      for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) plane[pidx(b, i, j)] = src[pi_(nx, i, j)];
      break;
    case 4:
      for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) plane[pidx(b, i, j)] = src[m + nv * pi_(nx, i, j)];
      break;
    case 2:
      for (int j = 1; j <= ny - 1; j++)
        for (int i = 1; i <= nx - 1; i++) plane[pidx(b, i, j)] = src[m + nv * ci_(nx, i, j)];
      break;
    case 3:
      for (int j = 0; j <= ny; j++)
        for (int i = 0; i <= nx; i++) plane[pidx(b, i, j)] = src[m + nv * wi_(nx, i, j)];
      break;
  }
}

void from_plane(const BlockDesc& b, const double* plane, double* dst, int kind, int nv, int m) {
  int nx = b.nx, ny = b.ny;
  switch (kind) {
    case 0:
      for (int j = -1; j <= ny + 1; j++)
        for (int i = -1; i <= nx + 1; i++) dst[m + nv * gi(nx, i, j)] = plane[pidx(b, i, j)];
      break;
    case 5:
      for (int j = -1; j <= ny + 1; j++)
        for (int i = -1; i <= nx + 1; i++) dst[gi(nx, i, j)] = plane[pidx(b, i, j)];
      break;
    case 1:
      for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) dst[pi_(nx, i, j)] = plane[pidx(b, i, j)];
      break;
    case 4:
      for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) dst[m + nv * pi_(nx, i, j)] = plane[pidx(b, i, j)];
      break;
    case 2:
      for (int j = 1; j <= ny - 1; j++)
        for (int i = 1; i <= nx - 1; i++) dst[m + nv * ci_(nx, i, j)] = plane[pidx(b, i, j)];
      break;
    case 3:
      for (int j = 0; j <= ny; j++)
        for (int i = 0; i <= nx; i++) dst[m + nv * wi_(nx, i, j)] = plane[pidx(b, i, j)];
      break;
  }
}

// which -> (first plane, components, layout kind)
int array_map(const DLevel& L, int which, int* slot, int* ncomp, int* kind) {
  const PlaneMap& p = L.pm;
  switch (which) {
    case EC2D_U: *slot = p.U; *ncomp = L.o.nvar; *kind = 0; return 0;
    case EC2D_RES: *slot = p.Res; *ncomp = L.o.nvar; *kind = 2; return 0;
    case EC2D_DT: *slot = p.dt; *ncomp = 1; *kind = 2; return 0;
    case EC2D_DU: *slot = p.dU; *ncomp = L.o.nvar; *kind = 4; return 0;
    case EC2D_LCI: *slot = p.Lci; *ncomp = 1; *kind = 1; return 0;
    case EC2D_LCJ: *slot = p.Lcj; *ncomp = 1; *kind = 1; return 0;
    case EC2D_LVI: *slot = p.Lvi; *ncomp = 1; *kind = 1; return 0;
    case EC2D_LVJ: *slot = p.Lvj; *ncomp = 1; *kind = 1; return 0;
    case EC2D_AMU: *slot = p.Amu; *ncomp = 1; *kind = 5; return 0;
    case EC2D_AMU_T: *slot = p.Amu_t; *ncomp = 1; *kind = 5; return 0;
    case EC2D_UN: *slot = p.Un; *ncomp = L.o.nvar_un; *kind = 3; return 0;
    case EC2D_QF: *slot = p.QF; *ncomp = 4; *kind = 2; return 0;
    case EC2D_DELTU: *slot = p.deltU; *ncomp = 4; *kind = 0; return 0;
    case EC2D_DW: *slot = p.dw; *ncomp = 1; *kind = 3; return 0;
    case EC2D_UN1: if (p.Un1 < 0) return 1; *slot = p.Un1; *ncomp = L.o.nvar; *kind = 3; return 0;
    case EC2D_CLOSURE_MUT: if (p.clo < 0) return 1; *slot = p.clo; *ncomp = 1; *kind = 5; return 0;
    case EC2D_CLOSURE_MASK: if (p.clo < 0) return 1; *slot = p.clo + 1; *ncomp = 1; *kind = 5; return 0;
    case EC2D_VOL: *slot = p.vol; *ncomp = 1; *kind = 1; return 0;
    case EC2D_SI: *slot = p.si; *ncomp = 1; *kind = 1; return 0;
    case EC2D_SJ: *slot = p.sj; *ncomp = 1; *kind = 1; return 0;
    case EC2D_NI1: *slot = p.ni1; *ncomp = 1; *kind = 1; return 0;
    case EC2D_NI2: *slot = p.ni2; *ncomp = 1; *kind = 1; return 0;
    case EC2D_NJ1: *slot = p.nj1; *ncomp = 1; *kind = 1; return 0;
    case EC2D_NJ2: *slot = p.nj2; *ncomp = 1; *kind = 1; return 0;
    case EC2D_X1: *slot = p.x1; *ncomp = 1; *kind = 5; return 0;
    case EC2D_Y1: *slot = p.y1; *ncomp = 1; *kind = 5; return 0;
  }
  return 1;
}

int global_block(int case_id, int level, int mblock) {
  DLevel& L = R.lv[level];
  int nb = (int)R.cases[case_id].levels[level].size();
  if (mblock < 1 || mblock > nb) return -1;
  return L.case_first_block[case_id] + mblock - 1;
}

int init_default_state(DLevel& L, int level, std::vector<double>& img) {
  for (size_t b = 0; b < L.hb.size(); b++) {
    const BlockDesc& bd = L.hb[b];
    const HCase& c = R.cases[L.blk_case[b]];
    double* base = img.data() + bd.base;
    auto plane = [&](int slot) { return base + (int64_t)slot * bd.plane; };
    int nv = L.o.nvar;
    for (int64_t k = 0; k < bd.plane; k++) {
      plane(L.pm.U + 0)[k] = 1.0;
      plane(L.pm.U + 3)[k] = 1.0;
      if (nv == 5) plane(L.pm.U + 4)[k] = 1.0 / c.ctl.Re;
      if (nv == 6) plane(L.pm.U + 5)[k] = 1.0;
      plane(L.pm.Un + 0)[k] = 1.0;
      if (level > 0) plane(L.pm.Un + 3)[k] = 1.0;
    }
  }
  return 0;
}

}  // namespace

// ====================================================================== C-ABI
// mean time (ms) of `reps` launches of fn after one untimed launch
template <class Fn>
int time_launches(int reps, Fn fn, double* ms) {
  if (fn()) return 1;
  HIPCHK(hipEventRecord(R.ev0, R.s));
  for (int k = 0; k < reps; k++)
    if (fn()) return 1;
  HIPCHK(hipEventRecord(R.ev1, R.s));
  HIPCHK(hipEventSynchronize(R.ev1));
  float t = 0;
  HIPCHK(hipEventElapsedTime(&t, R.ev0, R.ev1));
  *ms = t / reps;
  return 0;
}


// ---- learned closure: host side ------------------------------------------
unsigned short host_f2bf(float x) {
  unsigned u;
  std::memcpy(&u, &x, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

// theta (CL_NPARAM floats) + norm (mean, inv_std, lo, hi) -> device network
int closure_upload(const float* theta, const float* norm) {
  ClosureNet N;
  std::memset(&N, 0, sizeof(N));
  for (int k = 0; k < CL_NF; k++)
    for (int n = 0; n < 64; n++) N.w1t[n * 16 + k] = host_f2bf(theta[CL_OW1 + k * 64 + n]);
  for (int k = 0; k < 64; k++)
    for (int n = 0; n < 64; n++) {
      N.w2t[n * 64 + k] = host_f2bf(theta[CL_OW2 + k * 64 + n]);
      N.w3t[n * 64 + k] = host_f2bf(theta[CL_OW3 + k * 64 + n]);
    }
  for (int k = 0; k < 64; k++) N.w4t[k] = host_f2bf(theta[CL_OW4 + k]);
  for (int n = 0; n < 64; n++) {
    N.b1[n] = theta[CL_OB1 + n];
    N.b2[n] = theta[CL_OB2 + n];
    N.b3[n] = theta[CL_OB3 + n];
  }
  N.b4 = theta[CL_OB4];
  for (int k = 0; k < CL_NF; k++) {
    N.mean[k] = norm[k];
    N.inv_std[k] = norm[CL_NF + k];
    N.lo[k] = norm[2 * CL_NF + k];
    N.hi[k] = norm[3 * CL_NF + k];
  }
  if (!R.clo.dnet) HIPCHK(hipMalloc(&R.clo.dnet, sizeof(ClosureNet)));
  if (!R.clo.theta) {
    HIPCHK(hipMalloc(&R.clo.theta, sizeof(float) * CL_NPARAM));
    HIPCHK(hipMalloc(&R.clo.m1, sizeof(float) * CL_NPARAM));
    HIPCHK(hipMalloc(&R.clo.m2, sizeof(float) * CL_NPARAM));
  }
  // R.s is a non-blocking stream: every copy of the closure's state goes
  // through it, so that it orders with the kernels that read and write it
  HIPCHK(hipMemcpyAsync(R.clo.dnet, &N, sizeof(N), hipMemcpyHostToDevice, R.s));
  HIPCHK(hipMemcpyAsync(R.clo.theta, theta, sizeof(float) * CL_NPARAM, hipMemcpyHostToDevice, R.s));
  HIPCHK(hipMemsetAsync(R.clo.m1, 0, sizeof(float) * CL_NPARAM, R.s));
  HIPCHK(hipMemsetAsync(R.clo.m2, 0, sizeof(float) * CL_NPARAM, R.s));
  HIPCHK(hipStreamSynchronize(R.s));
  R.clo.step = 0;
  std::memcpy(R.clo.norm, norm, sizeof(R.clo.norm));
  R.clo.ready = true;
  return 0;
}

// a small deterministic generator for the initial weights and the sample order
struct Lcg {
  unsigned long long s;
  explicit Lcg(unsigned seed) : s(0x9E3779B97F4A7C15ull ^ seed) {}
  unsigned next() {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    return (unsigned)(s >> 33);
  }
  double uniform() { return (next() + 0.5) / 2147483648.0; }  // (0, 1)
};

// autotune report text of an LU-SGS shape (0: the fused launch)
std::string shape_text(int sp, int u) {
  if (sp == 0) return "fused preparation and sweeps, " + std::to_string(u == 3 ? 3 : 6) + " producer waves";
  return "workgroups per block " + std::to_string(sp) + ", steps per set " + std::to_string(u);
}

extern "C" {

int ec2d_init(void) {
  if (R.inited) return 0;
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (n < 1) return fail("no HIP device");
  HIPCHK(hipSetDevice(0));
  {
    // the device address of the no-allocation scratch array on this device
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= SCRATCH_DEVICES) return fail("device index out of range for the scratch table");
    if (!scratch_addr[dev]) {
      void* p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p) return fail(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      scratch_addr[dev] = p;
    }
  }
  HIPCHK(hipStreamCreateWithFlags(&R.s, hipStreamNonBlocking));
  HIPCHK(hipEventCreate(&R.ev0));
  HIPCHK(hipEventCreate(&R.ev1));
  R.inited = true;
  return 0;
}

int ec2d_finalize(void) {
  if (!R.inited) return 0;
  hipStreamSynchronize(R.s);
  if (R.graph) hipGraphExecDestroy(R.graph);
  for (auto& L : R.lv) {
    hipFree(L.arena); hipFree(L.db); hipFree(L.ds); hipFree(L.dunits); hipFree(L.dnorm); hipFree(L.dt); hipFree(L.dperm);
    hipFree(L.dcase_tile0); hipFree(L.dncell); hipFree(L.dpartials); hipFree(L.dnorms); hipFree(L.drows);
    hipFree(L.dblocks_all); hipFree(L.sk); hipFree(L.dwall); hipFree(L.dcase_wall0); hipFree(L.dforces);
    hipFree(L.dgring);
    hipFree(L.d_ilines); hipFree(L.d_frows);
  }
  hipFree(R.clo.dnet); hipFree(R.clo.theta); hipFree(R.clo.m1); hipFree(R.clo.m2); hipFree(R.clo.part);
  hipFree(R.clo.losses); hipFree(R.clo.data); hipFree(R.clo.cells); hipFree(R.clo.idx);
  hipFree(R.dcps);
  hipFree(R.derr);
  hipFree(R.dsweep_err);
  hipFree(R.dactive);
  hipFree(R.dconv);
  hipFree(R.dcold);
  for (auto e : R.pev) hipEventDestroy(e);
  hipEventDestroy(R.ev0);
  hipEventDestroy(R.ev1);
  hipStreamDestroy(R.s);
  R = Runtime();
  return 0;
}

int ec2d_last_error(char* buf, int len) {
  if (!buf || len <= 0) return 1;
  std::snprintf(buf, (size_t)len, "%s", R.err.c_str());
  return 0;
}

int ec2d_set_tile_rows(int rows) {
  if (R.committed) return fail("tile shape must be set before ec2d_commit");
  if (rows != 1 && rows != 2 && rows != 4 && rows != 8) return fail("tile rows must be 1, 2, 4 or 8");
  R.tile_rows = rows;
  R.tile_cols = 64;
  return 0;
}

int ec2d_set_tile_shape(int cols, int rows) {
  if (R.committed) return fail("tile shape must be set before ec2d_commit");
  if (cols == 64) return ec2d_set_tile_rows(rows);
  if (cols != 32 || (rows != 8 && rows != 16)) return fail("tile shapes: 64 x 1, 2, 4 or 8, or 32 x 8 or 16");
  R.tile_cols = 32;
  R.tile_rows = rows;
  return 0;
}

int ec2d_case_create(const ec2d_control* ctl) {
  if (!R.inited && ec2d_init()) return -1;
  if (R.committed) { fail("cases must be created before ec2d_commit"); return -1; }
  if (ctl->Iflag_turbulence_model == 4 && (ctl->Nvar != 4 || ctl->If_viscous != 1)) {
    fail("turbulence model 4 (learned closure) needs Nvar = 4 and If_viscous = 1");
    return -1;
  }
  HCase c;
  c.ctl = *ctl;
  CaseParams& p = c.cp;
  p.gamma = ctl->gamma; p.Ma = ctl->Ma; p.Cv = ctl->Cv; p.Cp = ctl->Cp; p.Re = ctl->Re; p.Pr = ctl->Pr;
  p.Tsb = 110.4 / ctl->T_inf;
  p.CFL = ctl->CFL; p.dtmax = ctl->dtmax; p.dtmin = ctl->dtmin; p.dt_global = ctl->dt_global;
  p.P_outlet = ctl->P_outlet;
  // freestream of boundary_Farfield_inlet_outlet (sub_boundary.f90:470), with the
  // host C library's cos / sin as the reference uses
  p.u_inf = 1.0 * std::cos(ctl->AoA);
  p.v_inf = 1.0 * std::sin(ctl->AoA);
  p.p_inf = 1.0 / (ctl->gamma * ctl->Ma * ctl->Ma);
  p.Twall = ctl->Twall; p.vt_inf = ctl->vt_inf; p.Kt_inf = ctl->Kt_inf; p.Wt_inf = ctl->Wt_inf;
  p.p00 = 1.0 / (ctl->gamma * ctl->Ma * ctl->Ma);
  p.inv_Re = 1.0 / ctl->Re;
  p.If_viscous = ctl->If_viscous; p.Iflag_local_dt = ctl->Iflag_local_dt; p.Time_Method = ctl->Time_Method;
  c.num_mesh = std::max(1, std::min(3, ctl->Num_Mesh));
  c.levels.resize(c.num_mesh);
  if (!R.cases.empty()) {
    const ec2d_control& f = R.cases[0].ctl;
    if (f.Nvar != ctl->Nvar || f.Iflag_Scheme != ctl->Iflag_Scheme || f.Iflag_Flux != ctl->Iflag_Flux ||
        f.IFlag_Reconstruction != ctl->IFlag_Reconstruction || f.If_viscous != ctl->If_viscous ||
        f.Time_Method != ctl->Time_Method || f.Num_Mesh != ctl->Num_Mesh ||
        f.Iflag_turbulence_model != ctl->Iflag_turbulence_model) {
      fail("all cases of a batch must share equations, schemes, time method and multigrid depth");
      return -1;
    }
  }
  R.cases.push_back(c);
  return (int)R.cases.size() - 1;
}

int ec2d_define_block(int case_id, int nMesh, int mBlock, int nx, int ny, int nsub, const int* bc_msg, int fvm_fdm,
                      const double* x, const double* y, const double* x1, const double* y1, const double* vol,
                      const double* si, const double* sj, const double* ni1, const double* ni2, const double* nj1,
                      const double* nj2, const double* dw, const double* fd) {
  if (case_id < 0 || case_id >= (int)R.cases.size()) return fail("bad case id");
  HCase& c = R.cases[case_id];
  if (nMesh < 1 || nMesh > c.num_mesh) return fail("bad mesh level");
  if (R.committed) return fail("blocks must be defined before ec2d_commit");
  if (nx < 3 || ny < 3) return fail("block needs at least 2x2 cells");
  if (nsub < 1 || nsub > EC2D_MAXSUB) return fail("bad subface count");
  auto& lev = c.levels[nMesh - 1];
  if ((int)lev.size() < mBlock) lev.resize(mBlock);
  HBlock& b = lev[mBlock - 1];
  b.nx = nx; b.ny = ny; b.nsub = nsub; b.fdm = fvm_fdm;
  b.bc.assign(bc_msg, bc_msg + 9 * nsub);
  size_t nn = (size_t)(nx + 4) * (ny + 4), ng = (size_t)(nx + 3) * (ny + 3), np = (size_t)nx * ny;
  if (!x || !y) return fail("block nodes missing");
  b.x.assign(x, x + nn); b.y.assign(y, y + nn);
  b.geom_dev = x1 == nullptr;
  if (b.geom_dev) {
    // the device makes the rest at commit from the interior nodes
    if (y1 || vol || si || sj || ni1 || ni2 || nj1 || nj2)
      return fail("give all of x1, y1, vol, si, sj, ni1, ni2, nj1, nj2, or none (made on the device)");
    if (fvm_fdm) return fail("finite-difference blocks need their metrics from the caller");
  } else {
    if (!y1 || !vol || !si || !sj || !ni1 || !ni2 || !nj1 || !nj2) return fail("block geometry array missing");
    b.x1.assign(x1, x1 + ng); b.y1.assign(y1, y1 + ng);
    b.vol.assign(vol, vol + np); b.si.assign(si, si + np); b.sj.assign(sj, sj + np);
    b.ni1.assign(ni1, ni1 + np); b.ni2.assign(ni2, ni2 + np); b.nj1.assign(nj1, nj1 + np); b.nj2.assign(nj2, nj2 + np);
  }
  if (dw) b.dw.assign(dw, dw + (size_t)(nx + 1) * (ny + 1)); else b.dw.clear();
  if (fd && fvm_fdm) b.fd.assign(fd, fd + 5 * (size_t)(nx - 1) * (ny - 1)); else b.fd.clear();
  b.defined = true;
  return 0;
}

// Blocks defined without their geometry: ghost nodes, cell centres and face
// metrics from the interior nodes on the device (k_geometry), one launch per
// mesh level over every such block of the batch; the results go to the host
// copies the rest of ec2d_commit uploads and reads.
int device_geometry() {
  int nlev = 0;
  for (auto& c : R.cases) nlev = std::max(nlev, c.num_mesh);
  for (int l = 0; l < nlev; l++) {
    std::vector<HBlock*> blks;
    std::vector<int> first;  // launch index of each case's block 1 on this level (-1: none)
    for (auto& c : R.cases) {
      if (l >= c.num_mesh) { first.push_back(-1); continue; }
      int ndev = 0;
      for (auto& b : c.levels[l]) ndev += b.geom_dev ? 1 : 0;
      if (ndev && ndev != (int)c.levels[l].size()) return fail("a case mixes blocks with and without geometry");
      first.push_back(ndev ? (int)blks.size() : -1);
      if (ndev)
        for (auto& b : c.levels[l]) blks.push_back(&b);
    }
    if (blks.empty()) continue;
    std::vector<int64_t> noff, coff, moff;
    std::vector<int> dims, bc, sub0, nbblk;
    std::vector<double> xs, ys;
    int64_t nc = 0, nm = 0;
    size_t k = 0;
    for (size_t ci = 0; ci < R.cases.size(); ci++) {
      if (first[ci] < 0) continue;
      for (const HBlock& h : R.cases[ci].levels[l]) {
        noff.push_back((int64_t)xs.size());
        xs.insert(xs.end(), h.x.begin(), h.x.end());
        ys.insert(ys.end(), h.y.begin(), h.y.end());
        dims.push_back(h.nx);
        dims.push_back(h.ny);
        sub0.push_back((int)k);
        for (int s = 0; s < h.nsub; s++, k++) {
          const int* m = &h.bc[9 * s];
          bc.insert(bc.end(), m, m + 9);
          int nb = -1;
          if (m[6] > 0) {
            if (m[6] > (int)R.cases[ci].levels[l].size()) return fail("interface to a missing block");
            const HBlock& o = R.cases[ci].levels[l][m[6] - 1];
            if (m[7] < 1 || m[7] > o.nsub) return fail("interface to a missing subface");
            nb = first[ci] + m[6] - 1;
          }
          nbblk.push_back(nb);
        }
        coff.push_back(nc);
        nc += 2 * (int64_t)(h.nx + 3) * (h.ny + 3);
        moff.push_back(nm);
        nm += 7 * (int64_t)h.nx * h.ny;
      }
    }
    sub0.push_back((int)k);
    GeomArgs a;
    double *dx = nullptr, *dy = nullptr, *dc = nullptr, *dm = nullptr;
    int64_t *dno = nullptr, *dco = nullptr, *dmo = nullptr;
    int *ddims = nullptr, *dbc = nullptr, *dsub0 = nullptr, *dnb = nullptr;
    auto up = [&](auto** dst, const auto& v) -> int {
      using T = std::remove_reference_t<decltype(v[0])>;
      HIPCHK(hipMalloc(dst, sizeof(T) * std::max<size_t>(v.size(), 1)));
      if (!v.empty()) HIPCHK(hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
      return 0;
    };
    int rc = up(&dx, xs) || up(&dy, ys) || up(&dno, noff) || up(&dco, coff) || up(&dmo, moff) || up(&ddims, dims) ||
             up(&dbc, bc) || up(&dsub0, sub0) || up(&dnb, nbblk);
    if (!rc && (hipMalloc(&dc, sizeof(double) * std::max<int64_t>(nc, 1)) != hipSuccess ||
                hipMalloc(&dm, sizeof(double) * std::max<int64_t>(nm, 1)) != hipSuccess))
      rc = fail("hipMalloc failed");
    std::vector<double> cen((size_t)nc), met((size_t)nm);
    if (!rc) {
      a.x = dx; a.y = dy; a.noff = dno; a.dims = ddims; a.bc = dbc; a.sub0 = dsub0; a.nbblk = dnb;
      a.cen = dc; a.coff = dco; a.met = dm; a.moff = dmo;
      hipLaunchKernelGGL(k_geometry, dim3((unsigned)blks.size()), dim3(256), 0, R.s, a);
      if (hipGetLastError() != hipSuccess || hipStreamSynchronize(R.s) != hipSuccess ||
          hipMemcpy(xs.data(), dx, sizeof(double) * xs.size(), hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(ys.data(), dy, sizeof(double) * ys.size(), hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(cen.data(), dc, sizeof(double) * cen.size(), hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(met.data(), dm, sizeof(double) * met.size(), hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail("geometry kernel failed");
    }
    hipFree(dx); hipFree(dy); hipFree(dc); hipFree(dm); hipFree(dno); hipFree(dco); hipFree(dmo);
    hipFree(ddims); hipFree(dbc); hipFree(dsub0); hipFree(dnb);
    if (rc) return rc;
    for (size_t b = 0; b < blks.size(); b++) {
      HBlock& h = *blks[b];
      const size_t nn = h.x.size(), ng = (size_t)(h.nx + 3) * (h.ny + 3), np = (size_t)h.nx * h.ny;
      std::copy(xs.begin() + noff[b], xs.begin() + noff[b] + nn, h.x.begin());
      std::copy(ys.begin() + noff[b], ys.begin() + noff[b] + nn, h.y.begin());
      const double* c0 = cen.data() + coff[b];
      h.x1.assign(c0, c0 + ng);
      h.y1.assign(c0 + ng, c0 + 2 * ng);
      const double* m0 = met.data() + moff[b];
      std::vector<double>* dst[7] = {&h.vol, &h.si, &h.sj, &h.ni1, &h.ni2, &h.nj1, &h.nj2};
      for (int q = 0; q < 7; q++) dst[q]->assign(m0 + q * np, m0 + (q + 1) * np);
      h.geom_dev = false;
    }
  }
  return 0;
}

bool wall_distance_needed(const ec2d_control& c) {
  const int t = c.Iflag_turbulence_model;
  return t == 2 || t == 3 || t == 4;
}

// comput_dw on the device (ec2d_dw.hip) for the finest-level blocks whose
// caller gave no dw: node coordinates and the wall nodes of each case go up
// once, k_dw_nodes / k_dw_cells write the dw plane.
int device_wall_distance(DLevel& L) {
  std::vector<int> blk, wall0;
  std::vector<int64_t> off;
  std::vector<double> xy, wx, wy;
  for (size_t b = 0; b < L.hb.size(); b++) {
    const HCase& cs = R.cases[L.blk_case[b]];
    const HBlock& hb = cs.levels[0][L.blk_local[b]];
    if (!hb.dw.empty() || !wall_distance_needed(cs.ctl)) continue;
    // the wall nodes of the block's case (every block, every wall subface)
    const int w0 = (int)wx.size();
    for (const HBlock& b1 : cs.levels[0])
      for (int k = 0; k < b1.nsub; k++) {
        const int* s = &b1.bc[9 * k];
        if (s[6] != -10) continue;
        for (int j1 = s[4]; j1 <= s[5]; j1++)
          for (int i1 = s[2]; i1 <= s[3]; i1++) {
            wx.push_back(b1.x[ni_(b1.nx, i1, j1)]);
            wy.push_back(b1.y[ni_(b1.nx, i1, j1)]);
          }
      }
    blk.push_back((int)b);
    wall0.push_back(w0);
    wall0.push_back((int)wx.size());
    off.push_back((int64_t)xy.size());
    const int nn = hb.nx * hb.ny;
    xy.resize(xy.size() + 2 * (size_t)nn);
    double* px = xy.data() + off.back();
    for (int j = 1; j <= hb.ny; j++)
      for (int i = 1; i <= hb.nx; i++) {
        px[(j - 1) * hb.nx + (i - 1)] = hb.x[ni_(hb.nx, i, j)];
        px[nn + (j - 1) * hb.nx + (i - 1)] = hb.y[ni_(hb.nx, i, j)];
      }
  }
  if (blk.empty()) return 0;
  int maxn = 0;
  for (int b : blk) maxn = std::max(maxn, L.hb[b].nx * L.hb[b].ny);
  DwArgs a;
  double *d_xy = nullptr, *d_wx = nullptr, *d_wy = nullptr, *d_node = nullptr;
  int64_t* d_off = nullptr;
  int *d_blk = nullptr, *d_w0 = nullptr;
  auto up = [&](auto** dst, const auto& v) -> int {
    using T = std::remove_reference_t<decltype(v[0])>;
    HIPCHK(hipMalloc(dst, sizeof(T) * std::max<size_t>(v.size(), 1)));
    if (!v.empty()) HIPCHK(hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    return 0;
  };
  int rc = up(&d_xy, xy) || up(&d_wx, wx) || up(&d_wy, wy) || up(&d_off, off) || up(&d_blk, blk) || up(&d_w0, wall0);
  if (!rc) rc = hipMalloc(&d_node, sizeof(double) * std::max<size_t>(xy.size() / 2, 1)) == hipSuccess ? 0 : fail("hipMalloc failed");
  if (!rc) {
    a.nx_y = d_xy; a.off = d_off; a.blk = d_blk; a.wall0 = d_w0; a.wx = d_wx; a.wy = d_wy;
    a.blocks = L.db; a.arena = L.arena; a.dw = L.pm.dw; a.node_dw = d_node;
    const dim3 g((unsigned)((maxn + 255) / 256), (unsigned)blk.size());
    hipLaunchKernelGGL(k_dw_nodes, g, dim3(256), 0, R.s, a);
    hipLaunchKernelGGL(k_dw_cells, g, dim3(256), 0, R.s, a);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(R.s) != hipSuccess) rc = fail("wall distance kernels failed");
  }
  hipFree(d_xy); hipFree(d_wx); hipFree(d_wy); hipFree(d_off); hipFree(d_blk); hipFree(d_w0); hipFree(d_node);
  return rc;
}

int ec2d_commit(void) {
  if (!R.inited) return fail("ec2d_init has not been called");
  if (R.cases.empty()) return fail("no case defined");
  if (R.committed) return fail("already committed");
  R.num_mesh = R.cases[0].num_mesh;
  if (R.cases[0].ctl.Time_Method == -1) {
    // Dual_time_LU_SGS: one mesh (opencfd_ec2d_v1.5.4.f90:216-221); the inner
    // loop ends per case, so a batch holds one case; pre-steps would run
    // dual-time steps before Init_flow sets Un and Un1 (sub_init.f90:396)
    if (R.num_mesh != 1) return fail("dual-time LU-SGS needs Num_Mesh = 1");
    if (R.cases.size() != 1) return fail("dual-time LU-SGS runs one case per batch");
    if (R.cases[0].ctl.Pre_Step_Mesh[0] != 0) return fail("dual-time LU-SGS with pre-steps is not supported");
    if (R.cases[0].ctl.Nstep_Inner_Limit < 1) return fail("dual-time LU-SGS needs Nstep_Inner_Limit >= 1");
  }
  for (auto& c : R.cases)
    for (int l = 0; l < c.num_mesh; l++) {
      if (c.levels[l].empty()) return fail("mesh level without blocks");
      for (auto& b : c.levels[l])
        if (!b.defined) return fail("undefined block");
    }
  // geometry the caller left to the device (before anything reads it)
  if (device_geometry()) return 1;
  std::vector<CaseParams> cps;
  for (auto& c : R.cases) cps.push_back(c.cp);
  HIPCHK(hipMalloc(&R.dcps, sizeof(CaseParams) * cps.size()));
  HIPCHK(hipMemcpy(R.dcps, cps.data(), sizeof(CaseParams) * cps.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&R.derr, sizeof(unsigned long long)));
  HIPCHK(hipMalloc(&R.dsweep_err, sizeof(int)));
  HIPCHK(hipMemset(R.dsweep_err, 0, sizeof(int)));
  unsigned long long none = ~0ull;
  HIPCHK(hipMemcpy(R.derr, &none, sizeof(none), hipMemcpyHostToDevice));
  const ec2d_control& c0 = R.cases[0].ctl;
  R.lv.resize(R.num_mesh);
  for (int l = 0; l < R.num_mesh; l++) {
    DLevel& L = R.lv[l];
    int nv = (l == 0) ? c0.Nvar : 4;
    bool fdm_any = false;
    for (auto& c : R.cases)
      for (auto& b : c.levels[l]) fdm_any |= (b.fdm != 0);
    L.o.nvar = nv;
    L.o.nvar_un = nv;
    L.o.scheme = (l == 0) ? c0.Iflag_Scheme : 0;
    L.o.global_scheme = c0.Iflag_Scheme;
    L.o.flux = c0.Iflag_Flux;
    L.o.recon = c0.IFlag_Reconstruction;
    L.o.turb = (l == 0) ? c0.Iflag_turbulence_model : 0;
    L.fdm = fdm_any;
    const int turb0 = c0.Iflag_turbulence_model;
    L.pm = make_planes(nv, nv, fdm_any, l == 0 && (turb0 == 2 || turb0 == 4), l == 0 && c0.Time_Method == -1,
                       l == 0 && turb0 == 1 && c0.If_viscous == 1);
    // blocks
    int64_t off = 0;
    std::vector<SubfaceDesc> subs;
    std::vector<double> norms;
    int max_rows = 1;
    for (size_t ci = 0; ci < R.cases.size(); ci++) {
      auto& c = R.cases[ci];
      L.case_first_block.push_back((int)L.hb.size());
      long long cells = 0;
      for (size_t bi = 0; bi < c.levels[l].size(); bi++) {
        auto& hb = c.levels[l][bi];
        BlockDesc d{};
        d.nx = hb.nx; d.ny = hb.ny; d.P = pitch_of(hb.nx); d.case_id = (int)ci;
        d.plane = (int64_t)(hb.ny + 3) * d.P;
        d.base = off;
        d.fdm = hb.fdm;
        d.bc_off = (int)subs.size();
        d.nsub = hb.nsub;
        off += d.plane * L.pm.nplanes;
        off = (off + 31) / 32 * 32;
        L.hb.push_back(d);
        L.blk_case.push_back((int)ci);
        L.blk_local.push_back((int)bi);
        cells += (long long)(hb.nx - 1) * (hb.ny - 1);
        max_rows = std::max(max_rows, hb.ny - 1);
        for (int k = 0; k < hb.nsub; k++) {
          const int* s = &hb.bc[9 * k];
          SubfaceDesc sf{};
          sf.face = s[1]; sf.ist = s[2]; sf.iend = s[3]; sf.jst = s[4]; sf.jend = s[5];
          // every kernel index derived from a subface stays inside the block's
          // ghosted arrays only for a well-formed subface: check it here
          {
            bool ok = sf.ist >= 1 && sf.iend <= hb.nx && sf.jst >= 1 && sf.jend <= hb.ny && sf.ist <= sf.iend &&
                      sf.jst <= sf.jend;
            if (sf.face == 1) ok = ok && sf.ist == 1 && sf.iend == 1;
            else if (sf.face == 3) ok = ok && sf.ist == hb.nx && sf.iend == hb.nx;
            else if (sf.face == 2) ok = ok && sf.jst == 1 && sf.jend == 1;
            else if (sf.face == 4) ok = ok && sf.jst == hb.ny && sf.jend == hb.ny;
            else ok = false;
            if (!ok) return fail("malformed subface in the boundary table");
          }
          sf.neighb = s[6]; sf.nsub_nb = s[7]; sf.orient = s[8];
          sf.nb_block = -1;
          sf.norm_off = -1;
          if (sf.neighb > 0) {
            if (sf.neighb > (int)c.levels[l].size()) return fail("interface to a missing block");
            auto& nb = c.levels[l][sf.neighb - 1];
            if (sf.nsub_nb < 1 || sf.nsub_nb > nb.nsub) return fail("interface to a missing subface");
            const int* s1 = &nb.bc[9 * (sf.nsub_nb - 1)];
            sf.nb_block = L.case_first_block[ci] + sf.neighb - 1;
            sf.nb_face = s1[1]; sf.nb_ist = s1[2]; sf.nb_iend = s1[3]; sf.nb_jst = s1[4]; sf.nb_jend = s1[5];
            int n1 = sf.iend - sf.ist + sf.jend - sf.jst;
            int n1b = s1[3] - s1[2] + s1[5] - s1[4];
            if (n1 != n1b) return fail("interface subfaces of different length");
            bool okb = s1[2] >= 1 && s1[3] <= nb.nx && s1[4] >= 1 && s1[5] <= nb.ny;
            if (s1[1] == 1) okb = okb && s1[2] == 1;
            else if (s1[1] == 3) okb = okb && s1[3] == nb.nx && nb.nx >= EC2D_LAP + 1;
            else if (s1[1] == 2) okb = okb && s1[4] == 1;
            else if (s1[1] == 4) okb = okb && s1[5] == nb.ny && nb.ny >= EC2D_LAP + 1;
            else okb = false;
            if (!okb) return fail("interface to a malformed subface");
          } else if (sf.neighb == -20 || sf.neighb == -40 || sf.neighb == -50 || sf.neighb == -30) {
            // outward normals per boundary point (sub_boundary.f90:496-508 far field,
            // :411-420 symmetry), from the nodes with the reference's operations
            sf.norm_off = (int)(norms.size() / 2);
            int nx = hb.nx;
            for (int j = sf.jst; j <= sf.jend; j++)
              for (int i = sf.ist; i <= sf.iend; i++) {
                double dx, dy, sl, n1, n2;
                if (sf.face == 1 || sf.face == 3) {
                  dx = hb.x[ni_(nx, i, j + 1)] - hb.x[ni_(nx, i, j)];
                  dy = hb.y[ni_(nx, i, j + 1)] - hb.y[ni_(nx, i, j)];
                } else {
                  dx = hb.x[ni_(nx, i + 1, j)] - hb.x[ni_(nx, i, j)];
                  dy = hb.y[ni_(nx, i + 1, j)] - hb.y[ni_(nx, i, j)];
                }
                sl = std::sqrt(dx * dx + dy * dy);
                if (sf.neighb == -30) {
                  if (sf.face == 1 || sf.face == 3) { n1 = dy / sl; n2 = -dx / sl; }
                  else { n1 = -dy / sl; n2 = dx / sl; }
                } else {
                  if (sf.face == 1) { n1 = -dy / sl; n2 = dx / sl; }
                  else if (sf.face == 3) { n1 = dy / sl; n2 = -dx / sl; }
                  else if (sf.face == 2) { n1 = dy / sl; n2 = -dx / sl; }
                  else { n1 = -dy / sl; n2 = dx / sl; }
                }
                norms.push_back(n1);
                norms.push_back(n2);
              }
          }
          subs.push_back(sf);
        }
      }
      L.case_cells.push_back(cells);
    }
    L.case_first_block.push_back((int)L.hb.size());
    L.arena_n = off;
    // two lanes per cell when one wave holds every block's rows that way;
    // otherwise one lane per row (fewer waves, fewer barriers per step)
    if (lusgs_layout(L, max_rows <= 32 ? 2 : 1, 1, false)) return 1;
    // host image of the arena
    std::vector<double> img((size_t)off, 0.0);
    init_default_state(L, l, img);
    for (size_t b = 0; b < L.hb.size(); b++) {
      const BlockDesc& bd = L.hb[b];
      const HBlock& hb = R.cases[L.blk_case[b]].levels[l][L.blk_local[b]];
      double* base = img.data() + bd.base;
      auto plane = [&](int slot) { return base + (int64_t)slot * bd.plane; };
      to_plane(bd, plane(L.pm.vol), hb.vol.data(), 1, 1, 0);
      to_plane(bd, plane(L.pm.si), hb.si.data(), 1, 1, 0);
      to_plane(bd, plane(L.pm.sj), hb.sj.data(), 1, 1, 0);
      to_plane(bd, plane(L.pm.ni1), hb.ni1.data(), 1, 1, 0);
      to_plane(bd, plane(L.pm.ni2), hb.ni2.data(), 1, 1, 0);
      to_plane(bd, plane(L.pm.nj1), hb.nj1.data(), 1, 1, 0);
      to_plane(bd, plane(L.pm.nj2), hb.nj2.data(), 1, 1, 0);
      to_plane(bd, plane(L.pm.x1), hb.x1.data(), 5, 1, 0);
      to_plane(bd, plane(L.pm.y1), hb.y1.data(), 5, 1, 0);
      if (!hb.dw.empty()) to_plane(bd, plane(L.pm.dw), hb.dw.data(), 3, 1, 0);
      else if (l == 0 && wall_distance_needed(R.cases[L.blk_case[b]].ctl)) {
        // computed on the device below; B%dw(:,:) = 1 around it (sub_turbulence_SST.f90:288-293)
        std::vector<double> one((size_t)(hb.nx + 1) * (hb.ny + 1), 1.0);
        to_plane(bd, plane(L.pm.dw), one.data(), 3, 1, 0);
      }
      if (!hb.fd.empty() && L.pm.Akx >= 0) {
        size_t n = (size_t)(hb.nx - 1) * (hb.ny - 1);
        int slots[5] = {L.pm.Akx, L.pm.Aky, L.pm.Aix, L.pm.Aiy, L.pm.Ajac};
        for (int k = 0; k < 5; k++) to_plane(bd, plane(slots[k]), hb.fd.data() + k * n, 2, 1, 0);
      }
      // wall ghosts of Amut_boundary (sub_turbulence_SA.f90:194-218)
      for (int k = 0; k < hb.nsub; k++) {
        const int* s = &hb.bc[9 * k];
        if (s[6] != -10) continue;
        int face = s[1], ib = s[2], ie = s[3], jb = s[4], je = s[5];
        double* fl = plane(L.pm.flag);
        if (face == 1) for (int j = jb; j <= je - 1; j++) fl[pidx(bd, 0, j)] = 1.0;
        else if (face == 3) for (int j = jb; j <= je - 1; j++) fl[pidx(bd, hb.nx, j)] = 1.0;
        else if (face == 2) for (int i = ib; i <= ie - 1; i++) fl[pidx(bd, i, 0)] = 1.0;
        else for (int i = ib; i <= ie - 1; i++) fl[pidx(bd, i, hb.ny)] = 1.0;
      }
    }
    HIPCHK(hipMalloc(&L.arena, sizeof(double) * (size_t)std::max<int64_t>(off, 1)));
    HIPCHK(hipMemcpy(L.arena, img.data(), sizeof(double) * (size_t)off, hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&L.db, sizeof(BlockDesc) * L.hb.size()));
    HIPCHK(hipMemcpy(L.db, L.hb.data(), sizeof(BlockDesc) * L.hb.size(), hipMemcpyHostToDevice));
    // after the block table: the wall-distance kernels read it
    if (l == 0 && device_wall_distance(L)) return 1;
    L.bl_walls = L.bl_lines = 0;
    L.max_plane = 0;
    for (auto& b : L.hb) {
      L.max_plane = std::max(L.max_plane, b.plane);
      int nw = 0;
      for (int k = 0; k < b.nsub; k++) {
        const SubfaceDesc& sf = subs[b.bc_off + k];
        if (sf.neighb != -10) continue;
        nw++;
        L.bl_lines = std::max(L.bl_lines, (sf.face == 1 || sf.face == 3) ? sf.jend - sf.jst : sf.iend - sf.ist);
      }
      L.bl_walls = std::max(L.bl_walls, nw);
    }
    std::vector<int> units(std::max<size_t>(subs.size(), 1), 0);
    for (auto& b : L.hb) ghost_phases(b, subs.data() + b.bc_off, units.data() + b.bc_off);
    HIPCHK(hipMalloc(&L.dunits, sizeof(int) * units.size()));
    HIPCHK(hipMemcpy(L.dunits, units.data(), sizeof(int) * units.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&L.ds, sizeof(SubfaceDesc) * std::max<size_t>(subs.size(), 1)));
    HIPCHK(hipMemcpy(L.ds, subs.data(), sizeof(SubfaceDesc) * subs.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&L.dnorm, sizeof(double) * std::max<size_t>(norms.size(), 2)));
    if (!norms.empty()) HIPCHK(hipMemcpy(L.dnorm, norms.data(), sizeof(double) * norms.size(), hipMemcpyHostToDevice));
    L.tile_rows = R.tile_rows;
    L.tile_cols = R.tile_cols;
    if (build_tiles(L)) return 1;
    std::vector<double> ncell;
    for (auto v : L.case_cells) ncell.push_back(1.0 * (double)v);
    HIPCHK(hipMalloc(&L.dncell, sizeof(double) * ncell.size()));
    HIPCHK(hipMemcpy(L.dncell, ncell.data(), sizeof(double) * ncell.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&L.dnorms, sizeof(double) * 12 * R.cases.size()));
    std::vector<double> z(12 * R.cases.size(), 0.0);
    HIPCHK(hipMemcpy(L.dnorms, z.data(), sizeof(double) * z.size(), hipMemcpyHostToDevice));
    // rows table (j = 0..ny) and block list
    std::vector<int2> rows;
    std::vector<int> all;
    for (size_t b = 0; b < L.hb.size(); b++) {
      all.push_back((int)b);
      for (int j = 0; j <= L.hb[b].ny; j++) rows.push_back(make_int2((int)b, j));
    }
    L.nrows = (int)rows.size();
    if (L.fdm) {
      std::vector<int2> il, jl, fr;
      for (size_t b = 0; b < L.hb.size(); b++) {
        if (!L.hb[b].fdm) continue;
        int nxp = L.hb[b].nx - 1, nyp = L.hb[b].ny - 1;
        if (nxp > FDM_MAXN || nyp > FDM_MAXN) return fail("finite-difference block longer than 992 cells");
        if (nxp < 2 * LFDM + 2 || nyp < 2 * LFDM + 2) return fail("finite-difference block too small");
        for (int j = LFDM; j <= nyp - LFDM; j++) il.push_back(make_int2((int)b, j));
        for (int i = LFDM; i <= nxp - LFDM; i++) jl.push_back(make_int2((int)b, i));
        for (int j = 1; j <= nyp; j++) fr.push_back(make_int2((int)b, j));
      }
      L.n_ilines = (int)il.size();
      L.n_jlines = (int)jl.size();
      L.n_frows = (int)fr.size();
      il.insert(il.end(), jl.begin(), jl.end());  // i-lines, then j-lines: one launch covers both
      HIPCHK(hipMalloc(&L.d_ilines, sizeof(int2) * il.size()));
      HIPCHK(hipMemcpy(L.d_ilines, il.data(), sizeof(int2) * il.size(), hipMemcpyHostToDevice));
      HIPCHK(hipMalloc(&L.d_frows, sizeof(int2) * fr.size()));
      HIPCHK(hipMemcpy(L.d_frows, fr.data(), sizeof(int2) * fr.size(), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMalloc(&L.drows, sizeof(int2) * rows.size()));
    HIPCHK(hipMemcpy(L.drows, rows.data(), sizeof(int2) * rows.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&L.dblocks_all, sizeof(int) * all.size()));
    HIPCHK(hipMemcpy(L.dblocks_all, all.data(), sizeof(int) * all.size(), hipMemcpyHostToDevice));
    // wall faces for the force integration (finest level only)
    if (l == 0) {
      std::vector<int4> wf;
      std::vector<int> w0;
      for (size_t ci = 0; ci < R.cases.size(); ci++) {
        w0.push_back((int)wf.size());
        for (int b = L.case_first_block[ci]; b < L.case_first_block[ci + 1]; b++) {
          const HBlock& hb = R.cases[ci].levels[0][L.blk_local[b]];
          for (int k = 0; k < hb.nsub; k++) {
            const int* sb = &hb.bc[9 * k];
            if (sb[6] != -10) continue;
            const int face = sb[1];
            if (face == 1 || face == 3)
              for (int j = sb[4]; j < sb[5]; j++) wf.push_back(make_int4(b, face, face == 1 ? 1 : hb.nx, j));
            else
              for (int i = sb[2]; i < sb[3]; i++) wf.push_back(make_int4(b, face, i, face == 2 ? 1 : hb.ny));
          }
        }
      }
      w0.push_back((int)wf.size());
      HIPCHK(hipMalloc(&L.dwall, sizeof(int4) * std::max<size_t>(wf.size(), 1)));
      if (!wf.empty()) HIPCHK(hipMemcpy(L.dwall, wf.data(), sizeof(int4) * wf.size(), hipMemcpyHostToDevice));
      HIPCHK(hipMalloc(&L.dcase_wall0, sizeof(int) * w0.size()));
      HIPCHK(hipMemcpy(L.dcase_wall0, w0.data(), sizeof(int) * w0.size(), hipMemcpyHostToDevice));
      HIPCHK(hipMalloc(&L.dforces, sizeof(double) * 4 * R.cases.size()));
    }
  }
  R.committed = true;
  return 0;
}

int ec2d_put_array(int case_id, int nMesh, int mBlock, int which, const double* src) {
  if (check_ready()) return 1;
  if (case_id < 0 || case_id >= (int)R.cases.size() || nMesh < 1 || nMesh > R.num_mesh) return fail("bad case / mesh");
  int gb = global_block(case_id, nMesh - 1, mBlock);
  if (gb < 0) return fail("bad block");
  if ((which >= EC2D_VOL && which <= EC2D_NJ2) || which == EC2D_X1 || which == EC2D_Y1)
    return fail("the geometry is fixed at ec2d_commit (read-only selector)");
  DLevel& L = R.lv[nMesh - 1];
  int slot, nc, kind;
  if (array_map(L, which, &slot, &nc, &kind)) return fail("bad array selector");
  const BlockDesc& bd = L.hb[gb];
  std::vector<double> plane((size_t)bd.plane);
  for (int m = 0; m < nc; m++) {
    double* dptr = L.arena + bd.base + (int64_t)(slot + m) * bd.plane;
    HIPCHK(hipMemcpyAsync(plane.data(), dptr, sizeof(double) * plane.size(), hipMemcpyDeviceToHost, R.s));
    HIPCHK(hipStreamSynchronize(R.s));
    to_plane(bd, plane.data(), src, kind == 5 ? 5 : kind, nc, m);
    HIPCHK(hipMemcpyAsync(dptr, plane.data(), sizeof(double) * plane.size(), hipMemcpyHostToDevice, R.s));
    HIPCHK(hipStreamSynchronize(R.s));
  }
  if (which == EC2D_U) R.clip_pending[nMesh - 1] = false;
  return 0;
}

int ec2d_get_array(int case_id, int nMesh, int mBlock, int which, double* dst) {
  if (check_ready(which != EC2D_U)) return 1;
  if (case_id < 0 || case_id >= (int)R.cases.size() || nMesh < 1 || nMesh > R.num_mesh) return fail("bad case / mesh");
  int gb = global_block(case_id, nMesh - 1, mBlock);
  if (gb < 0) return fail("bad block");
  if (which == EC2D_U && flush_clip(nMesh - 1)) return 1;
  DLevel& L = R.lv[nMesh - 1];
  int slot, nc, kind;
  if (array_map(L, which, &slot, &nc, &kind)) return fail("bad array selector");
  const BlockDesc& bd = L.hb[gb];
  std::vector<double> plane((size_t)bd.plane);
  for (int m = 0; m < nc; m++) {
    const double* dptr = L.arena + bd.base + (int64_t)(slot + m) * bd.plane;
    HIPCHK(hipMemcpyAsync(plane.data(), dptr, sizeof(double) * plane.size(), hipMemcpyDeviceToHost, R.s));
    HIPCHK(hipStreamSynchronize(R.s));
    from_plane(bd, plane.data(), dst, kind, nc, m);
  }
  // the arrays of a run stopped by T < 1e-5 stay readable; a timed-out sweep fails
  return device_errors(false);
}

int ec2d_init_flow(void) {
  if (check_ready()) return 1;
  HIPCHK(hipStreamSynchronize(R.s));  // the plane fills below copy on the null stream
  // a fresh start: the rings go back to the sentinel if a sweep timed out,
  // then both error words are cleared
  if (device_errors(false) == 1) R.err.clear();
  if (reset_device_errors()) return 1;
  const ec2d_control& c0 = R.cases[0].ctl;
  R.residual_calls = 0;
  R.dual_steps = 0;
  R.dual_kt = 0;
  // Init_flow ends with Un = Un1 = U for dual-time stepping (sub_init.f90:396-409)
  auto dual_init = [&]() { return c0.Time_Method == -1 ? launch_pw(0, PW_DUAL_INIT) : 0; };
  if (c0.Iflag_init != 0) {
    if (launch_ghost(0, 1, 1)) return 1;
    return dual_init();
  }
  // init_flow_zero (sub_init.f90:423-495)
  int nm = R.num_mesh;
  for (size_t ci = 0; ci < R.cases.size(); ci++) {
    const HCase& c = R.cases[ci];
    const ec2d_control& k = c.ctl;
    double d0 = 1.0, u0 = 1.0 * std::cos(k.AoA), v0 = 1.0 * std::sin(k.AoA);
    double p0 = 1.0 / (k.gamma * k.Ma * k.Ma);
    for (int l = 0; l < nm; l++) {
      DLevel& L = R.lv[l];
      for (int b = L.case_first_block[ci]; b < L.case_first_block[ci + 1]; b++) {
        const BlockDesc& bd = L.hb[b];
        int nv = L.o.nvar;
        std::vector<double> plane((size_t)bd.plane);
        auto fill = [&](int slot, double v) -> int {
          double* dptr = L.arena + bd.base + (int64_t)slot * bd.plane;
          HIPCHK(hipMemcpy(plane.data(), dptr, sizeof(double) * plane.size(), hipMemcpyDeviceToHost));
          for (int j = 1; j <= bd.ny - 1; j++)
            for (int i = 1; i <= bd.nx - 1; i++) plane[pidx(bd, i, j)] = v;
          HIPCHK(hipMemcpy(dptr, plane.data(), sizeof(double) * plane.size(), hipMemcpyHostToDevice));
          return 0;
        };
        if (l == 0 && nv == 5 && fill(L.pm.U + 4, k.vt_inf / k.Re)) return 1;
        if (l == 0 && nv == 6 && (fill(L.pm.U + 4, 10.0 * k.Kt_inf) || fill(L.pm.U + 5, k.Wt_inf))) return 1;
        if (l == nm - 1) {
          if (fill(L.pm.U + 0, d0) || fill(L.pm.U + 1, d0 * u0) || fill(L.pm.U + 2, d0 * v0) ||
              fill(L.pm.U + 3, p0 / (k.gamma - 1.0) + 0.5 * d0 * (u0 * u0 + v0 * v0)))
            return 1;
        }
      }
    }
  }
  if (launch_ghost(nm - 1, 1, 1)) return 1;
  for (int n = nm; n >= 1; n--) {
    for (int st = 0; st < c0.Pre_Step_Mesh[n - 1]; st++)
      if (time_advance(n - 1)) return 1;
    if (n > 1) {
      if (flush_clip(n - 1)) return 1;
      if (launch_xfer(n - 2, n - 1, 1, false)) return 1;
      if (launch_ghost(n - 2, 1, 1)) return 1;
    }
  }
  return dual_init();
}

int ec2d_dual_inner(int* kt_in) {
  if (check_ready()) return 1;
  if (kt_in) *kt_in = R.dual_kt;
  return 0;
}

int ec2d_comput_residual_one_mesh(int nMesh) {
  if (check_ready() || nMesh < 1 || nMesh > R.num_mesh) return fail("bad mesh level");
  return residual(nMesh - 1);
}

int ec2d_du_lu_sgs_2d(int nMesh, int mBlock, double alfa1) {
  if (check_ready() || nMesh < 1 || nMesh > R.num_mesh) return fail("bad mesh level");
  DLevel& L = R.lv[nMesh - 1];
  if (flush_clip(nMesh - 1)) return 1;
  if (mBlock == 0) return launch_lusgs(nMesh - 1, L.dblocks_all, (int)L.hb.size(), alfa1, 0);
  if (R.cases.size() != 1) return fail("a single block can only be selected in a one-case batch");
  if (mBlock < 1 || mBlock > (int)L.hb.size()) return fail("bad block");
  return launch_lusgs(nMesh - 1, L.dblocks_all + (mBlock - 1), 1, alfa1, 0);
}

int ec2d_du_lu_sgs_2d_noalloc(int nMesh, int mBlock, double alfa1) {
  if (check_ready() || nMesh < 1 || nMesh > R.num_mesh) return fail("bad mesh level");
  if (R.profile) return fail("ec2d_du_lu_sgs_2d_noalloc: profiling creates events; switch it off first");
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  void* base = (dev >= 0 && dev < SCRATCH_DEVICES) ? scratch_addr[dev] : nullptr;
  if (!base) return fail("ec2d_du_lu_sgs_2d_noalloc: no scratch address for this device (ec2d_init sets it)");
  DLevel& L = R.lv[nMesh - 1];
  if (mBlock != 0) {
    if (R.cases.size() != 1) return fail("a single block can only be selected in a one-case batch");
    if (mBlock < 1 || mBlock > (int)L.hb.size()) return fail("bad block");
  }
  const size_t rec_bytes = ((size_t)L.sk_total * 8 + 255) / 256 * 256;
  const int64_t nring = L.dgring ? (int64_t)L.hb.size() * (L.lusgs_split - 1) * 2 * L.gstride : 0;
  const size_t need = rec_bytes + (size_t)nring * 8;
  if (need > SCRATCH_BYTES) {
    char msg[200];
    std::snprintf(msg, sizeof(msg),
                  "ec2d_du_lu_sgs_2d_noalloc: level %d needs %.1f MiB of LU-SGS records, more than the %zu MiB "
                  "scratch (use ec2d_du_lu_sgs_2d)", nMesh, need / 1048576.0, SCRATCH_BYTES >> 20);
    return fail(msg);
  }
  double* sk = static_cast<double*>(base);
  unsigned long long* gr = nring ? reinterpret_cast<unsigned long long*>(static_cast<char*>(base) + rec_bytes) : nullptr;
  // the array keeps the last call's contents: records start at zero (padding
  // positions are read, never written), ring slots at the sentinel
  hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((L.sk_total + 255) / 256)), dim3(256), 0, R.s,
                     reinterpret_cast<unsigned long long*>(sk), L.sk_total, 0ull);
  if (nring)
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((nring + 255) / 256)), dim3(256), 0, R.s, gr, nring, GSENT);
  HIPCHK(hipGetLastError());
  if (flush_clip(nMesh - 1)) return 1;
  if (mBlock == 0) return launch_lusgs(nMesh - 1, L.dblocks_all, (int)L.hb.size(), alfa1, 0, sk, gr);
  return launch_lusgs(nMesh - 1, L.dblocks_all + (mBlock - 1), 1, alfa1, 0, sk, gr);
}

int ec2d_update_u(int nMesh) {
  if (check_ready() || nMesh < 1 || nMesh > R.num_mesh) return fail("bad mesh level");
  if (flush_clip(nMesh - 1)) return 1;
  return launch_pw(nMesh - 1, PW_UPDATE);
}

int ec2d_boundary_condition_onemesh(int nMesh) {
  if (check_ready() || nMesh < 1 || nMesh > R.num_mesh) return fail("bad mesh level");
  if (flush_clip(nMesh - 1)) return 1;
  return launch_ghost(nMesh - 1, 1, 0);
}

int ec2d_update_buffer_onemesh(int nMesh) {
  if (check_ready() || nMesh < 1 || nMesh > R.num_mesh) return fail("bad mesh level");
  if (flush_clip(nMesh - 1)) return 1;
  return launch_ghost(nMesh - 1, 0, 1);
}

int ec2d_force_vt_kw(int nMesh) {
  if (check_ready() || nMesh < 1 || nMesh > R.num_mesh) return fail("bad mesh level");
  R.clip_pending[nMesh - 1] = false;
  if (R.lv[nMesh - 1].o.nvar < 5) return 0;
  return launch_pw(nMesh - 1, PW_CLIP);
}

int ec2d_ns_time_advance(int nMesh) {
  if (check_ready() || nMesh < 1 || nMesh > R.num_mesh) return fail("bad mesh level");
  return time_advance(nMesh - 1);
}

int ec2d_ns_multigrid(void) {
  if (check_ready()) return 1;
  if (R.num_mesh < 2) return fail("multigrid needs at least two mesh levels");
  for (int l = 0; l < R.num_mesh; l++)
    if (flush_clip(l)) return 1;
  return multigrid();
}

int ec2d_interpolation2h(int m1, int m2, int flag) {
  if (check_ready() || m1 < 1 || m2 != m1 + 1 || m2 > R.num_mesh) return fail("bad mesh levels");
  if (flag != 1 && flag != 2) return fail("flag must be 1 (U) or 2 (Res)");
  if (flush_clip(m1 - 1)) return 1;
  return launch_xfer(m1 - 1, m2 - 1, flag, true);
}

int ec2d_prolong_u(int m1, int m2, int flag) {
  if (check_ready() || m2 < 1 || m1 != m2 + 1 || m1 > R.num_mesh) return fail("bad mesh levels");
  if (flag != 1 && flag != 2) return fail("flag must be 1 (U) or 2 (deltU)");
  return launch_xfer(m2 - 1, m1 - 1, flag, false);
}

int ec2d_comput_force_function(int n) {
  if (check_ready() || n < 2 || n > R.num_mesh) return fail("bad mesh level");
  return launch_pw(n - 1, PW_FORCE);
}
int ec2d_add_force_function(int n) {
  if (check_ready() || n < 2 || n > R.num_mesh) return fail("bad mesh level");
  return launch_pw(n - 1, PW_ADD_FORCE);
}
int ec2d_comput_delt_u(int n) {
  if (check_ready() || n < 1 || n > R.num_mesh) return fail("bad mesh level");
  return launch_pw(n - 1, PW_DELT_U);
}
int ec2d_set_un(int n) {
  if (check_ready() || n < 1 || n > R.num_mesh) return fail("bad mesh level");
  return launch_pw(n - 1, PW_SET_UN);
}
int ec2d_comput_new_u(int n) {
  if (check_ready() || n < 1 || n > R.num_mesh) return fail("bad mesh level");
  return launch_pw(n - 1, PW_NEW_U);
}

int ec2d_run(int n_iter) {
  if (check_ready(false)) return 1;
  if (n_iter <= 0) return 0;
  R.res_ms = R.lusgs_ms = R.ghost_ms = 0.0;
  HIPCHK(hipEventRecord(R.ev0, R.s));
  // dual-time stepping reads the norms on the host inside every step: eager
  if (R.profile || R.cases[0].ctl.Time_Method == -1) {
    for (auto e : R.pev) hipEventDestroy(e);
    R.pev.clear();
    R.pev_kind.clear();
    for (int k = 0; k < n_iter; k++)
      if (main_step()) return 1;
  } else if (R.graph && std::memcmp(R.graph_pend, R.clip_pending, sizeof(R.graph_pend)) == 0) {
    // the graph of an earlier call still holds this loop body: replay it
    for (int k = 0; k < n_iter; k++) HIPCHK(hipGraphLaunch(R.graph, R.s));
    for (int k = 0; k < n_iter; k++)
      for (int l = 0; l < 3; l++)
        for (int t = 0; t < R.graph_ticks[l]; t++) tick(l);
    if (R.graph_sko) R.lv[0].planar_stale = true;  // as advance_lusgs marks it in an eager pass
  } else {
    // capture one pass of the loop body; the clip flag must be in a steady
    // state (the body leaves it as it found it from the second pass on)
    if (main_step()) return 1;
    int rest = n_iter - 1;
    if (rest > 0) {
      bool pend_before[3];
      std::memcpy(pend_before, R.clip_pending, sizeof(pend_before));
      hipGraph_t g;
      int before[3] = {tick_count[0], tick_count[1], tick_count[2]};
      const long long sko_before = R.sko_passes;
      HIPCHK(hipStreamBeginCapture(R.s, hipStreamCaptureModeThreadLocal));
      int rc = main_step();
      const bool body_sko = R.sko_passes != sko_before;
      int pass_ticks[3] = {tick_count[0] - before[0], tick_count[1] - before[1], tick_count[2] - before[2]};
      hipError_t e = hipStreamEndCapture(R.s, &g);
      if (rc) return rc;
      if (e != hipSuccess) return fail(std::string("graph capture: ") + hipGetErrorString(e));
      bool steady = std::memcmp(pend_before, R.clip_pending, sizeof(pend_before)) == 0;
      if (!steady) return fail("loop body is not in a steady state for graph replay");
      hipGraphExec_t ge;
      HIPCHK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
      HIPCHK(hipGraphDestroy(g));
      // capture ran the host side of one pass (its tick() calls) without
      // executing it; the graph runs `rest` passes, so account for the other
      // rest-1 passes' time counters here
      for (int k = 0; k < rest; k++) HIPCHK(hipGraphLaunch(ge, R.s));
      for (int k = 1; k < rest; k++)
        for (int l = 0; l < 3; l++)
          for (int t = 0; t < pass_ticks[l]; t++) tick(l);
      if (R.graph) HIPCHK(hipGraphExecDestroy(R.graph));
      R.graph = ge;
      std::memcpy(R.graph_pend, pend_before, sizeof(R.graph_pend));
      std::memcpy(R.graph_ticks, pass_ticks, sizeof(R.graph_ticks));
      R.graph_sko = body_sko;
    }
  }
  HIPCHK(hipEventRecord(R.ev1, R.s));
  HIPCHK(hipEventSynchronize(R.ev1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, R.ev0, R.ev1));
  R.last_run_ms = ms;
  if (int rc = device_errors(true)) return rc;
  if (R.profile) {
    // mean launch time per (kernel, level); ec2d_timing reports the finest level
    double acc[40] = {0};
    int cnt[40] = {0};
    // marks nest (the LU-SGS bracket holds its three kernels): pair each
    // opening mark with the next closing mark of the same kind
    for (size_t k = 0; k + 1 < R.pev.size(); k++) {
      int kind = R.pev_kind[k];
      if (kind <= 0 || kind >= 40) continue;
      for (size_t e = k + 1; e < R.pev.size(); e++) {
        if (R.pev_kind[e] == kind) break;
        if (R.pev_kind[e] != -kind) continue;
        float t = 0;
        hipEventElapsedTime(&t, R.pev[k], R.pev[e]);
        acc[kind] += t;
        cnt[kind]++;
        break;
      }
    }
    for (int k = 0; k < 40; k++) {
      R.prof_mean[k] = cnt[k] ? acc[k] / cnt[k] : 0.0;
      R.prof_count[k] = cnt[k];
    }
    R.res_ms = R.prof_mean[1];
    R.lusgs_ms = R.prof_mean[2];
    R.ghost_ms = R.prof_mean[3];
  }
  return 0;
}

int ec2d_set_wait_budget(int polls) {
  if (polls < 0) return fail("the wait budget is a number of polls >= 0");
  if (polls == R.wait_budget) return 0;
  R.wait_budget = polls;
  return drop_graph();  // a captured graph holds the old budget in its launch arguments
}

int ec2d_set_profile(int on) {
  R.profile = on;
  return 0;
}

int ec2d_run_to_drop(double drop, int max_iter, int check_every, int* iters, int* status, double* ms) {
  if (check_ready(false)) return 1;
  if (!(drop > 0.0 && drop < 1.0) || max_iter < 1 || check_every < 1)
    return fail("ec2d_run_to_drop: need 0 < drop < 1, max_iter >= 1 and check_every >= 1");
  if (R.cases[0].ctl.Time_Method == -1) return fail("ec2d_run_to_drop: dual-time stepping runs a fixed physical time");
  if (R.profile) return fail("ec2d_run_to_drop: switch profiling off first (ec2d_set_profile(0))");
  const int nc = (int)R.cases.size();
  if (!R.dactive) {
    HIPCHK(hipMalloc(&R.dactive, sizeof(int) * nc));
    HIPCHK(hipMalloc(&R.dconv, sizeof(ConvState) * nc));
    HIPCHK(hipMalloc(&R.dcold, sizeof(int) * nc));
  }
  HIPCHK(hipMemset(R.dcold, 0, sizeof(int) * nc));
  std::vector<int> act(nc, 1);
  std::vector<ConvState> st(nc);
  std::memset(st.data(), 0, sizeof(ConvState) * nc);
  HIPCHK(hipMemcpy(R.dactive, act.data(), sizeof(int) * nc, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(R.dconv, st.data(), sizeof(ConvState) * nc, hipMemcpyHostToDevice));
  const int* mask = R.dactive;
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(g_case_active), &mask, sizeof(mask)));
  R.conv_on = true;
  R.conv_drop = drop;
  R.conv_max = max_iter;
  if (drop_graph()) return 1;
  std::vector<std::array<int, 3>> k0(nc);
  std::vector<std::array<double, 3>> t0(nc);
  for (int c = 0; c < nc; c++)
    for (int l = 0; l < 3; l++) k0[c][l] = R.cases[c].kstep[l], t0[c][l] = R.cases[c].tt[l];
  std::vector<double> t_done(nc, 0.0);
  std::vector<char> seen(nc, 0);
  int rc = 0, passes = 0, left = nc;
  const auto start = std::chrono::steady_clock::now();
  const char* pe = std::getenv("EC2D_PROGRESS");
  const bool progress = pe && pe[0] == '1';
  double last_print = 0.0;
  while (left > 0 && passes < max_iter + check_every) {
    if ((rc = ec2d_run(check_every))) break;
    passes += check_every;
    if (hipMemcpy(st.data(), R.dconv, sizeof(ConvState) * nc, hipMemcpyDeviceToHost) != hipSuccess) {
      rc = fail("ec2d_run_to_drop: reading the case states failed");
      break;
    }
    const double el = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start).count();
    for (int c = 0; c < nc; c++)
      if (!seen[c] && st[c].status != 0) seen[c] = 1, t_done[c] = el, left--;
    if (progress && (el - last_print > 10000.0 || left == 0)) {
      std::fprintf(stderr, "ec2d_run_to_drop: %d passes, %d of %d cases stopped, %.1f s\n", passes, nc - left, nc,
                   el / 1e3);
      last_print = el;
    }
  }
  // leave the mode: every case runs again in later calls
  R.conv_on = false;
  mask = nullptr;
  if (hipMemcpyToSymbol(HIP_SYMBOL(g_case_active), &mask, sizeof(mask)) != hipSuccess && !rc)
    rc = fail("ec2d_run_to_drop: clearing the case mask failed");
  if (drop_graph() && !rc) rc = 1;
  // Kstep and tt of each case count the passes that updated it (one fewer
  // than its residual checks); the host counted every pass for every case
  for (int c = 0; c < nc && passes > 0; c++) {
    const int upd = std::max(st[c].iters - 1, 0);
    for (int l = 0; l < R.num_mesh; l++) {
      const int per = (R.cases[c].kstep[l] - k0[c][l]) / passes;
      R.cases[c].kstep[l] = k0[c][l];
      R.cases[c].tt[l] = t0[c][l];
      for (int k = 0; k < upd * per; k++) {
        R.cases[c].tt[l] = R.cases[c].tt[l] + R.cases[c].ctl.dt_global;
        R.cases[c].kstep[l] += 1;
      }
    }
  }
  for (int c = 0; c < nc; c++) {
    if (iters) iters[c] = st[c].iters;
    if (status) status[c] = st[c].status;
    if (ms) ms[c] = t_done[c];
  }
  if (rc) return rc;
  // the reference's T < 1e-5 stop and a split sweep's timeout, as ec2d_res_norms reports them
  double mx[6], rms[6];
  return ec2d_res_norms(0, 1, mx, rms);
}

int ec2d_timing(double* run_ms, double* residual_ms, double* lusgs_ms, double* ghost_ms) {
  if (run_ms) *run_ms = R.last_run_ms;
  if (residual_ms) *residual_ms = R.res_ms;
  if (lusgs_ms) *lusgs_ms = R.lusgs_ms;
  if (ghost_ms) *ghost_ms = R.ghost_ms;
  return 0;
}

int ec2d_lusgs_shape(int nMesh, int lanes_per_cell, int workgroups_per_block) {
  if (check_ready() || nMesh < 1 || nMesh > R.num_mesh) return fail("bad mesh level");
  if (lanes_per_cell != 1 && lanes_per_cell != 2) return fail("lanes per cell must be 1 or 2");
  HIPCHK(hipStreamSynchronize(R.s));
  DLevel& L = R.lv[nMesh - 1];
  // 0: fused with 6 producer waves, -3: fused with 3
  const bool fused = workgroups_per_block == 0 || workgroups_per_block == -3;
  if (workgroups_per_block < 0 && !fused) return fail("workgroups per block 1, 2, 4, or 0 / -3 (fused)");
  if (fused) {
    if (lanes_per_cell != 1) return fail("the fused LU-SGS launch takes one lane per cell");
    for (auto& b : L.hb)
      if (b.ny - 1 > 64) return fail("the fused LU-SGS launch needs blocks of at most 64 rows of cells");
  }
  if (lusgs_layout(L, lanes_per_cell, fused ? 1 : workgroups_per_block, true, fused)) return 1;
  L.lusgs_u = 2;
  L.lusgs_fused = fused;
  L.fused_np = workgroups_per_block == -3 ? 3 : 6;
  if (R.graph) {  // a captured graph holds the old launch shape
    HIPCHK(hipGraphExecDestroy(R.graph));
    R.graph = nullptr;
  }
  return 0;
}

int ec2d_autotune(const char* cache_path, int reps, char* report, int report_len) {
  if (check_ready()) return 1;
  if (R.profile) return fail("autotune needs the profile mode off");
  if (reps < 1) reps = 5;
  HIPCHK(hipStreamSynchronize(R.s));  // the layouts below are rewritten with null-stream copies
  // cache: one line per level configuration, "key tile_rows lpc split steps"
  struct Choice { int tr, lp, sp, u; };
  std::vector<std::pair<std::string, Choice>> cache;
  if (cache_path && cache_path[0]) {
    if (FILE* f = std::fopen(cache_path, "r")) {
      char key[512];
      int tr, lp, sp, u;
      while (std::fscanf(f, "%511s %d %d %d %d", key, &tr, &lp, &sp, &u) == 5) cache.push_back({key, {tr, lp, sp, u}});
      std::fclose(f);
    }
  }
  std::string rep_txt;
  std::vector<std::pair<std::string, Choice>> added;
  for (int l = 0; l < R.num_mesh; l++) {
    DLevel& L = R.lv[l];
    // key: the level's block shapes, equations and discretisation
    std::string key = "nv" + std::to_string(L.o.nvar) + "_s" + std::to_string(L.o.scheme) + "_f" +
                      std::to_string(L.o.flux) + "_r" + std::to_string(L.o.recon) + "_t" + std::to_string(L.o.turb) +
                      "_v" + std::to_string(R.cases[0].cp.If_viscous) + (L.fdm ? "_fdm" : "");
    // the block shapes in order, as a count and a 64-bit FNV-1a hash (a key
    // stays one short word however many blocks the level has)
    {
      std::string shapes;
      for (auto& b : L.hb) shapes += "_" + std::to_string(b.nx) + "x" + std::to_string(b.ny);
      uint64_t h = 1469598103934665603ull;
      for (unsigned char ch : shapes) h = (h ^ ch) * 1099511628211ull;
      char hx[48];
      std::snprintf(hx, sizeof(hx), "_b%zu_%016llx", L.hb.size(), (unsigned long long)h);
      key += hx;
    }
    int best_tr = -1, best_lp = -1, best_sp = 1, best_u = 2;
    for (auto& e : cache)
      if (e.first == key) { best_tr = e.second.tr; best_lp = e.second.lp; best_sp = e.second.sp; best_u = e.second.u; }
    char line[1024];
    if (best_tr < 0) {
      // the timed launches write the level's state (with a learned closure,
      // its under-relaxed mu_t carries over to the next residual): keep a copy
      // of the level's planes and put it back after timing
      double* keep = nullptr;
      HIPCHK(hipMalloc(&keep, sizeof(double) * (size_t)L.arena_n));
      HIPCHK(hipMemcpyAsync(keep, L.arena, sizeof(double) * (size_t)L.arena_n, hipMemcpyDeviceToDevice, R.s));
      // residual tile shape: 64 x 1, 2, 4 or 8 cells, or 32 x 8 or 16 (the
      // cache holds the rows, plus 100 for 32 columns)
      double best = 1e30;
      for (int tr : {1, 2, 4, 8, 108, 116}) {
        L.tile_cols = tr >= 100 ? 32 : 64;
        L.tile_rows = tr % 100;
        if (build_tiles(L)) return 1;
        double ms;
        if (time_launches(reps, [&] { return launch_residual(l, 0); }, &ms)) return 1;
        if (ms < best) { best = ms; best_tr = tr; }
      }
      // LU-SGS lanes per cell and workgroups per block (the state after a
      // residual; dU only is written)
      best = 1e30;
      L.lusgs_fused = false;
      std::string cand;  // every timed LU-SGS shape, for the report
      auto note = [&](const char* tag, double ms) {
        char t[64];
        std::snprintf(t, sizeof(t), " %s %.3f", tag, ms);
        cand += t;
      };
      for (int sp : {1, 2, 4})
        for (int lp : {1, 2}) {
          if (lusgs_layout(L, lp, sp, true)) {
            R.err.clear();
            continue;  // layout not possible for these blocks
          }
          // record prefetch depth: steps per register set (deeper sets for
          // up to four waves with two lanes per cell)
          for (int u : {2, 3, 4}) {
            if (u != 2 && !(lp == 2 && L.lusgs_threads <= 256 && (u == 4 || L.lusgs_threads > 64))) continue;
            L.lusgs_u = u;
            double ms;
            if (time_launches(reps, [&] { return launch_lusgs(l, L.dblocks_all, (int)L.hb.size(), 0.0, 0); }, &ms))
              return 1;
            char tag[32];
            std::snprintf(tag, sizeof(tag), "w%dl%du%d", sp, lp, u);
            note(tag, ms);
            if (ms < best) { best = ms; best_lp = lp; best_sp = sp; best_u = u; }
          }
          L.lusgs_u = 2;
        }
      // the fused variant: records made in the sweep workgroup (rows <= 64)
      bool fits = true;
      for (auto& b : L.hb) fits = fits && b.ny - 1 <= 64;
      if (fits && !lusgs_layout(L, 1, 1, true, true)) {
        L.lusgs_fused = true;
        for (int np : {3, 6}) {
          L.fused_np = np;
          double ms;
          if (time_launches(reps, [&] { return launch_lusgs(l, L.dblocks_all, (int)L.hb.size(), 0.0, 0); }, &ms))
            return 1;
          note(np == 3 ? "fused3" : "fused6", ms);
          // the cache's fourth field holds the producer waves of a fused choice
          if (ms < best) { best = ms; best_lp = 1; best_sp = 0; best_u = np; }
        }
        L.lusgs_fused = false;
      }
      R.err.clear();
      HIPCHK(hipMemcpyAsync(L.arena, keep, sizeof(double) * (size_t)L.arena_n, hipMemcpyDeviceToDevice, R.s));
      HIPCHK(hipStreamSynchronize(R.s));
      HIPCHK(hipFree(keep));
      if (best_lp < 0) return fail("no LU-SGS layout fits this level");
      added.push_back({key, {best_tr, best_lp, best_sp, best_u}});
      std::snprintf(line, sizeof(line), "level %d: measured tile %dx%d, lanes per cell %d, %s;", l + 1,
                    best_tr >= 100 ? 32 : 64, best_tr % 100, best_lp, shape_text(best_sp, best_u).c_str());
      rep_txt += line;
      // LU-SGS times of every shape (w: workgroups per block, l: lanes per cell, u: steps per set)
      std::snprintf(line, sizeof(line), " LU-SGS ms:%s\n", cand.c_str());
    } else {
      std::snprintf(line, sizeof(line), "level %d: cached tile %dx%d, lanes per cell %d, %s\n", l + 1,
                    best_tr >= 100 ? 32 : 64, best_tr % 100, best_lp, shape_text(best_sp, best_u).c_str());
    }
    rep_txt += line;
    L.tile_cols = best_tr >= 100 ? 32 : 64;
    L.tile_rows = best_tr % 100;
    if (build_tiles(L)) return 1;
    if (lusgs_layout(L, best_lp, best_sp == 0 ? 1 : best_sp, true, best_sp == 0)) return 1;
    L.lusgs_fused = best_sp == 0;
    L.lusgs_u = L.lusgs_fused ? 2 : best_u;
    if (L.lusgs_fused) L.fused_np = best_u == 3 ? 3 : 6;
  }
  if (!added.empty() && cache_path && cache_path[0]) {
    if (FILE* f = std::fopen(cache_path, "a")) {
      for (auto& e : added)
        std::fprintf(f, "%s %d %d %d %d\n", e.first.c_str(), e.second.tr, e.second.lp, e.second.sp, e.second.u);
      std::fclose(f);
    }
  }
  if (report && report_len > 0) {
    std::strncpy(report, rep_txt.c_str(), report_len - 1);
    report[report_len - 1] = 0;
  }
  // a captured graph holds the old launch shapes
  if (R.graph) {
    HIPCHK(hipGraphExecDestroy(R.graph));
    R.graph = nullptr;
  }
  return 0;
}


int ec2d_closure_collect(int* nsamples) {
  if (check_ready()) return 1;
  if (R.lv[0].o.nvar != 5 || R.lv[0].o.turb != 2) return fail("the closure trains on SA cases");
  if (closure_cells()) return 1;
  if (!R.clo.data) HIPCHK(hipMalloc(&R.clo.data, sizeof(float) * 8 * (size_t)R.clo.ncell));
  ClosureArgs a = closure_args();
  hipLaunchKernelGGL(k_closure_collect, dim3((unsigned)((R.clo.ncell + 255) / 256)), dim3(256), 0, R.s, a);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(R.s));
  if (nsamples) *nsamples = R.clo.ncell;
  return 0;
}

int ec2d_closure_train(int epochs, int batch, double lr, unsigned seed, double* first_epoch_loss,
                       double* last_epoch_loss) {
  if (check_ready()) return 1;
  if (!R.clo.data) return fail("no training data (ec2d_closure_collect)");
  if (epochs < 1 || batch < 64 || batch % 64 != 0) return fail("epochs >= 1, batch a multiple of 64");
  const int n = R.clo.ncell;
  std::vector<float> d((size_t)8 * n);
  HIPCHK(hipMemcpyAsync(d.data(), R.clo.data, sizeof(float) * d.size(), hipMemcpyDeviceToHost, R.s));
  HIPCHK(hipStreamSynchronize(R.s));
  std::vector<int> valid;
  for (int c = 0; c < n; c++)
    if (d[(size_t)8 * c + 7] != 0.f) valid.push_back(c);
  if ((int)valid.size() < batch) return fail("fewer valid samples than one batch");
  Lcg rng(seed);
  if (!R.clo.ready) {
    // normalisation and range of the training features
    float norm[4 * CL_NF];
    for (int k = 0; k < CL_NF; k++) {
      double s1 = 0, s2 = 0, lo = 1e300, hi = -1e300;
      for (int c : valid) {
        const double x = d[(size_t)8 * c + k];
        s1 += x;
        s2 += x * x;
      }
      const double mean = s1 / valid.size();
      const double var = std::max(s2 / valid.size() - mean * mean, 1e-12);
      const double inv = 1.0 / std::sqrt(var);
      for (int c : valid) {
        const double x = (d[(size_t)8 * c + k] - mean) * inv;
        lo = std::min(lo, x);
        hi = std::max(hi, x);
      }
      const double pad = 0.05 * (hi - lo);
      norm[k] = (float)mean;
      norm[CL_NF + k] = (float)inv;
      norm[2 * CL_NF + k] = (float)(lo - pad);
      norm[3 * CL_NF + k] = (float)(hi + pad);
    }
    // Xavier-uniform weights, zero biases
    std::vector<float> th(CL_NPARAM, 0.f);
    auto fill = [&](int off, int fin, int fout) {
      const double lim = std::sqrt(6.0 / (fin + fout));
      for (int k = 0; k < fin * fout; k++) th[off + k] = (float)((2.0 * rng.uniform() - 1.0) * lim);
    };
    fill(CL_OW1, CL_NF, 64);
    fill(CL_OW2, 64, 64);
    fill(CL_OW3, 64, 64);
    fill(CL_OW4, 64, 1);
    if (closure_upload(th.data(), norm)) return 1;
  }
  const int steps_per_epoch = (int)valid.size() / batch;
  const int nsteps = epochs * steps_per_epoch;
  // sample order: one permutation per epoch
  std::vector<int> order((size_t)nsteps * batch);
  for (int e = 0; e < epochs; e++) {
    std::vector<int> p = valid;
    for (int k = (int)p.size() - 1; k > 0; k--) std::swap(p[k], p[rng.next() % (unsigned)(k + 1)]);
    std::copy(p.begin(), p.begin() + (size_t)steps_per_epoch * batch, order.begin() + (size_t)e * steps_per_epoch * batch);
  }
  hipFree(R.clo.idx);
  hipFree(R.clo.part);
  hipFree(R.clo.losses);
  HIPCHK(hipMalloc(&R.clo.idx, sizeof(int) * order.size()));
  HIPCHK(hipMemcpyAsync(R.clo.idx, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice, R.s));
  const int nwg = batch / 64;
  HIPCHK(hipMalloc(&R.clo.part, sizeof(float) * (size_t)nwg * (CL_NPARAM + 1)));
  HIPCHK(hipMalloc(&R.clo.losses, sizeof(float) * nsteps));
  for (int st = 0; st < nsteps; st++) {
    TrainArgs ta;
    ta.net = R.clo.dnet;
    ta.data = R.clo.data;
    ta.idx = R.clo.idx + (size_t)st * batch;
    ta.nbatch = batch;
    ta.part = R.clo.part;
    hipLaunchKernelGGL(k_closure_train, dim3(nwg), dim3(256), 0, R.s, ta);
    AdamArgs aa;
    aa.part = R.clo.part;
    aa.nparts = nwg;
    aa.theta = R.clo.theta;
    aa.m1 = R.clo.m1;
    aa.m2 = R.clo.m2;
    aa.step = ++R.clo.step;
    aa.lr = (float)lr;
    aa.beta1 = 0.9f;
    aa.beta2 = 0.999f;
    aa.eps = 1e-8f;
    aa.net = R.clo.dnet;
    aa.loss = R.clo.losses + st;
    hipLaunchKernelGGL(k_closure_adam, dim3((CL_NPARAM + 1 + 255) / 256), dim3(256), 0, R.s, aa);
  }
  HIPCHK(hipGetLastError());
  std::vector<float> L(nsteps);
  HIPCHK(hipMemcpyAsync(L.data(), R.clo.losses, sizeof(float) * nsteps, hipMemcpyDeviceToHost, R.s));
  HIPCHK(hipStreamSynchronize(R.s));
  double f = 0, l = 0;
  for (int k = 0; k < steps_per_epoch; k++) {
    f += L[k];
    l += L[(size_t)(epochs - 1) * steps_per_epoch + k];
  }
  if (first_epoch_loss) *first_epoch_loss = f / steps_per_epoch;
  if (last_epoch_loss) *last_epoch_loss = l / steps_per_epoch;
  return 0;
}

int ec2d_closure_gradient(int n, float* grad, double* loss) {
  if (check_ready()) return 1;
  if (!R.clo.data) return fail("no training data (ec2d_closure_collect)");
  if (!R.clo.ready) return fail("closure has no parameters (train or set them first)");
  if (n < 64 || n % 64 != 0) return fail("the gradient batch must be a positive multiple of 64");
  std::vector<float> d((size_t)8 * R.clo.ncell);
  HIPCHK(hipMemcpyAsync(d.data(), R.clo.data, sizeof(float) * d.size(), hipMemcpyDeviceToHost, R.s));
  HIPCHK(hipStreamSynchronize(R.s));
  std::vector<int> idx;
  for (int c = 0; c < R.clo.ncell && (int)idx.size() < n; c++)
    if (d[(size_t)8 * c + 7] != 0.f) idx.push_back(c);
  if ((int)idx.size() < n) return fail("fewer valid samples than the gradient batch");
  int* didx = nullptr;
  float *part = nullptr, *out = nullptr;
  HIPCHK(hipMalloc(&didx, sizeof(int) * n));
  HIPCHK(hipMalloc(&part, sizeof(float) * (size_t)(n / 64) * (CL_NPARAM + 1)));
  HIPCHK(hipMalloc(&out, sizeof(float) * (CL_NPARAM + 1)));
  HIPCHK(hipMemcpyAsync(didx, idx.data(), sizeof(int) * n, hipMemcpyHostToDevice, R.s));
  TrainArgs ta;
  ta.net = R.clo.dnet;
  ta.data = R.clo.data;
  ta.idx = didx;
  ta.nbatch = n;
  ta.part = part;
  hipLaunchKernelGGL(k_closure_train, dim3(n / 64), dim3(256), 0, R.s, ta);
  hipLaunchKernelGGL(k_closure_gsum, dim3((CL_NPARAM + 1 + 255) / 256), dim3(256), 0, R.s, part, n / 64, out);
  std::vector<float> g(CL_NPARAM + 1);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(g.data(), out, sizeof(float) * g.size(), hipMemcpyDeviceToHost, R.s);
  if (e == hipSuccess) e = hipStreamSynchronize(R.s);
  hipFree(didx);
  hipFree(part);
  hipFree(out);
  if (e != hipSuccess) return fail(std::string("closure gradient: ") + hipGetErrorString(e));
  if (grad) std::memcpy(grad, g.data(), sizeof(float) * CL_NPARAM);
  if (loss) *loss = g[CL_NPARAM];
  return 0;
}

int ec2d_closure_get_params(float* theta, float* norm) {
  if (!R.clo.ready) return fail("closure has no parameters");
  if (theta) {
    HIPCHK(hipMemcpyAsync(theta, R.clo.theta, sizeof(float) * CL_NPARAM, hipMemcpyDeviceToHost, R.s));
    HIPCHK(hipStreamSynchronize(R.s));
  }
  if (norm) std::memcpy(norm, R.clo.norm, sizeof(R.clo.norm));
  return 0;
}

int ec2d_closure_set_params(const float* theta, const float* norm) {
  if (check_ready()) return 1;
  if (!theta || !norm) return fail("null parameters");
  if (R.lv[0].pm.clo < 0) return fail("the closure runs on SA and model-4 cases");
  // the cell list exists before a run graph is captured (no allocation under capture)
  if (closure_cells()) return 1;
  R.clo.stalled = false;
  R.clo.best_rms = 0.0;
  return closure_upload(theta, norm);
}

int ec2d_closure_relaxation(double w) {
  if (!(w >= 0.0 && w <= 1.0)) return fail("relaxation factor must lie in [0, 1]");
  R.clo.relax = w;
  if (R.graph) {  // the factor is a kernel argument of the captured graph
    HIPCHK(hipGraphExecDestroy(R.graph));
    R.graph = nullptr;
  }
  return 0;
}

int ec2d_closure_infer(void) {
  if (check_ready()) return 1;
  if (R.lv[0].pm.clo < 0) return fail("the closure runs on SA cases");
  return closure_infer();
}

int ec2d_closure_enable(int on) {
  if (check_ready()) return 1;
  if (on && R.lv[0].pm.clo < 0) return fail("the closure runs on SA cases");
  if (on && !R.clo.ready) return fail("closure has no parameters (train or set them first)");
  // the cell list exists before a run graph is captured (no allocation under capture)
  if (on && closure_cells()) return 1;
  R.clo.enabled = on != 0;
  R.clo.best_rms = 0.0;
  R.clo.stalled = false;
  if (R.graph) {
    HIPCHK(hipGraphExecDestroy(R.graph));
    R.graph = nullptr;
  }
  return 0;
}

int ec2d_closure_monitor(double growth, int* active) {
  if (check_ready()) return 1;
  // model-4 cases have no SA to switch to inside the batch: a stall is
  // reported (active = 0) and the caller restarts the case with SA
  const bool standin = R.lv[0].o.turb == 4;
  if (R.clo.enabled || (standin && !R.clo.stalled)) {
    double worst = 0.0;
    for (int c = 0; c < (int)R.cases.size(); c++) {
      double mx[6], rms[6];
      int rc = ec2d_res_norms(c, 1, mx, rms);
      if (rc == 1) return 1;
      worst = std::max(worst, rms[0]);
    }
    if (R.clo.best_rms == 0.0 || worst < R.clo.best_rms) R.clo.best_rms = worst;
    else if (worst > growth * R.clo.best_rms || !std::isfinite(worst)) {
      // the residual stalls or grows: fall back to SA everywhere
      if (standin) R.clo.stalled = true;
      else R.clo.enabled = false;
      if (R.graph) {
        HIPCHK(hipGraphExecDestroy(R.graph));
        R.graph = nullptr;
      }
    }
  }
  if (active) *active = standin ? (R.clo.stalled ? 0 : 1) : (R.clo.enabled ? 1 : 0);
  return 0;
}

int ec2d_force_coefficients(int case_id, double* cl, double* cd, double* forces) {
  if (check_ready()) return 1;
  if (case_id < 0 || case_id >= (int)R.cases.size()) return fail("bad case id");
  DLevel& L = R.lv[0];
  ForceArgs a;
  a.blocks = L.db;
  a.faces = L.dwall;
  a.case_face0 = L.dcase_wall0;
  a.cps = R.dcps;
  a.arena = L.arena;
  a.pm = L.pm;
  a.viscous = R.cases[0].cp.If_viscous == 1;
  a.out = L.dforces;
  hipLaunchKernelGGL(k_forces, dim3((unsigned)R.cases.size()), dim3(256), 0, R.s, a);
  HIPCHK(hipGetLastError());
  double F[4];
  HIPCHK(hipMemcpyAsync(F, L.dforces + 4 * case_id, sizeof(F), hipMemcpyDeviceToHost, R.s));
  HIPCHK(hipStreamSynchronize(R.s));
  const double al = R.cases[case_id].ctl.AoA;
  const double fx = F[0] + F[2], fy = F[1] + F[3];
  if (cl) *cl = (-fx * std::sin(al) + fy * std::cos(al)) / 0.5;
  if (cd) *cd = (fx * std::cos(al) + fy * std::sin(al)) / 0.5;
  if (forces)
    for (int k = 0; k < 4; k++) forces[k] = F[k];
  return 0;
}

int ec2d_kernel_time(int nMesh, int kind, double* ms, int* launches) {
  if (nMesh < 1 || nMesh > 3) return fail("bad mesh level");
  if (kind < 1 || kind > 6) return fail("bad kernel kind");
  const int k = 10 * (nMesh - 1) + kind;
  if (ms) *ms = R.prof_mean[k];
  if (launches) *launches = R.prof_count[k];
  return 0;
}

int ec2d_timing_level(int nMesh, double* residual_ms, double* lusgs_ms, double* ghost_ms, int* launches) {
  if (nMesh < 1 || nMesh > 3) return fail("bad mesh level");
  int b = 10 * (nMesh - 1);
  if (residual_ms) *residual_ms = R.prof_mean[b + 1];
  if (lusgs_ms) *lusgs_ms = R.prof_mean[b + 2];
  if (ghost_ms) *ghost_ms = R.prof_mean[b + 3];
  if (launches) {
    launches[0] = R.prof_count[b + 1];
    launches[1] = R.prof_count[b + 2];
    launches[2] = R.prof_count[b + 3];
  }
  return 0;
}

int ec2d_synchronize(void) {
  if (!R.inited) return 0;
  HIPCHK(hipStreamSynchronize(R.s));
  return R.committed ? device_errors(true) : 0;
}

int ec2d_res_norms(int case_id, int nMesh, double* res_max, double* res_rms) {
  if (check_ready(false)) return 1;
  if (case_id < 0 || case_id >= (int)R.cases.size() || nMesh < 1 || nMesh > R.num_mesh) return fail("bad case / mesh");
  DLevel& L = R.lv[nMesh - 1];
  double buf[12];
  if (L.residual_done && reduce_norms(L)) return 1;
  HIPCHK(hipMemcpyAsync(buf, L.dnorms + 12 * case_id, sizeof(buf), hipMemcpyDeviceToHost, R.s));
  HIPCHK(hipStreamSynchronize(R.s));
  int se = 0;
  HIPCHK(hipMemcpy(&se, R.dsweep_err, sizeof(se), hipMemcpyDeviceToHost));
  if (se) return device_errors(true);
  for (int m = 0; m < 6; m++) {
    res_max[m] = buf[m];
    res_rms[m] = buf[6 + m];
  }
  return device_errors(true);
}

int ec2d_time(int case_id, int nMesh, int* kstep, double* tt) {
  if (case_id < 0 || case_id >= (int)R.cases.size() || nMesh < 1 || nMesh > 3) return fail("bad case / mesh");
  *kstep = R.cases[case_id].kstep[nMesh - 1];
  *tt = R.cases[case_id].tt[nMesh - 1];
  return 0;
}

int ec2d_set_time(int case_id, int nMesh, int kstep, double tt) {
  if (case_id < 0 || case_id >= (int)R.cases.size() || nMesh < 1 || nMesh > 3) return fail("bad case / mesh");
  if (kstep < 0) return fail("Kstep must not be negative");
  R.cases[case_id].kstep[nMesh - 1] = kstep;
  R.cases[case_id].tt[nMesh - 1] = tt;
  return 0;
}

int ec2d_get_run_counters(long long* state) {
  if (!state) return fail("null state");
  state[0] = R.residual_calls;
  state[1] = R.dual_steps;
  state[2] = R.dual_kt;
  return 0;
}

int ec2d_set_run_counters(const long long* state) {
  if (!state) return fail("null state");
  if (state[0] < 0 || state[1] < 0 || state[2] < 0 || state[2] > INT32_MAX) return fail("bad run counters");
  R.residual_calls = state[0];
  R.dual_steps = state[1];
  R.dual_kt = (int)state[2];
  return 0;
}

long long ec2d_num_cells(int case_id, int nMesh) {
  if (!R.committed || nMesh < 1 || nMesh > R.num_mesh) return -1;
  DLevel& L = R.lv[nMesh - 1];
  if (case_id < 0) {
    long long s = 0;
    for (auto v : L.case_cells) s += v;
    return s;
  }
  if (case_id >= (int)L.case_cells.size()) return -1;
  return L.case_cells[case_id];
}

}  // extern "C"
