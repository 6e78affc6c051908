/* ec2d.h - C-ABI of the MI355X residual / implicit-update path of OpenCFD-EC2D.
 *
 * Plain C types only (int, double, pointers); every entry point returns 0 on
 * success and a nonzero status otherwise, with the message available from
 * ec2d_last_error().  Arrays cross the boundary in the reference's own
 * Fortran layout (column-major, the reference's bounds, the equation index
 * fastest for U), so a Fortran caller passes c_loc() of its module arrays.
 *
 * Model: the library holds a batch of cases (one reference run each); every
 * solver entry point below acts on all cases of the batch at once, so a
 * single-case caller gets exactly the reference's subroutine and an ensemble
 * caller gets one launch per stage for all of its cases.  Per-case results
 * (residual norms, errors, time counters) are read back by case id.
 *
 * Reference interfaces replaced (file:line in OpenCFD-EC2D v1.5.4):
 *   see the comment of each function.
 */
#ifndef EC2D_H
#define EC2D_H

#ifdef __cplusplus
extern "C" {
#endif

/* control.in after read_parameter's conversions (opencfd_ec2d_v1.5.4.f90:197-235):
 * AoA in radians, Twall divided by T_inf, Cv and Cp derived. */
typedef struct {
  double Ma, Re, gamma, AoA, Pr, Cv, Cp;
  double t_end, dt_global, CFL, dtmax, dtmin, P_outlet, T_inf, Twall, vt_inf, Kt_inf, Wt_inf;
  int If_viscous, Iflag_turbulence_model, Iflag_init, Iflag_local_dt, Time_Method;
  int Iflag_Scheme, Iflag_Flux, IFlag_Reconstruction, Num_Mesh, Nvar;
  int Pre_Step_Mesh[3];
  int Nstep_Inner_Limit;  /* dual-time stepping (Time_Method -1): inner iterations at most */
  double Res_Inner_Limit; /* and the max_m Res_rms(m) that ends them (opencfd_ec2d_v1.5.4.f90:209) */
} ec2d_control;

/* array selectors for ec2d_put_array / ec2d_get_array (reference bounds) */
enum {
  EC2D_U = 1,      /* U(NV,-1:nx+1,-1:ny+1) */
  EC2D_RES = 2,    /* Res(NV,1:nx-1,1:ny-1) */
  EC2D_DT = 3,     /* dt(1:nx-1,1:ny-1) */
  EC2D_DU = 4,     /* dU(NV,1:nx,1:ny): written by ec2d_du_lu_sgs_2d; a main-loop
                      pass (ec2d_ns_time_advance, ec2d_run) on the fused LU-SGS
                      path updates U without rewriting it (nothing reads it there) */
  EC2D_LCI = 5, EC2D_LCJ = 6, EC2D_LVI = 7, EC2D_LVJ = 8, /* (1:nx,1:ny) */
  EC2D_AMU = 9, EC2D_AMU_T = 10,                          /* (-1:nx+1,-1:ny+1); after a residual Amu holds
                                                             the interior, Amu_t also the layer 0:nx x 0:ny
                                                             (Amut_boundary's ghost values, which output writes) */
  EC2D_VOL = 11, EC2D_SI = 12, EC2D_SJ = 13,                 /* (1:nx,1:ny), read only: cell volume, face areas */
  EC2D_NI1 = 14, EC2D_NI2 = 15, EC2D_NJ1 = 16, EC2D_NJ2 = 17, /* and face normals (sub_init.f90:180-204) */
  EC2D_X1 = 20, EC2D_Y1 = 21,                                /* (-1:nx+1,-1:ny+1), read only: cell centres */
  EC2D_DW = 22,    /* dw(0:nx,0:ny) wall distance (SA / SST / model 4) */
  EC2D_UN = 23,    /* Un(NVun,0:nx,0:ny) */
  EC2D_QF = 24,    /* QF(4,1:nx-1,1:ny-1) */
  EC2D_DELTU = 25, /* deltU(4,-1:nx+1,-1:ny+1) */
  EC2D_UN1 = 31,   /* Un1(NV,0:nx,0:ny), dual-time stepping only (interior set) */
  EC2D_CLOSURE_MUT = 40,  /* learned closure: mu_t of the last inference (-1:nx+1,-1:ny+1; SA cases: the
                             interior, model-4 cases: 0:nx,0:ny) */
  EC2D_CLOSURE_MASK = 41  /* learned closure: 1 where the features lay inside the training range */
};

/* ---- life cycle --------------------------------------------------------- */
int ec2d_init(void);     /* select the GPU, create the stream; no reference counterpart */
int ec2d_finalize(void); /* free everything */
int ec2d_last_error(char *buf, int len);

/* A new case with its control parameters (read_parameter,
 * opencfd_ec2d_v1.5.4.f90:197-235, and set_control_para, :239-260: the
 * coarse levels use UD1, no turbulence model and 4 equations).
 * Returns the case id (>= 0) or a negative status. */
int ec2d_case_create(const ec2d_control *ctl);

/* Geometry of block mBlock (1-based) of mesh level nMesh (1 = finest) of a
 * case, as the reference holds it after its set-up (Creat_Mesh1 /
 * Creat_Mesh, sub_init.f90:111-380; Update_coordinate_buffer,
 * sub_boundary.f90:159-301; comput_dw, sub_turbulence_SST.f90:274-365;
 * comput_Jacobian, sub_Finite_Difference.f90:314-382).
 *   bc_msg  9 ints per subface: f_no face ist iend jst jend neighb subface orient
 *   x, y    (-1:nx+2,-1:ny+2) nodes incl. ghosts
 *   x1, y1  (-1:nx+1,-1:ny+1) cell centres
 *   vol, si, sj, ni1, ni2, nj1, nj2  (1:nx,1:ny)
 *           x1 .. nj2 all NULL: ec2d_commit makes them on the device from the
 *           interior nodes x, y (1:nx,1:ny; the ghost entries of x, y are not
 *           read): the ghost nodes of Update_coordinate_buffer
 *           (sub_boundary.f90:170-281), the cell centres and Creat_Mesh's
 *           metrics (sub_init.f90:180-204); not for fvm_fdm = 1
 *   dw      (0:nx,0:ny) wall distance, NULL unless SA / SST / model 4; NULL for one of those: ec2d_commit
 *           computes it on the device (comput_dw, sub_turbulence_SST.f90:286-337)
 *   fd      5 arrays Akx,Aky,Aix,Aiy,Ajac of (1:nx-1,1:ny-1) back to back, NULL unless fvm_fdm = 1 */
int ec2d_define_block(int case_id, int nMesh, int mBlock, int nx, int ny, int nsub, const int *bc_msg, int fvm_fdm,
                      const double *x, const double *y, const double *x1, const double *y1, const double *vol,
                      const double *si, const double *sj, const double *ni1, const double *ni2, const double *nj1,
                      const double *nj2, const double *dw, const double *fd);

/* Allocate the device arenas of the batch and upload geometry; U starts at
 * the reference's allocation defaults (sub_init.f90:160-174, 310-318). */
int ec2d_commit(void);

int ec2d_put_array(int case_id, int nMesh, int mBlock, int which, const double *src);
/* Synchronises the stream; fails when an LU-SGS wait timed out (see
 * ec2d_synchronize).  The arrays of a run stopped by T < 1e-5 stay readable. */
int ec2d_get_array(int case_id, int nMesh, int mBlock, int which, double *dst);

/* ---- the per-iteration path (acts on all cases of the batch) ------------ */

/* Init_flow (sub_init.f90:383-495): freestream start with coarse-level
 * pre-iterations and prolongation (Iflag_init = 0), or ghost fill of a state
 * put with ec2d_put_array (Iflag_init = 1, as init_flow_read :499-560).
 * Clears both device error words (see ec2d_synchronize). */
int ec2d_init_flow(void);

/* Comput_Residual_one_mesh (sub_Residual.f90:13-107): Res, dt, Lci..Lvj and
 * the residual norms of every case. */
int ec2d_comput_residual_one_mesh(int nMesh);

/* du_LU_SGS_2D (sub_LU_SGS.f90:4-108); mBlock = 0 sweeps every block. */
int ec2d_du_lu_sgs_2d(int nMesh, int mBlock, double alfa1);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The same as ec2d_du_lu_sgs_2d, result for result, but
 * the LU-SGS records (and the rings of split sweeps) live in a static device
 * array of the library, at most 64 MiB per GPU (held until the process ends):
 * no hipMalloc, hipFree or allocating routine is called.  Fails with a message
 * when the level's records need more (a level of up to about 100,000 SA cells
 * fits).  Every call on a GPU shares that GPU's array and clears it first, so
 * calls that overlap on one GPU (several host threads, several streams) race. */
int ec2d_du_lu_sgs_2d_noalloc(int nMesh, int mBlock, double alfa1);

/* the update loop U = U + dU of NS_Time_advance_LU_SGS (sub_NS_singlegrid.f90:203-211) */
int ec2d_update_u(int nMesh);

/* Boundary_condition_onemesh (sub_boundary.f90:10-34) */
int ec2d_boundary_condition_onemesh(int nMesh);

/* update_buffer_onemesh (sub_boundary.f90:45-153) */
int ec2d_update_buffer_onemesh(int nMesh);

/* force_vt_kw (sub_NS_singlegrid.f90:30-60) */
int ec2d_force_vt_kw(int nMesh);

/* NS_Time_advance (sub_NS_singlegrid.f90:8-26): one LU-SGS, Euler or RK3
 * step including ghost fill and the nu_t / k / omega clip. */
int ec2d_ns_time_advance(int nMesh);

/* NS_2stge_multigrid / NS_3stge_multigrid (sub_NS_multigrid.f90:14-142) */
int ec2d_ns_multigrid(void);

/* multigrid pieces (sub_NS_multigrid.f90:156-398) */
int ec2d_interpolation2h(int m1, int m2, int flag);
int ec2d_prolong_u(int m1, int m2, int flag);
int ec2d_comput_force_function(int nMesh);
int ec2d_add_force_function(int nMesh);
int ec2d_comput_delt_u(int nMesh);
int ec2d_set_un(int nMesh);
int ec2d_comput_new_u(int nMesh);

/* n passes of the main-loop body (opencfd_ec2d_v1.5.4.f90:172-186) as one
 * HIP graph replayed n times; no host synchronisation inside, one at the end,
 * where the device error words are checked as in ec2d_synchronize.  On a
 * single level with the fused LU-SGS launch, the passes leave Res, dt and the
 * spectral radii in the update's hyperplane copy only.  The next call that
 * may read them restores their planes first, so every array reads as the
 * reference's. */
int ec2d_run(int n_iter);

/* Runs main-loop passes (as ec2d_run, graph replays of check_every passes)
 * until every case of the batch has reached its own residual drop: after the
 * first fine-level residual of each pass a device check reduces each case's
 * density Res_rms(1) (sub_Residual.f90:104-106) and stops the case
 * (status 1) once it is at most drop times its value at the run's first
 * check; a case whose Res_rms of any equation (density, momenta, energy,
 * turbulence) is not finite stops with status 2 (the NaN guard; the
 * reference has none), and one checked max_iter times without reaching
 * the drop with status 3.  A stopped case is frozen: every kernel skips its
 * blocks, so its state stays the one whose residual met the test and the
 * other cases run on at the same cost per pass for fewer cells.  Per case
 * (arrays of the batch's case count, each may be NULL): iters = residual
 * checks (the state has had iters - 1 updates in this call, and its Kstep and
 * tt count those), status, ms = host wall time from the call's start to the
 * check that saw it stop (resolution: check_every passes).  The reference's
 * T < 1e-5 stop (sub_Residual.f90:143-146) ends only the case it occurs in:
 * that case stops, frozen with the state of that residual, with status 4.
 * No reference counterpart: the reference runs a fixed
 * t_end / dt_global passes (opencfd_ec2d_v1.5.4.f90:172) and prints the
 * residual; this is its per-case stop for ensembles.  With the environment
 * variable EC2D_PROGRESS=1 it prints a progress line to stderr about every
 * 10 s. */
int ec2d_run_to_drop(double drop, int max_iter, int check_every, int *iters, int *status, double *ms);

/* Res_max(1:6), Res_rms(1:6) of mesh nMesh of a case (after a residual);
 * synchronises the stream.  Returns the T < 1e-5 error of the reference
 * (sub_Residual.f90:143-146) as status 2 with the cell in ec2d_last_error. */
int ec2d_res_norms(int case_id, int nMesh, double *res_max, double *res_rms);

/* Kstep and tt of a case's mesh level (Mesh_TYPE, opencfd_ec2d_v1.5.4.f90:88-90) */
int ec2d_time(int case_id, int nMesh, int *kstep, double *tt);

/* Dual-time stepping (Time_Method -1): the inner-loop index kt_in of the last
 * physical step, the value Dual_time_LU_SGS prints (sub_NS_singlegrid.f90:277;
 * Nstep_Inner_Limit + 1 when the inner loop ran out).  No reference
 * counterpart as an interface: the reference only prints it. */
int ec2d_dual_inner(int *kt_in);

/* Full-precision checkpoint support: the run state kept outside the arrays.
 * The arrays (U, Un, Un1, Amu_t of every level and block) travel through
 * ec2d_get_array / ec2d_put_array; these restore the rest, so that a run
 * resumed from a checkpoint continues bit for bit as the uninterrupted run.
 *   ec2d_set_time          sets Kstep and tt of a case's level (the values
 *                          ec2d_time reads, Mesh_TYPE, opencfd_ec2d_v1.5.4.f90:88-90);
 *   ec2d_get/set_run_counters  state[0] residual evaluations so far (the first
 *                          one takes Sfac = 1/(3 dt): the `save` flag Iflag1 of
 *                          Comput_Residual_one_mesh, sub_Residual.f90:19,28-34),
 *                          state[1] physical dual-time steps so far (alfa1 =
 *                          1/dt on the first: the flag Iflag of Dual_time_LU_SGS,
 *                          sub_NS_singlegrid.f90:239-245), state[2] kt_in of the
 *                          last physical step.
 * No reference counterpart as an interface: the reference restarts from
 * flow2d.dat, written at F20.10 (sub_init.f90:499-560), with Kstep, tt and
 * both flags starting afresh. */
int ec2d_set_time(int case_id, int nMesh, int kstep, double tt);
int ec2d_get_run_counters(long long *state);
int ec2d_set_run_counters(const long long *state);

/* Waits for the stream, then checks the two device error words, which stay
 * set until ec2d_init_flow: an LU-SGS wait between waves or workgroups that
 * ran out of its budget (status 1, results invalid; the split-sweep rings are
 * reset) and the reference's T < 1e-5 stop (sub_Residual.f90:143-146,
 * status 2 with the cell in ec2d_last_error). */
int ec2d_synchronize(void);

/* Cells of the batch, and of one case, on a mesh level (Num_Cell, sub_init.f90:134). */
long long ec2d_num_cells(int case_id, int nMesh);

/* Force coefficients of a case on the finest mesh, from its current state.
 * The reference has no force integration (only wall Cp, in the separate
 * post-processing tool read_flow-1.2.f90:36-51).  Definition: every wall face
 * contributes -(p - p_inf) n s (p of the adjacent interior cell) and the
 * viscous traction (tau . n) s with tau from the face-centred gradient
 * stencil of Residual_FVM (sub_Residual.f90:301-343, 408-450) and
 * mu = Amu + Amu_t of that cell; n points into the fluid.  The total is
 * rotated by AoA: Cl = (-Fx sin a + Fy cos a) / q, Cd = (Fx cos a + Fy sin a) / q,
 * q = 1/2 (unit chord).  forces (4 doubles, may be NULL): pressure Fx, Fy,
 * viscous Fx, Fy.  Synchronises the stream. */
int ec2d_force_coefficients(int case_id, double *cl, double *cd, double *forces);

/* ---- learned turbulence closure (no reference counterpart) --------------
 * A 6 -> 64 -> 64 -> 64 -> 1 MLP (tanh, softplus output) predicts nu_t / nu
 * per cell of the finest level from six local dimensionless features; BF16
 * weights and activations, FP32 accumulation, MFMA.  It runs inside SA cases:
 * SA keeps evolving nu~, and where a cell's features lie in the training
 * range the MLP's mu_t replaces SA's in the mean-flow viscous flux and the
 * spectral radii; elsewhere SA's mu_t stays (per-cell fallback).
 * Parameters: CL_NPARAM = 8833 floats (W1 6x64, b1, W2 64x64, b2, W3 64x64, b3,
 * W4 64, b4; W row-major [in][out]) and 24 floats of feature normalisation
 * (mean, 1/std, normalised lower and upper bounds of the valid range). */
int ec2d_closure_collect(int *nsamples);  /* samples from the current SA state */
/* Adam over the collected samples (first call: normalisation from the data,
 * Xavier-uniform start from seed); mean mini-batch loss of the first and of
 * the last epoch.  batch: a multiple of 64. */
int ec2d_closure_train(int epochs, int batch, double lr, unsigned seed, double *first_epoch_loss,
                       double *last_epoch_loss);
/* The gradient of the training loss with respect to the CL_NPARAM parameters
 * (layout of ec2d_closure_get_params) over the first n valid collected
 * samples as one mini-batch (n a multiple of 64), exactly as a training step
 * of ec2d_closure_train computes it before its Adam update, and that batch's
 * loss; the parameters do not change. */
int ec2d_closure_gradient(int n, float *grad, double *loss);
int ec2d_closure_get_params(float *theta, float *norm);
int ec2d_closure_set_params(const float *theta, const float *norm);
int ec2d_closure_infer(void);             /* fills EC2D_CLOSURE_MUT / _MASK from the current state */
int ec2d_closure_enable(int on);          /* use the closure in every following residual */
/* Model-4 cases: under-relax the network's mu_t from one residual to the next,
 * mu_t <- mu_t + w (mu_t(network) - mu_t), w in [0, 1] (1, the default: none;
 * the plane starts at 0, so mu_t ramps up from laminar; 0: mu_t frozen at its
 * current values, the network is not evaluated). */
int ec2d_closure_relaxation(double w);
/* Fallback on a stall: if the largest density Res_rms of the batch exceeds
 * growth times the smallest seen since enabling, switch the closure off
 * (SA everywhere).  active: 1 while the closure is on.  Synchronises. */
int ec2d_closure_monitor(double growth, int *active);

/* Device time of the last ec2d_run call in milliseconds (HIP events on the
 * library stream), and the mean time per launch of the residual and LU-SGS
 * kernels measured with events around them during that call (0 if not
 * measured).  No reference counterpart: measurement support. */
int ec2d_timing(double *run_ms, double *residual_ms, double *lusgs_ms, double *ghost_ms);

/* Per mesh level: mean launch time (ms) of the residual, LU-SGS and ghost
 * kernels and their launch counts (3 ints) during the last profiled ec2d_run.
 * No reference counterpart: measurement support. */
int ec2d_timing_level(int nMesh, double *residual_ms, double *lusgs_ms, double *ghost_ms, int *launches);

/* Per mesh level and kernel kind: mean launch time (ms) and launch count
 * during the last profiled ec2d_run.  Kinds: 1 residual (all its kernels),
 * 2 LU-SGS update (the three below together), 3 ghost fill, 4 LU-SGS record
 * preparation, 5 LU-SGS sweeps, 6 LU-SGS scatter.  No reference
 * counterpart: measurement support. */
int ec2d_kernel_time(int nMesh, int kind, double *ms, int *launches);

/* 1: ec2d_run launches eagerly and times every residual / LU-SGS / ghost
 * launch with HIP events (see ec2d_timing); 0 (default): one HIP graph.
 * No reference counterpart: measurement support. */
int ec2d_set_profile(int on);

/* The budget of the LU-SGS sweeps' bounded waits between waves and between
 * workgroups, in polls (default 2^20, about a second); past it the wave gives
 * up and the device error word is set (see ec2d_synchronize).  With 0 every
 * receiving wave of a split sweep gives up at its first slot: a host-forced
 * way to exercise the error path.  No reference counterpart. */
int ec2d_set_wait_budget(int polls);

/* No reference counterpart: the LU-SGS launch shape of one level, chosen by
 * hand (ec2d_autotune chooses it by timing otherwise): lanes per cell 1 or 2,
 * workgroups per block 1, 2 or 4 (each block's rows split between that many
 * CUs, for levels of at most 256/split blocks taller than split/2 waves), or
 * 0 with one lane per cell: the fused launch, whose workgroups make each
 * block's records themselves (blocks of at most 64 rows of cells), so that
 * they never go through device memory, with 6 producer waves and a writer
 * wave beside the sweep wave (-3: with 3 producers).  Results do not depend
 * on it. */
int ec2d_lusgs_shape(int nMesh, int lanes_per_cell, int workgroups_per_block);

/* Autotune the launch shapes of every mesh level on the committed batch:
 * residual tile rows (1, 2, 4, 8 x 64 cells), LU-SGS lanes per cell (1 or 2),
 * workgroups per block (1, 2 or 4, see ec2d_lusgs_shape) and sweep steps per
 * prefetched register set (2, 3 or 4 with two lanes per cell and at most four
 * waves), each timed with HIP events over `reps` launches on the current
 * state; the fastest is kept.  Results are cached per level configuration
 * (equations, scheme, flux, reconstruction, turbulence model, viscosity,
 * block shapes) in the text file cache_path (NULL or "": no cache).  report
 * receives one line per level.  Changes no result bit: the level's planes are
 * copied before the timed launches and put back after them (a learned
 * closure's under-relaxed mu_t included).  No reference counterpart. */
int ec2d_autotune(const char *cache_path, int reps, char *report, int report_len);

/* Tile shape of the fused residual kernel (rows x 64 cells); no reference
 * counterpart.  Call before ec2d_commit.  Valid rows: 1, 2, 4, 8. */
int ec2d_set_tile_rows(int rows);
/* The same with the width: 64 x 1, 2, 4 or 8 cells, or 32 x 8 or 32 x 16
 * (smaller halos per workgroup for the viscous kernels).  Results do not
 * depend on it (the residual norms' summation tree does, at round-off). */
int ec2d_set_tile_shape(int cols, int rows);

#ifdef __cplusplus
}
#endif
#endif
