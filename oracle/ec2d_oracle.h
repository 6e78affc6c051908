/* CPU restatement of the OpenCFD-EC2D residual / implicit-update path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle the HIP path is checked
 * against; it is never linked into, or called by, the product library.
 *
 * Arrays use the reference's own layout (column-major, same bounds, the
 * variable index fastest for U), so that inputs and outputs can be compared
 * element by element with the reference build in oracle/_ref:
 *   U(NV,-1:nx+1,-1:ny+1)  Res(NV,1:nx-1,1:ny-1)  dt(1:nx-1,1:ny-1)
 *   dU(NV,1:nx,1:ny)       Lci..Lvj, vol, si, sj, ni*, nj* (1:nx,1:ny)
 *   x, y (-1:nx+2,-1:ny+2)  x1, y1, Amu, Amu_t (-1:nx+1,-1:ny+1)
 *   dw (0:nx,0:ny)  Un(NVun,0:nx,0:ny)  QF(4,1:nx-1,1:ny-1)
 *   deltU(4,-1:nx+1,-1:ny+1)  Akx..Ajac (1:nx-1,1:ny-1)
 * (opencfd_ec2d_v1.5.4.f90:66-83, sub_init.f90:138-156, 291-303).
 */
#ifndef EC2D_ORACLE_H
#define EC2D_ORACLE_H

typedef struct {
  int nx, ny, nsub, fvm_fdm;
  const int *bc; /* 9 ints per subface: f_no face ist iend jst jend neighb subface orient */
  double *U, *Res, *dt, *dU, *Lci, *Lcj, *Lvi, *Lvj, *Amu, *Amu_t;
  const double *vol, *si, *sj, *ni1, *ni2, *nj1, *nj2, *x, *y, *x1, *y1, *dw;
  double *Un, *QF, *deltU;
  const double *Akx, *Aky, *Aix, *Aiy, *Ajac;
  double *Un1; /* (NVun,0:nx,0:ny), dual-time stepping only (sub_init.f90:154-156) */
} OBlock;

typedef struct {
  int nvar, scheme, flux, recon, turb, nblock, num_cell, nvar_un;
  OBlock *blocks;
  double res_max[6], res_rms[6];
} OMesh;

typedef struct {
  double Ma, Re, gamma, AoA, Pr, Cv, Cp, dt_global, CFL, dtmax, dtmin, P_outlet, T_inf, Twall, vt_inf,
      Kt_inf, Wt_inf;
  int If_viscous, Iflag_local_dt, Time_Method, Iflag_Scheme; /* Iflag_Scheme: the global one */
  double sfac; /* dual-time source factor of this residual call (sub_Residual.f90:28-34), Time_Method -1 */
} OParams;

typedef struct { int code, i, j, block; double T; } OError;

int o_residual(OMesh *M, const OParams *P, OError *err);
void o_lusgs_block(OMesh *M, const OParams *P, int mblock, double alfa1);
void o_update_lusgs(OMesh *M);
void o_bc(OMesh *M, const OParams *P);
void o_exchange(OMesh *M, const OParams *P);
void o_force_vt_kw(OMesh *M);
void o_euler1_update(OMesh *M);
void o_rk3_stage(OMesh *M, int krk);
void o_copy_un(OMesh *M);
void o_interp2h(OMesh *Mf, OMesh *Mc, int flag);
void o_prolong(OMesh *Mc, OMesh *Mf, int flag);
void o_force_function(OMesh *M);
void o_add_force(OMesh *M);
void o_delt_u(OMesh *M);
void o_set_un(OMesh *M);
void o_new_u(OMesh *M);
void o_dual_init(OMesh *M);
void o_dual_shift(OMesh *M);

/* wall forces of one mesh: F[0..3] = pressure (Fx, Fy), viscous (Fx, Fy) */
void o_forces(const OMesh *M, const OParams *P, double *F);

/* pointwise pieces, for closed-form tests of the oracle */
void o_pt_flux(int iflux, const double *QL, const double *QR, double *F, double g);
double o_pt_scheme(int side, const double *u, int flag);
void o_pt_recon(int kind, const double *U0, double *UL, double *UR, double g, int flag);
void o_pt_weno5(int side, int n, const double *v, double *h);
void o_pt_split_sw(double d, double u, double v, double cc, double A1, double A2, double g, double *fp, double *fm);
void o_pt_dfn(int NV, double *DF, const double *U, const double *DU, double n1, double n2, double g);

#endif
