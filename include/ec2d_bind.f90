! ec2d_bind.f90 - Fortran 2003 binding (ISO_C_BINDING) of include/ec2d.h.
!
! The reference's driver (OpenCFD-EC2D v1.5.4, Fortran) uses this module to
! call the MI355X path in place of its own subroutines; every interface below
! is one entry point of ec2d.h with the same name and argument order.
! Scalars pass by value, arrays as c_loc of their first element in the
! reference's own bounds (c_loc(B%U(1,-1,-1)), c_loc(B%vol(1,1)), ...), so
! nothing is repacked.  ec2d_check turns a nonzero status into the
! reference's error convention: print the message and stop.
!
! Build: flang -c ec2d_bind.f90 (the module file ec2d_bind.mod is what a
! caller's `use ec2d_bind` reads); link with -lec2d_hip.
module ec2d_bind
  use iso_c_binding
  implicit none

  ! control.in after read_parameter's conversions (opencfd_ec2d_v1.5.4.f90:197-235);
  ! the same members in the same order as the C struct ec2d_control
  type, bind(C) :: ec2d_control
    real(c_double) :: Ma, Re, gamma, AoA, Pr, Cv, Cp
    real(c_double) :: t_end, dt_global, CFL, dtmax, dtmin, P_outlet, T_inf, Twall, vt_inf, Kt_inf, Wt_inf
    integer(c_int) :: If_viscous, Iflag_turbulence_model, Iflag_init, Iflag_local_dt, Time_Method
    integer(c_int) :: Iflag_Scheme, Iflag_Flux, IFlag_Reconstruction, Num_Mesh, Nvar
    integer(c_int) :: Pre_Step_Mesh(3)
    integer(c_int) :: Nstep_Inner_Limit
    real(c_double) :: Res_Inner_Limit
  end type

  ! array selectors of ec2d_put_array / ec2d_get_array
  integer(c_int), parameter :: EC2D_U = 1, EC2D_RES = 2, EC2D_DT = 3, EC2D_DU = 4
  integer(c_int), parameter :: EC2D_LCI = 5, EC2D_LCJ = 6, EC2D_LVI = 7, EC2D_LVJ = 8
  integer(c_int), parameter :: EC2D_AMU = 9, EC2D_AMU_T = 10, EC2D_DW = 22, EC2D_UN = 23
  integer(c_int), parameter :: EC2D_QF = 24, EC2D_DELTU = 25, EC2D_UN1 = 31
  integer(c_int), parameter :: EC2D_VOL = 11, EC2D_SI = 12, EC2D_SJ = 13, EC2D_NI1 = 14, EC2D_NI2 = 15
  integer(c_int), parameter :: EC2D_NJ1 = 16, EC2D_NJ2 = 17, EC2D_X1 = 20, EC2D_Y1 = 21
  integer(c_int), parameter :: EC2D_CLOSURE_MUT = 40, EC2D_CLOSURE_MASK = 41

  interface
    ! ---- life cycle
    integer(c_int) function ec2d_init() bind(C, name="ec2d_init")
      import
    end function
    integer(c_int) function ec2d_finalize() bind(C, name="ec2d_finalize")
      import
    end function
    integer(c_int) function ec2d_last_error(buf, n) bind(C, name="ec2d_last_error")
      import
      character(kind=c_char) :: buf(*)
      integer(c_int), value :: n
    end function
    integer(c_int) function ec2d_case_create(ctl) bind(C, name="ec2d_case_create")
      import
      type(ec2d_control), intent(in) :: ctl
    end function
    integer(c_int) function ec2d_define_block(case_id, nMesh, mBlock, nx, ny, nsub, bc_msg, fvm_fdm, &
        x, y, x1, y1, vol, si, sj, ni1, ni2, nj1, nj2, dw, fd) bind(C, name="ec2d_define_block")
      import
      integer(c_int), value :: case_id, nMesh, mBlock, nx, ny, nsub, fvm_fdm
      integer(c_int), intent(in) :: bc_msg(*)
      type(c_ptr), value :: x, y, x1, y1, vol, si, sj, ni1, ni2, nj1, nj2, dw, fd
    end function
    integer(c_int) function ec2d_commit() bind(C, name="ec2d_commit")
      import
    end function
    integer(c_int) function ec2d_put_array(case_id, nMesh, mBlock, which, src) bind(C, name="ec2d_put_array")
      import
      integer(c_int), value :: case_id, nMesh, mBlock, which
      type(c_ptr), value :: src
    end function
    integer(c_int) function ec2d_get_array(case_id, nMesh, mBlock, which, dst) bind(C, name="ec2d_get_array")
      import
      integer(c_int), value :: case_id, nMesh, mBlock, which
      type(c_ptr), value :: dst
    end function

    ! ---- the per-iteration path
    integer(c_int) function ec2d_init_flow() bind(C, name="ec2d_init_flow")
      import
    end function
    integer(c_int) function ec2d_comput_residual_one_mesh(nMesh) bind(C, name="ec2d_comput_residual_one_mesh")
      import
      integer(c_int), value :: nMesh
    end function
    integer(c_int) function ec2d_du_lu_sgs_2d(nMesh, mBlock, alfa1) bind(C, name="ec2d_du_lu_sgs_2d")
      import
      integer(c_int), value :: nMesh, mBlock
      real(c_double), value :: alfa1
    end function
    integer(c_int) function ec2d_du_lu_sgs_2d_noalloc(nMesh, mBlock, alfa1) bind(C, name="ec2d_du_lu_sgs_2d_noalloc")
      import
      integer(c_int), value :: nMesh, mBlock
      real(c_double), value :: alfa1
    end function
    integer(c_int) function ec2d_update_u(nMesh) bind(C, name="ec2d_update_u")
      import
      integer(c_int), value :: nMesh
    end function
    integer(c_int) function ec2d_boundary_condition_onemesh(nMesh) bind(C, name="ec2d_boundary_condition_onemesh")
      import
      integer(c_int), value :: nMesh
    end function
    integer(c_int) function ec2d_update_buffer_onemesh(nMesh) bind(C, name="ec2d_update_buffer_onemesh")
      import
      integer(c_int), value :: nMesh
    end function
    integer(c_int) function ec2d_force_vt_kw(nMesh) bind(C, name="ec2d_force_vt_kw")
      import
      integer(c_int), value :: nMesh
    end function
    integer(c_int) function ec2d_ns_time_advance(nMesh) bind(C, name="ec2d_ns_time_advance")
      import
      integer(c_int), value :: nMesh
    end function
    integer(c_int) function ec2d_ns_multigrid() bind(C, name="ec2d_ns_multigrid")
      import
    end function
    integer(c_int) function ec2d_interpolation2h(m1, m2, flag) bind(C, name="ec2d_interpolation2h")
      import
      integer(c_int), value :: m1, m2, flag
    end function
    integer(c_int) function ec2d_prolong_u(m1, m2, flag) bind(C, name="ec2d_prolong_u")
      import
      integer(c_int), value :: m1, m2, flag
    end function
    integer(c_int) function ec2d_comput_force_function(nMesh) bind(C, name="ec2d_comput_force_function")
      import
      integer(c_int), value :: nMesh
    end function
    integer(c_int) function ec2d_add_force_function(nMesh) bind(C, name="ec2d_add_force_function")
      import
      integer(c_int), value :: nMesh
    end function
    integer(c_int) function ec2d_comput_delt_u(nMesh) bind(C, name="ec2d_comput_delt_u")
      import
      integer(c_int), value :: nMesh
    end function
    integer(c_int) function ec2d_set_un(nMesh) bind(C, name="ec2d_set_un")
      import
      integer(c_int), value :: nMesh
    end function
    integer(c_int) function ec2d_comput_new_u(nMesh) bind(C, name="ec2d_comput_new_u")
      import
      integer(c_int), value :: nMesh
    end function
    integer(c_int) function ec2d_run(n_iter) bind(C, name="ec2d_run")
      import
      integer(c_int), value :: n_iter
    end function
    integer(c_int) function ec2d_run_to_drop(drop, max_iter, check_every, iters, status, ms) &
        bind(C, name="ec2d_run_to_drop")
      import
      real(c_double), value :: drop
      integer(c_int), value :: max_iter, check_every
      integer(c_int) :: iters(*), status(*)
      real(c_double) :: ms(*)
    end function
    integer(c_int) function ec2d_res_norms(case_id, nMesh, rmax, rrms) bind(C, name="ec2d_res_norms")
      import
      integer(c_int), value :: case_id, nMesh
      real(c_double) :: rmax(6), rrms(6)
    end function
    integer(c_int) function ec2d_time(case_id, nMesh, kstep, tt) bind(C, name="ec2d_time")
      import
      integer(c_int), value :: case_id, nMesh
      integer(c_int) :: kstep
      real(c_double) :: tt
    end function
    integer(c_int) function ec2d_dual_inner(kt_in) bind(C, name="ec2d_dual_inner")
      import
      integer(c_int) :: kt_in
    end function
    integer(c_int) function ec2d_set_time(case_id, nMesh, kstep, tt) bind(C, name="ec2d_set_time")
      import
      integer(c_int), value :: case_id, nMesh, kstep
      real(c_double), value :: tt
    end function
    integer(c_int) function ec2d_get_run_counters(state) bind(C, name="ec2d_get_run_counters")
      import
      integer(c_long_long) :: state(3)
    end function
    integer(c_int) function ec2d_set_run_counters(state) bind(C, name="ec2d_set_run_counters")
      import
      integer(c_long_long) :: state(3)
    end function
    integer(c_int) function ec2d_synchronize() bind(C, name="ec2d_synchronize")
      import
    end function
    integer(c_long_long) function ec2d_num_cells(case_id, nMesh) bind(C, name="ec2d_num_cells")
      import
      integer(c_int), value :: case_id, nMesh
    end function
    integer(c_int) function ec2d_force_coefficients(case_id, cl, cd, forces) bind(C, name="ec2d_force_coefficients")
      import
      integer(c_int), value :: case_id
      real(c_double) :: cl, cd, forces(4)
    end function

    ! ---- launch shapes and measurement (no reference counterpart)
    integer(c_int) function ec2d_autotune(cache_path, reps, report, report_len) bind(C, name="ec2d_autotune")
      import
      character(kind=c_char) :: cache_path(*), report(*)
      integer(c_int), value :: reps, report_len
    end function
    integer(c_int) function ec2d_lusgs_shape(nMesh, lanes_per_cell, workgroups_per_block) &
        bind(C, name="ec2d_lusgs_shape")
      import
      integer(c_int), value :: nMesh, lanes_per_cell, workgroups_per_block
    end function
    integer(c_int) function ec2d_set_tile_rows(rows) bind(C, name="ec2d_set_tile_rows")
      import
      integer(c_int), value :: rows
    end function
    integer(c_int) function ec2d_set_tile_shape(cols, rows) bind(C, name="ec2d_set_tile_shape")
      import
      integer(c_int), value :: cols, rows
    end function
    integer(c_int) function ec2d_set_profile(on) bind(C, name="ec2d_set_profile")
      import
      integer(c_int), value :: on
    end function
    integer(c_int) function ec2d_set_wait_budget(polls) bind(C, name="ec2d_set_wait_budget")
      import
      integer(c_int), value :: polls
    end function
    integer(c_int) function ec2d_timing(run_ms, residual_ms, lusgs_ms, ghost_ms) bind(C, name="ec2d_timing")
      import
      real(c_double) :: run_ms, residual_ms, lusgs_ms, ghost_ms
    end function
    integer(c_int) function ec2d_timing_level(nMesh, residual_ms, lusgs_ms, ghost_ms, launches) &
        bind(C, name="ec2d_timing_level")
      import
      integer(c_int), value :: nMesh
      real(c_double) :: residual_ms, lusgs_ms, ghost_ms
      integer(c_int) :: launches(3)
    end function
    integer(c_int) function ec2d_kernel_time(nMesh, kind, ms, launches) bind(C, name="ec2d_kernel_time")
      import
      integer(c_int), value :: nMesh, kind
      real(c_double) :: ms
      integer(c_int) :: launches
    end function

    ! ---- learned turbulence closure (no reference counterpart)
    integer(c_int) function ec2d_closure_collect(nsamples) bind(C, name="ec2d_closure_collect")
      import
      integer(c_int) :: nsamples
    end function
    integer(c_int) function ec2d_closure_train(epochs, batch, lr, seed, first_epoch_loss, last_epoch_loss) &
        bind(C, name="ec2d_closure_train")
      import
      integer(c_int), value :: epochs, batch
      real(c_double), value :: lr
      integer(c_int), value :: seed
      real(c_double) :: first_epoch_loss, last_epoch_loss
    end function
    integer(c_int) function ec2d_closure_gradient(n, grad, loss) bind(C, name="ec2d_closure_gradient")
      import
      integer(c_int), value :: n
      real(c_float) :: grad(*)
      real(c_double) :: loss
    end function
    integer(c_int) function ec2d_closure_get_params(theta, norm) bind(C, name="ec2d_closure_get_params")
      import
      real(c_float) :: theta(*), norm(*)
    end function
    integer(c_int) function ec2d_closure_set_params(theta, norm) bind(C, name="ec2d_closure_set_params")
      import
      real(c_float), intent(in) :: theta(*), norm(*)
    end function
    integer(c_int) function ec2d_closure_infer() bind(C, name="ec2d_closure_infer")
      import
    end function
    integer(c_int) function ec2d_closure_enable(on) bind(C, name="ec2d_closure_enable")
      import
      integer(c_int), value :: on
    end function
    integer(c_int) function ec2d_closure_relaxation(w) bind(C, name="ec2d_closure_relaxation")
      import
      real(c_double), value :: w
    end function
    integer(c_int) function ec2d_closure_monitor(growth, active) bind(C, name="ec2d_closure_monitor")
      import
      real(c_double), value :: growth
      integer(c_int) :: active
    end function
  end interface

contains

  ! The reference's error convention: print the message and stop.
  subroutine ec2d_check(rc)
    integer(c_int), intent(in) :: rc
    character(kind=c_char) :: msg(512)
    character(len=512) :: text
    integer :: k, ios
    if (rc == 0) return
    ios = ec2d_last_error(msg, 512)
    text = ' '
    do k = 1, 512
      if (msg(k) == c_null_char) exit
      text(k:k) = msg(k)
    end do
    print *, 'ec2d: ', trim(text), ' (status ', rc, ')'
    stop 1
  end subroutine

end module ec2d_bind
