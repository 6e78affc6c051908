! Test-infrastructure harness around the OpenCFD-EC2D reference build.
!
! This file is part of the ORACLE, not of the product: it lets tests (and the
! CPU-baseline leg of the benchmark) drive the unmodified reference solver
! (opencfd_ec2d_v1.5.4.f90 and the files it includes, compiled by flang from
! the reference tree) in-process, and read its arrays at full FP64 precision.
! The reference writes results only as 10-digit text (opencfd_ec2d_v1.5.4.f90:
! 266-279, 284-371), which is too coarse for parity checks.
!
! Apart from ref_op(3), which is the three-line `U = U + dU` loop that the
! reference keeps inline in NS_Time_advance_LU_SGS (sub_NS_singlegrid.f90:
! 203-211), nothing here re-implements solver arithmetic: every entry point
! calls a reference subroutine or copies one of the reference's module arrays.
! The reference's own `program main` (opencfd_ec2d_v1.5.4.f90:142-191) is
! compiled but never called; `ref_setup` makes the same set-up calls in the
! same order (opencfd_ec2d_v1.5.4.f90:151-165) and `ref_op(20,...)` performs
! one pass of its main loop body (opencfd_ec2d_v1.5.4.f90:174-180).
module ref_harness
  use iso_c_binding
  use Global_Var
  implicit none
contains

  ! Set-up sequence of the reference main program; the working directory must
  ! hold control.in, Mesh2d.dat and bc2d_1.in (and optionally FDM.in).
  subroutine ref_setup() bind(C, name="ref_setup")
    call read_parameter
    call check_mesh_multigrid
    call Init
    call set_control_para
    call Update_coordinate_buffer
    call Init_FiniteDifference
    if (Iflag_turbulence_model .eq. Turbulence_SST .or. &
        Iflag_turbulence_model .eq. Turbulence_SA) call comput_dw
    call Init_flow
  end subroutine

  integer(c_int) function ref_num_mesh() bind(C, name="ref_num_mesh")
    ref_num_mesh = Num_Mesh
  end function

  subroutine ref_mesh_info(nMesh, nblock, nvar, kstep, info) bind(C, name="ref_mesh_info")
    integer(c_int), value :: nMesh
    integer(c_int) :: nblock, nvar, kstep
    integer(c_int) :: info(4)
    nblock = Mesh(nMesh)%Num_Block
    nvar = Mesh(nMesh)%Nvar
    kstep = Mesh(nMesh)%Kstep
    info(1) = Mesh(nMesh)%Iflag_Scheme
    info(2) = Mesh(nMesh)%IFlag_flux
    info(3) = Mesh(nMesh)%IFlag_Reconstruction
    info(4) = Mesh(nMesh)%Iflag_turbulence_model
  end subroutine

  subroutine ref_block_info(nMesh, mBlock, nx, ny, nsub, fvm_fdm) bind(C, name="ref_block_info")
    integer(c_int), value :: nMesh, mBlock
    integer(c_int) :: nx, ny, nsub, fvm_fdm
    nx = Mesh(nMesh)%Block(mBlock)%nx
    ny = Mesh(nMesh)%Block(mBlock)%ny
    nsub = Mesh(nMesh)%Block(mBlock)%subface
    fvm_fdm = Mesh(nMesh)%Block(mBlock)%FVM_FDM
  end subroutine

  subroutine ref_bc_msg(nMesh, mBlock, ksub, v) bind(C, name="ref_bc_msg")
    integer(c_int), value :: nMesh, mBlock, ksub
    integer(c_int) :: v(9)
    type(BC_MSG_TYPE), pointer :: Bc
    Bc => Mesh(nMesh)%Block(mBlock)%bc_msg(ksub)
    v = [Bc%f_no, Bc%face, Bc%ist, Bc%iend, Bc%jst, Bc%jend, Bc%neighb, Bc%subface, Bc%orient]
  end subroutine

  ! Array selector shared by ref_size / ref_get / ref_set.
  subroutine pick(nMesh, mBlock, which, p2, p3)
    integer(c_int), value :: nMesh, mBlock, which
    real*8, pointer :: p2(:,:), p3(:,:,:)
    type(Block_TYPE), pointer :: B
    B => Mesh(nMesh)%Block(mBlock)
    nullify(p2); nullify(p3)
    select case (which)
    case (1);  p3 => B%U
    case (2);  p3 => B%Res
    case (3);  p2 => B%dt
    case (4);  p3 => B%dU
    case (5);  p2 => B%Lci
    case (6);  p2 => B%Lcj
    case (7);  p2 => B%Lvi
    case (8);  p2 => B%Lvj
    case (9);  p2 => B%Amu
    case (10); p2 => B%Amu_t
    case (11); p2 => B%vol
    case (12); p2 => B%si
    case (13); p2 => B%sj
    case (14); p2 => B%ni1
    case (15); p2 => B%ni2
    case (16); p2 => B%nj1
    case (17); p2 => B%nj2
    case (18); p2 => B%x
    case (19); p2 => B%y
    case (20); p2 => B%x1
    case (21); p2 => B%y1
    case (22); if (associated(B%dw)) p2 => B%dw
    case (23); if (associated(B%Un)) p3 => B%Un
    case (24); if (associated(B%QF)) p3 => B%QF
    case (25); if (associated(B%deltU)) p3 => B%deltU
    case (26); if (associated(B%Akx)) p2 => B%Akx
    case (27); if (associated(B%Aky)) p2 => B%Aky
    case (28); if (associated(B%Aix)) p2 => B%Aix
    case (29); if (associated(B%Aiy)) p2 => B%Aiy
    case (30); if (associated(B%Ajac)) p2 => B%Ajac
    case (31); if (associated(B%Un1)) p3 => B%Un1
    end select
  end subroutine

  integer(c_int) function ref_size(nMesh, mBlock, which) bind(C, name="ref_size")
    integer(c_int), value :: nMesh, mBlock, which
    real*8, pointer :: p2(:,:), p3(:,:,:)
    call pick(nMesh, mBlock, which, p2, p3)
    ref_size = 0
    if (associated(p2)) ref_size = size(p2)
    if (associated(p3)) ref_size = size(p3)
  end function

  ! Lower/upper bounds of the selected array (up to 3 dims), for the caller's reshape.
  subroutine ref_bounds(nMesh, mBlock, which, lb, ub, rank) bind(C, name="ref_bounds")
    integer(c_int), value :: nMesh, mBlock, which
    integer(c_int) :: lb(3), ub(3), rank
    real*8, pointer :: p2(:,:), p3(:,:,:)
    call pick(nMesh, mBlock, which, p2, p3)
    lb = 0; ub = -1; rank = 0
    if (associated(p2)) then
      rank = 2; lb(1:2) = lbound(p2); ub(1:2) = ubound(p2)
    end if
    if (associated(p3)) then
      rank = 3; lb = lbound(p3); ub = ubound(p3)
    end if
  end subroutine

  subroutine ref_get(nMesh, mBlock, which, buf, n) bind(C, name="ref_get")
    integer(c_int), value :: nMesh, mBlock, which, n
    real(c_double) :: buf(n)
    real*8, pointer :: p2(:,:), p3(:,:,:)
    call pick(nMesh, mBlock, which, p2, p3)
    if (associated(p2)) buf = reshape(p2, [n])
    if (associated(p3)) buf = reshape(p3, [n])
  end subroutine

  subroutine ref_set(nMesh, mBlock, which, buf, n) bind(C, name="ref_set")
    integer(c_int), value :: nMesh, mBlock, which, n
    real(c_double) :: buf(n)
    real*8, pointer :: p2(:,:), p3(:,:,:)
    call pick(nMesh, mBlock, which, p2, p3)
    if (associated(p2)) p2 = reshape(buf, shape(p2))
    if (associated(p3)) p3 = reshape(buf, shape(p3))
  end subroutine

  subroutine ref_res_norms(nMesh, rmax, rrms) bind(C, name="ref_res_norms")
    integer(c_int), value :: nMesh
    real(c_double) :: rmax(6), rrms(6)
    rmax = Mesh(nMesh)%Res_max
    rrms = Mesh(nMesh)%Res_rms
  end subroutine

  ! Control parameters (after read_parameter's unit conversions).
  subroutine ref_params(r, iv) bind(C, name="ref_params")
    real(c_double) :: r(20)
    integer(c_int) :: iv(20)
    r = 0.d0; iv = 0
    r(1) = Ma; r(2) = Re; r(3) = gamma; r(4) = AoA; r(5) = Pr; r(6) = Cv; r(7) = Cp
    r(8) = t_end; r(9) = dt_global; r(10) = CFL; r(11) = dtmax; r(12) = dtmin
    r(13) = P_outlet; r(14) = T_inf; r(15) = Twall; r(16) = vt_inf; r(17) = Kt_inf
    r(18) = Wt_inf; r(19) = Mesh(1)%tt; r(20) = Res_Inner_Limit
    iv(1) = Nvar; iv(2) = If_viscous; iv(3) = Iflag_turbulence_model; iv(4) = Iflag_init
    iv(5) = Iflag_local_dt; iv(6) = Time_Method; iv(7) = Iflag_Scheme; iv(8) = IFlag_flux
    iv(9) = IFlag_Reconstruction; iv(10) = Kstep_show; iv(11) = Num_Mesh
    iv(12) = Kstep_save; iv(13) = Nstep_Inner_Limit
  end subroutine

  ! Change numerical options after set-up so that one state can be run
  ! through every scheme / flux / reconstruction.  which: 1 scheme (mesh 1 and
  ! the global Iflag_Scheme), 2 flux, 3 reconstruction, 4 Time_Method,
  ! 5 global Iflag_Scheme only, 6 mesh-level scheme only (nMesh).
  subroutine ref_set_option(nMesh, which, val) bind(C, name="ref_set_option")
    integer(c_int), value :: nMesh, which, val
    select case (which)
    case (1); Iflag_Scheme = val; Mesh(nMesh)%Iflag_Scheme = val
    case (2); IFlag_flux = val; Mesh(nMesh)%IFlag_flux = val
    case (3); IFlag_Reconstruction = val; Mesh(nMesh)%IFlag_Reconstruction = val
    case (4); Time_Method = val
    case (5); Iflag_Scheme = val
    case (6); Mesh(nMesh)%Iflag_Scheme = val
    end select
  end subroutine

  subroutine ref_set_real(which, val) bind(C, name="ref_set_real")
    integer(c_int), value :: which
    real(c_double), value :: val
    select case (which)
    case (1); CFL = val
    case (2); dt_global = val
    case (3); dtmax = val
    case (4); dtmin = val
    end select
  end subroutine

  ! Call one reference subroutine.  a, b, c are its integer arguments.
  subroutine ref_op(op, a, b, c) bind(C, name="ref_op")
    integer(c_int), value :: op, a, b, c
    integer :: i, j, m
    type(Block_TYPE), pointer :: BK
    select case (op)
    case (1); call Comput_Residual_one_mesh(a)
    case (2); call du_LU_SGS_2D(a, b, real(c, 8) * 0.125d0)   ! alfa1 = c/8
    case (3)
      ! U(n+1) = U(n) + dU on block b of mesh a (sub_NS_singlegrid.f90:203-211)
      BK => Mesh(a)%Block(b)
      do j = 1, BK%ny - 1
        do i = 1, BK%nx - 1
          do m = 1, Mesh(a)%Nvar
            BK%U(m, i, j) = BK%U(m, i, j) + BK%dU(m, i, j)
          end do
        end do
      end do
    case (4); call Boundary_condition_onemesh(a)
    case (5); call update_buffer_onemesh(a)
    case (6); call force_vt_kw(a)
    case (7); call NS_Time_advance(a)
    case (8); call NS_2stge_multigrid
    case (9); call NS_3stge_multigrid
    case (10); call NS_Time_advance_1Euler(a)
    case (11); call NS_Time_advance_RK3(a)
    case (12); call interpolation2h(a, b, c)
    case (13); call prolong_U(a, b, c)
    case (14); call comput_force_function(a)
    case (15); call Add_force_function(a)
    case (16); call comput_delt_U(a)
    case (17); call Set_Un(a)
    case (18); call comput_new_U(a)
    case (19); call NS_Time_advance_LU_SGS(a)
    case (20)
      if (Num_Mesh .eq. 1) then
        call NS_Time_advance(1)
      else if (Num_Mesh .eq. 2) then
        call NS_2stge_multigrid
      else
        call NS_3stge_multigrid
      end if
    case (22); call output(a)
    case (23); call output_Res(a)
    case (21)
      ! run a iterations of the main loop body (timing helper; no output)
      do i = 1, a
        if (Num_Mesh .eq. 1) then
          call NS_Time_advance(1)
        else if (Num_Mesh .eq. 2) then
          call NS_2stge_multigrid
        else
          call NS_3stge_multigrid
        end if
      end do
    end select
  end subroutine

end module ref_harness
