! Test infrastructure: a Fortran caller of the MI355X path through the
! ISO_C_BINDING module include/ec2d_bind.f90, as the reference's driver would
! call it (its arrays in the reference's bounds, passed by c_loc).
!
!   ec2d_fdriver IN OUT
!
! IN (stream, written by tests/test_gpu_fortran.py): int32 nblock, nsteps,
! mode (1: nsteps calls of ec2d_ns_time_advance(1); 2: one ec2d_run(nsteps));
! the ec2d_control members in order; per block int32 nx, ny, nsub, fvm_fdm,
! has_dw, has_u0, bc_msg(9*nsub), then x, y (-1:nx+2,-1:ny+2), x1, y1
! (-1:nx+1,-1:ny+1), vol, si, sj, ni1, ni2, nj1, nj2 (1:nx,1:ny), when
! has_dw dw (0:nx,0:ny), and when has_u0 the initial U(Nvar,-1:nx+1,-1:ny+1)
! of a restart (Iflag_init = 1), column-major.
! OUT (stream): per block U(Nvar,-1:nx+1,-1:ny+1), then Res_max(6),
! Res_rms(6) of mesh 1 and Kstep.
program ec2d_fdriver
  use iso_c_binding
  use ec2d_bind
  implicit none
  type blk
    integer(c_int) :: nx, ny, nsub, fvm, has_dw, has_u0
    integer(c_int), allocatable :: bc(:)
    real(c_double), allocatable :: x(:,:), y(:,:), x1(:,:), y1(:,:), vol(:,:), si(:,:), sj(:,:)
    real(c_double), allocatable :: ni1(:,:), ni2(:,:), nj1(:,:), nj2(:,:), dw(:,:), U(:,:,:)
  end type
  type(blk), allocatable, target :: bl(:)
  type(ec2d_control), target :: ctl
  character(len=1024) :: fin, fout
  integer(c_int) :: nblock, nsteps, mode, cid, k, n, kstep
  real(c_double) :: rmax(6), rrms(6), tt
  type(c_ptr) :: pdw

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) nblock, nsteps, mode
  read(10) ctl%Ma, ctl%Re, ctl%gamma, ctl%AoA, ctl%Pr, ctl%Cv, ctl%Cp
  read(10) ctl%t_end, ctl%dt_global, ctl%CFL, ctl%dtmax, ctl%dtmin, ctl%P_outlet, ctl%T_inf, ctl%Twall, &
           ctl%vt_inf, ctl%Kt_inf, ctl%Wt_inf
  read(10) ctl%If_viscous, ctl%Iflag_turbulence_model, ctl%Iflag_init, ctl%Iflag_local_dt, ctl%Time_Method
  read(10) ctl%Iflag_Scheme, ctl%Iflag_Flux, ctl%IFlag_Reconstruction, ctl%Num_Mesh, ctl%Nvar
  read(10) ctl%Pre_Step_Mesh, ctl%Nstep_Inner_Limit, ctl%Res_Inner_Limit
  allocate(bl(nblock))
  do k = 1, nblock
    associate (b => bl(k))
      read(10) b%nx, b%ny, b%nsub, b%fvm, b%has_dw, b%has_u0
      allocate(b%bc(9 * b%nsub))
      read(10) b%bc
      allocate(b%x(-1:b%nx+2, -1:b%ny+2), b%y(-1:b%nx+2, -1:b%ny+2))
      allocate(b%x1(-1:b%nx+1, -1:b%ny+1), b%y1(-1:b%nx+1, -1:b%ny+1))
      allocate(b%vol(b%nx, b%ny), b%si(b%nx, b%ny), b%sj(b%nx, b%ny), b%ni1(b%nx, b%ny), b%ni2(b%nx, b%ny))
      allocate(b%nj1(b%nx, b%ny), b%nj2(b%nx, b%ny))
      allocate(b%U(ctl%Nvar, -1:b%nx+1, -1:b%ny+1))
      read(10) b%x, b%y, b%x1, b%y1, b%vol, b%si, b%sj, b%ni1, b%ni2, b%nj1, b%nj2
      if (b%has_dw /= 0) then
        allocate(b%dw(0:b%nx, 0:b%ny))
        read(10) b%dw
      end if
      if (b%has_u0 /= 0) read(10) b%U
    end associate
  end do
  close(10)

  ! set-up (opencfd_ec2d_v1.5.4.f90:151-165 with the device in place of Init's arrays)
  call ec2d_check(ec2d_init())
  cid = ec2d_case_create(ctl)
  if (cid < 0) call ec2d_check(cid)
  do k = 1, nblock
    pdw = c_null_ptr
    if (bl(k)%has_dw /= 0) pdw = c_loc(bl(k)%dw(0, 0))
    call ec2d_check(ec2d_define_block(cid, 1, k, bl(k)%nx, bl(k)%ny, bl(k)%nsub, bl(k)%bc, bl(k)%fvm, &
        c_loc(bl(k)%x(-1, -1)), c_loc(bl(k)%y(-1, -1)), c_loc(bl(k)%x1(-1, -1)), c_loc(bl(k)%y1(-1, -1)), &
        c_loc(bl(k)%vol(1, 1)), c_loc(bl(k)%si(1, 1)), c_loc(bl(k)%sj(1, 1)), c_loc(bl(k)%ni1(1, 1)), &
        c_loc(bl(k)%ni2(1, 1)), c_loc(bl(k)%nj1(1, 1)), c_loc(bl(k)%nj2(1, 1)), pdw, c_null_ptr))
  end do
  call ec2d_check(ec2d_commit())
  ! a restart state goes in before Init_flow, which then fills its ghosts
  do k = 1, nblock
    if (bl(k)%has_u0 /= 0) call ec2d_check(ec2d_put_array(cid, 1, k, EC2D_U, c_loc(bl(k)%U(1, -1, -1))))
  end do
  call ec2d_check(ec2d_init_flow())

  ! the main loop (opencfd_ec2d_v1.5.4.f90:172-186)
  if (mode == 1) then
    do n = 1, nsteps
      call ec2d_check(ec2d_ns_time_advance(1))
    end do
  else
    call ec2d_check(ec2d_run(nsteps))
  end if
  call ec2d_check(ec2d_res_norms(cid, 1, rmax, rrms))
  call ec2d_check(ec2d_time(cid, 1, kstep, tt))

  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  do k = 1, nblock
    call ec2d_check(ec2d_get_array(cid, 1, k, EC2D_U, c_loc(bl(k)%U(1, -1, -1))))
    write(11) bl(k)%U
  end do
  write(11) rmax, rrms, kstep
  close(11)
  call ec2d_check(ec2d_finalize())
  print *, 'fdriver OK', nsteps, kstep
end program
