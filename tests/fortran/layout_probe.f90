! Test infrastructure: the storage layout of the Fortran binding's
! ec2d_control (include/ec2d_bind.f90) as flang lays it out, for comparison
! with the C struct of include/ec2d.h (tests/fortran/layout_probe.c).
module layout_probe
  use iso_c_binding
  use ec2d_bind
  implicit none
contains
  ! n = c_sizeof(ec2d_control); offs(k) = byte offset of member k in
  ! declaration order (Pre_Step_Mesh counts as one member)
  subroutine fprobe_layout(n, offs) bind(C, name="fprobe_layout")
    integer(c_int), intent(out) :: n
    integer(c_int), intent(out) :: offs(32)
    type(ec2d_control), target :: c
    integer(c_intptr_t) :: b
    b = transfer(c_loc(c), b)
    n = int(c_sizeof(c), c_int)
    offs = -1
    offs(1) = off(c_loc(c%Ma)); offs(2) = off(c_loc(c%Re)); offs(3) = off(c_loc(c%gamma))
    offs(4) = off(c_loc(c%AoA)); offs(5) = off(c_loc(c%Pr)); offs(6) = off(c_loc(c%Cv))
    offs(7) = off(c_loc(c%Cp)); offs(8) = off(c_loc(c%t_end)); offs(9) = off(c_loc(c%dt_global))
    offs(10) = off(c_loc(c%CFL)); offs(11) = off(c_loc(c%dtmax)); offs(12) = off(c_loc(c%dtmin))
    offs(13) = off(c_loc(c%P_outlet)); offs(14) = off(c_loc(c%T_inf)); offs(15) = off(c_loc(c%Twall))
    offs(16) = off(c_loc(c%vt_inf)); offs(17) = off(c_loc(c%Kt_inf)); offs(18) = off(c_loc(c%Wt_inf))
    offs(19) = off(c_loc(c%If_viscous)); offs(20) = off(c_loc(c%Iflag_turbulence_model))
    offs(21) = off(c_loc(c%Iflag_init)); offs(22) = off(c_loc(c%Iflag_local_dt))
    offs(23) = off(c_loc(c%Time_Method)); offs(24) = off(c_loc(c%Iflag_Scheme))
    offs(25) = off(c_loc(c%Iflag_Flux)); offs(26) = off(c_loc(c%IFlag_Reconstruction))
    offs(27) = off(c_loc(c%Num_Mesh)); offs(28) = off(c_loc(c%Nvar)); offs(29) = off(c_loc(c%Pre_Step_Mesh))
    offs(30) = off(c_loc(c%Nstep_Inner_Limit)); offs(31) = off(c_loc(c%Res_Inner_Limit))
  contains
    integer(c_int) function off(p)
      type(c_ptr), intent(in) :: p
      off = int(transfer(p, b) - b, c_int)
    end function
  end subroutine
end module
