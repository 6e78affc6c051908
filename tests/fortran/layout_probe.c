/* Test infrastructure: the C layout of ec2d_control (include/ec2d.h), member
 * offsets in declaration order, for comparison with the Fortran binding. */
#include <stddef.h>
#include "../../include/ec2d.h"

void cprobe_layout(int *n, int *offs) {
  int k = 0;
  *n = (int)sizeof(ec2d_control);
#define O(m) offs[k++] = (int)offsetof(ec2d_control, m)
  O(Ma); O(Re); O(gamma); O(AoA); O(Pr); O(Cv); O(Cp); O(t_end); O(dt_global); O(CFL); O(dtmax); O(dtmin);
  O(P_outlet); O(T_inf); O(Twall); O(vt_inf); O(Kt_inf); O(Wt_inf); O(If_viscous); O(Iflag_turbulence_model);
  O(Iflag_init); O(Iflag_local_dt); O(Time_Method); O(Iflag_Scheme); O(Iflag_Flux); O(IFlag_Reconstruction);
  O(Num_Mesh); O(Nvar); O(Pre_Step_Mesh); O(Nstep_Inner_Limit); O(Res_Inner_Limit);
#undef O
  while (k < 32) offs[k++] = -1;
}
