"""Wall-force integration of the oracle (CPU).

The reference has no force integration, so the definition (pressure of the
wall-adjacent cell over the wall faces plus the viscous traction from
Residual_FVM's face-gradient stencil) is pinned here by:
* a uniform state: no net pressure or viscous force on a closed body;
* an independent integral of the pressure over the wall segments built from
  the node coordinates (not from the stored face normals);
* the rotation into lift and drag for a state rotated with the angle of attack.
"""
import math

import numpy as np

import oracle as O
from ec2d import cases
from ec2d import formats as F
from ec2d import geometry as G


def naca_oracle(ctl, nx=65, ny=17, ds1=1e-3):
    x, y, bc, info = cases.naca_cgrid(nx, ny, ds1=ds1)
    need_dw = ctl.Iflag_turbulence_model in (F.TURB_SA, F.TURB_SST)
    oc = O.OracleCase(control=ctl, levels=G.build_levels([x], [y], [bc], 1, need_dw=need_dw))
    return oc, x, y, bc


def test_uniform_state_gives_no_force():
    c = cases.control_sa_naca(iters=1, Re=1e5, Ma=0.5)
    c.Iflag_turbulence_model = F.TURB_NONE
    oc, x, y, bc = naca_oracle(c)
    oc.init_flow()
    b = oc.block()
    # the freestream in every cell, ghosts included (the wall condition would
    # mirror the velocity into the ghosts and shear the wall faces)
    b.U[...] = b.U[b.ny // 2, b.nx // 2]
    b.Amu[...] = 1e-5
    Fv = oc.forces()
    assert np.all(np.abs(Fv[:2]) < 1e-13), Fv
    assert np.all(Fv[2:] == 0.0), Fv


def test_pressure_force_from_node_coordinates():
    c = cases.control_euler_naca(iters=5)
    oc, x, y, bc = naca_oracle(c)
    oc.init_flow()
    for _ in range(5):
        oc.main_step()
    b = oc.block()
    g, Ma = c.gamma, c.Ma
    wall = [s for s in bc if s[6] == F.BC_WALL][0]
    assert wall[1] == 2
    fx = fy = 0.0
    for i in range(wall[2], wall[3]):
        # segment from node (i, 1) to node (i+1, 1); the fluid lies on its left
        dx, dy = x[0, i] - x[0, i - 1], y[0, i] - y[0, i - 1]
        U = b.U[2, i + 1]  # interior cell (i, 1): numpy index (j+1, i+1)
        d, u, v = U[0], U[1] / U[0], U[2] / U[0]
        p = (g - 1.0) * (U[3] - 0.5 * d * (u * u + v * v))
        dp = p - 1.0 / (g * Ma * Ma)
        # -(p - p_inf) n s with n s = (-dy, dx)
        fx += dp * dy
        fy -= dp * dx
    F4 = oc.forces()
    np.testing.assert_allclose(F4[:2], [fx, fy], rtol=1e-9, atol=1e-14)
    assert np.all(F4[2:] == 0.0)


def test_lift_drag_rotation():
    c = cases.control_euler_naca(iters=5, aoa=3.0)
    oc, x, y, bc = naca_oracle(c)
    oc.init_flow()
    for _ in range(3):
        oc.main_step()
    Fx, Fy = oc.forces()[:2]
    cl, cd = oc.force_coefficients()
    a = math.radians(3.0)
    assert cl == (-Fx * math.sin(a) + Fy * math.cos(a)) / 0.5
    assert cd == (Fx * math.cos(a) + Fy * math.sin(a)) / 0.5
    # the resultant is invariant under the rotation
    assert math.isclose(math.hypot(cl, cd), math.hypot(Fx, Fy) / 0.5, rel_tol=1e-13)
