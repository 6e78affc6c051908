"""Baldwin-Lomax (turbulence model 1, sub_turbulence_BL.f90) on the GPU
against the oracle, which tests/test_oracle_reference.py pins bit for bit to
the reference build.

k_bl computes mu_t ahead of the residual kernel.  The model evaluates exp
with the GPU's math library, so U, Res, dt and Amu_t must match to 1e-11 per
step (relative; absolute 1e-300), like the SA path.  Cases: the C-grid airfoil
(its wall ends inside face 2, so the ghost copies reach the first wake
columns) and a multi-block layout with walls on two faces of a block (the
smaller mu_t of two wall lines wins).
"""
import numpy as np
import pytest

import oracle as O
from ec2d import cases
from ec2d import formats as F
from ec2d import solver
from helpers import naca_case, vortex_case

pytestmark = pytest.mark.gpu


def _walls(bcs):
    for subs in bcs:
        for s in subs:
            if s[6] == F.BC_FARFIELD and s[1] in (1, 2):
                s[6] = F.BC_WALL


def _case(kind):
    if kind == "mg2":
        c = cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5)
        c.Iflag_turbulence_model, c.Time_Method, c.Num_Mesh, c.CFL = F.TURB_BL, F.TIME_RK3, 2, 0.3
        return naca_case("g_bl_mg", c, ds1=1e-4)
    if kind == "naca":
        c = cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5)
        c.Iflag_turbulence_model = F.TURB_BL
        return naca_case("g_bl", c, ds1=1e-4)
    c = cases.control_vortex(iters=10)
    c.If_viscous, c.Re, c.Iflag_turbulence_model = 1, 1e4, F.TURB_BL
    return vortex_case("g_bl_mb", c, bc_edit=_walls)


def _check(b, ref, k):
    nblock = len(b.cases[0].levels[0])
    for mb in range(1, nblock + 1):
        rb = ref.block(1, mb)
        for a in ("U", "Res", "dt", "Amu_t"):
            np.testing.assert_allclose(b.get(a, 0, 1, mb), getattr(rb, a), rtol=1e-11, atol=1e-300,
                                       err_msg=f"step {k} block {mb} {a}")


@pytest.mark.parametrize("kind", ["naca", "two_walls", "mg2"])
def test_baldwin_lomax_steps(kind):
    d = _case(kind)
    ref = O.OracleCase(d)
    ref.init_flow()
    with solver.Batch([solver.Case.from_dir(d)]) as b:
        b.init_flow()
        for k in range(3):
            ref.main_step()
            b.main_step()
            _check(b, ref, k)
        assert np.abs(b.get("Amu_t")).max() > 0.0


def test_baldwin_lomax_graph_run():
    """The model inside the run graph: three replayed passes equal three oracle steps."""
    d = _case("naca")
    ref = O.OracleCase(d)
    ref.init_flow()
    with solver.Batch([solver.Case.from_dir(d)]) as b:
        b.init_flow()
        b.run(3)
        b.synchronize()
        for _ in range(3):
            ref.main_step()
        _check(b, ref, 2)
