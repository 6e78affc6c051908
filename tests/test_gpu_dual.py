"""Dual-time LU-SGS (Time_Method -1, Dual_time_LU_SGS, sub_NS_singlegrid.f90:
231-310) on the GPU against the oracle, which tests/test_oracle_reference.py
pins bit for bit to the reference build.

Per physical step: U, Res, dt and the interior of Un and Un1 (the cells the
reference sets), the inner-loop index the reference prints, and the norms.
Euler must match bit for bit; SA to 1e-11 per step (the SA path evaluates exp
and pow with the GPU's math library).  The 3.5e-3 inner limit makes inner
loops exit early; the max_m Res_rms values the exit tests lie at least 7% from
it, so the GPU's tree-summed norms (1e-12 relative) take the same exits.
"""
import numpy as np
import pytest

import oracle as O
from ec2d import solver
from helpers import dual_case, first_diff

pytestmark = pytest.mark.gpu


def _check(b, ref, k, exact):
    rb = ref.block()
    for a in ("U", "Res", "dt", "Un", "Un1"):
        g, r = b.get(a), getattr(rb, a)
        if a in ("Un", "Un1"):
            g, r = g[1:-1, 1:-1], r[1:-1, 1:-1]
        if exact:
            diff = first_diff(g, r)
            assert diff is None, f"step {k} {a}: {diff}"
        else:
            np.testing.assert_allclose(g, r, rtol=1e-11, atol=1e-300, err_msg=f"step {k} {a}")
    assert b.dual_inner() == ref.inner_iterations[-1], f"step {k} inner iterations"
    gmax, grms = b.res_norms()
    rmax, rrms = ref.norms()
    nv = ref.ctl.nvar
    np.testing.assert_allclose(grms[:nv], rrms[:nv], rtol=1e-12 if exact else 1e-11, atol=0)


@pytest.mark.parametrize("sa,limit,fdm", [(False, 1e-30, ()), (False, 3.5e-3, ()), (True, 1e-30, ()), (True, 1e-30, (1,))])
def test_dual_time_steps(sa, limit, fdm):
    """The FD-WENO5 variant applies the source after the finite-difference
    interior of the residual, as the reference does."""
    n_inner = 8 if limit > 1e-20 else 6
    d = dual_case(f"g_dual_{int(sa)}_{int(limit > 1e-20)}_{len(fdm)}", sa, limit, n_inner, fdm)
    ref = O.OracleCase(d)
    ref.init_flow()
    with solver.Batch([solver.Case.from_dir(d)]) as b:
        b.init_flow()
        for k in range(3):
            ref.main_step()
            b.main_step()
            _check(b, ref, k, not sa)
        if limit > 1e-20:
            assert min(ref.inner_iterations) <= n_inner  # early exits happened


def test_dual_time_run_matches_steps():
    """ec2d_run steps dual time eagerly (its exit test reads the norms on the
    host): two passes through it equal two single steps of the oracle."""
    d = dual_case("g_dual_run", False, 3.5e-3, 8)
    ref = O.OracleCase(d)
    ref.init_flow()
    with solver.Batch([solver.Case.from_dir(d)]) as b:
        b.init_flow()
        b.run(2)
        b.synchronize()
        ref.main_step()
        ref.main_step()
        _check(b, ref, 1, True)
        assert b.time() == (2, ref.tt[0]) and ref.kstep[0] == 2
