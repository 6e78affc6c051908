"""Full-precision checkpoint (Batch.save_checkpoint / load_checkpoint over
ec2d_get_array, ec2d_set_time and ec2d_get/set_run_counters).

The property: a run interrupted after k steps, written to a checkpoint and
resumed in a fresh batch built from the same case directory continues bit for
bit as the uninterrupted run, on every level, with the same Kstep, tt and
norms.  The dual-time case covers the run counters: without them the resumed
run would take the first-call Sfac = 1/(3 dt) and alfa1 = 1/dt again
(sub_Residual.f90:28-34, sub_NS_singlegrid.f90:241-245) and diverge from the
uninterrupted run in the first step.
"""
import numpy as np
import pytest

from ec2d import cases
from ec2d import formats as F
from ec2d import solver
from helpers import dual_case, first_diff, naca_case

pytestmark = pytest.mark.gpu


def _state(b):
    out = {}
    c = b.cases[0]
    for lev in range(1, c.num_mesh + 1):
        for mb in range(1, len(c.levels[lev - 1]) + 1):
            out[(lev, mb)] = b.get("U", 0, lev, mb)
    return out


def _resume_matches(d, k, m, use_run, tmp_path):
    path = str(tmp_path / "state.npz")
    with solver.Batch([solver.Case.from_dir(d)]) as b:
        b.init_flow()
        for _ in range(k):
            b.main_step()
        b.save_checkpoint(path)
        if use_run:
            b.run(m)
        else:
            for _ in range(m):
                b.main_step()
        want, want_t, want_n, want_kt = _state(b), b.time(), b.res_norms(), b.dual_inner()
    with solver.Batch([solver.Case.from_dir(d)]) as b:
        b.init_flow()
        b.load_checkpoint(path)
        assert b.time() == (k, b.time()[1])
        if use_run:
            b.run(m)
        else:
            for _ in range(m):
                b.main_step()
        got = _state(b)
        for key in want:
            diff = first_diff(got[key], want[key])
            assert diff is None, f"level/block {key}: {diff}"
        assert b.time() == want_t
        for g, w in zip(b.res_norms(), want_n):
            assert first_diff(g, w) is None
        assert b.dual_inner() == want_kt


def test_checkpoint_sa_lusgs_multigrid(tmp_path):
    """SA with an FD-WENO5 block, LU-SGS-smoothed 3-level V-cycle, resumed
    through the HIP graph of ec2d_run."""
    c = cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5, num_mesh=3, time_method=F.TIME_LU_SGS)
    d = naca_case("g_ckpt_mg", c, ds1=1e-4, fdm=(1,))
    _resume_matches(d, 3, 4, True, tmp_path)


def test_checkpoint_euler_rk3(tmp_path):
    c = cases.control_euler_naca(iters=10)
    c.Time_Method = F.TIME_RK3
    c.CFL = 1.0
    d = naca_case("g_ckpt_rk3", c)
    _resume_matches(d, 2, 3, False, tmp_path)


def test_checkpoint_dual_time_sa(tmp_path):
    d = dual_case("g_ckpt_dual", True, 1e-30, 3)
    _resume_matches(d, 1, 2, False, tmp_path)


def test_checkpoint_refuses_other_cases(tmp_path):
    path = str(tmp_path / "state.npz")
    d1 = naca_case("g_ckpt_a", cases.control_euler_naca(iters=10))
    d2 = naca_case("g_ckpt_b", cases.control_euler_naca(iters=10), nx=49)
    with solver.Batch([solver.Case.from_dir(d1)]) as b:
        b.init_flow()
        b.main_step()
        b.save_checkpoint(path)
    with solver.Batch([solver.Case.from_dir(d2)]) as b:
        b.init_flow()
        with pytest.raises(ValueError, match="shape"):
            b.load_checkpoint(path)
    with np.load(path) as z:
        assert all(z[k].dtype != object for k in z.files)


@pytest.mark.parametrize("field,value", [("dt_global", 0.02), ("Time_Method", 1), ("Ma", 0.55)])
def test_checkpoint_refuses_another_run(field, value, tmp_path):
    """Same blocks, but a run-defining record of control.in differs (the
    physical step, the time method, the flow condition): refused by name."""
    import dataclasses

    path = str(tmp_path / "state.npz")
    c1 = cases.control_euler_naca(iters=10)
    d1 = naca_case("g_ckpt_run_a", c1)
    d2 = naca_case(f"g_ckpt_run_{field}", dataclasses.replace(c1, **{field: value}))
    with solver.Batch([solver.Case.from_dir(d1)]) as b:
        b.init_flow()
        b.main_step()
        b.save_checkpoint(path)
    with solver.Batch([solver.Case.from_dir(d2)]) as b:
        b.init_flow()
        with pytest.raises(ValueError, match=field):
            b.load_checkpoint(path)
