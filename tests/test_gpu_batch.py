"""Batches of the sizes the batching path is for, GPU against the oracle.

An 8 x 8 layout of 64 multi-block pieces of uneven sizes (Euler, AUSM,
MUSCL3, LU-SGS, vortex start) must match bit for bit; an ensemble of
independent SA airfoil cases of uneven grid sizes in one batch must match
each case's own oracle run to 1e-11 per step (the SA path evaluates exp and
pow with the GPU's math library), with and without the autotuned launch
shapes.  Uneven blocks in one launch exercise the per-block wave counts of
the LU-SGS sweep (the launch is sized for the tallest block; the waves past
a smaller block's rows sit out).
"""
import os

import numpy as np
import pytest

import oracle as O
from ec2d import cases
from ec2d import solver
from helpers import case_dir, first_diff

pytestmark = pytest.mark.gpu


def _check(b, refs, arrays, exact, rtol):
    for ci, ref in enumerate(refs):
        c = b.cases[ci]
        for mb in range(1, len(c.levels[0]) + 1):
            rb = ref.block(1, mb)
            for a in arrays:
                g = b.get(a, ci, 1, mb)
                r = getattr(rb, a)
                if exact:
                    diff = first_diff(g, r)
                    assert diff is None, f"case {ci} block {mb} {a}: {diff}"
                else:
                    np.testing.assert_allclose(g, r, rtol=rtol, atol=1e-300, err_msg=f"case {ci} block {mb} {a}")


def test_multiblock_8x8_uneven_layout():
    d = case_dir("g_c3_layout")
    info = cases.write_vortex_case(d, cases.control_vortex(iters=10), seed=3, nbx=8, nby=8, wmin=9, wmax=41,
                                   hmin=9, hmax=25, h=0.25)
    assert len(set(info["W"])) > 3 and len(set(info["H"])) > 3  # uneven
    ref = O.OracleCase(d)
    ref.init_flow()
    case = solver.Case.from_dir(d)
    assert len(case.levels[0]) == 64
    with solver.Batch([case]) as b:
        b.init_flow()
        _check(b, [ref], ("U",), True, 0)
        for _ in range(3):
            ref.main_step()
            b.main_step()
            _check(b, [ref], ("U", "Res", "dt"), True, 0)
        gmax, grms = b.res_norms()
        rmax, rrms = ref.norms()
        np.testing.assert_allclose(grms[:4], rrms[:4], rtol=1e-12, atol=0)
        np.testing.assert_array_equal(gmax[:4], rmax[:4])


@pytest.mark.parametrize("tuned", [False, True])
def test_ensemble_uneven_sa_cases(tuned, tmp_path):
    prms = cases.ensemble_params(seed=11, n=6)
    assert len({(p["nx"], p["ny"]) for p in prms}) > 2
    dirs = [cases.write_ensemble_case(case_dir(f"g_c4_{tuned}_{k}"), p) for k, p in enumerate(prms)]
    refs = []
    for d in dirs:
        r = O.OracleCase(d)
        r.init_flow()
        refs.append(r)
    with solver.Batch([solver.Case.from_dir(d) for d in dirs]) as b:
        b.init_flow()
        if tuned:
            b.autotune(str(tmp_path / "tune.txt"), reps=2)
        for _ in range(3):
            for r in refs:
                r.main_step()
            b.main_step()
            _check(b, refs, ("U", "Res", "dt"), False, 1e-11)
        for ci, r in enumerate(refs):
            gmax, grms = b.res_norms(ci)
            rmax, rrms = r.norms()
            np.testing.assert_allclose(grms[:5], rrms[:5], rtol=1e-11, atol=0)


def test_fused_run_restores_residual_planes():
    """Single-level fused LU-SGS passes write Res, dt and the spectral radii
    to the update's hyperplane copy alone; reading them (or any entry point
    that may) puts them back into their planes first.  Through graph capture,
    a later replay of the captured pass and a read of U alone in between, the
    planes must equal the oracle's after every run: 1e-11 relative to each
    plane's largest value (SA's device exp / pow differ from the oracle's at
    about 1e-16 absolute per step, so entries near zero have no relative
    bound after eight passes)."""
    prms = cases.ensemble_params(seed=5, n=3)
    dirs = [cases.write_ensemble_case(case_dir(f"g_sko_{k}"), p) for k, p in enumerate(prms)]
    refs = []
    for d in dirs:
        r = O.OracleCase(d)
        r.init_flow()
        refs.append(r)
    planes = ("Res", "dt", "Lci", "Lcj", "Lvi", "Lvj")
    with solver.Batch([solver.Case.from_dir(d) for d in dirs]) as b:
        b.lusgs_shape(1, 1, 0)
        b.init_flow()
        for n, u_only in ((4, False), (3, True), (2, False)):
            b.run(n)
            for r in refs:
                for _ in range(n):
                    r.main_step()
            if u_only:
                _check(b, refs, ("U",), False, 1e-11)
                b.run(1)
                for r in refs:
                    r.main_step()
            for ci, r in enumerate(refs):
                rb = r.block(1, 1)
                np.testing.assert_allclose(b.get("U", ci), rb.U, rtol=1e-11, atol=1e-300, err_msg=f"case {ci} U")
                for a in planes:
                    g, want = b.get(a, ci), getattr(rb, a)
                    if a.startswith("L"):
                        g, want = g[:-1, :-1], want[:-1, :-1]
                    scale = float(np.abs(want).max())
                    np.testing.assert_allclose(g, want, rtol=1e-11, atol=1e-11 * scale,
                                               err_msg=f"case {ci} {a} after run({n})")
