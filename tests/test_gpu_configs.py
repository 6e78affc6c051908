"""Every benchmark configuration at its stated size, GPU against the oracle.

The benchmark cases (bench.py --workload c1..c4) and the reference's own
sample case run here at full size with the launch shapes the benchmark uses,
and the device state is compared with the oracle's after each checked pass:

* C1  NACA 0012 C-grid 257x65, Euler, Roe + MUSCL3, LU-SGS CFL 5: 5 passes,
      bit-exact.
* C2  NACA 0012 C-grid 513x129, SA-RANS, FD-WENO5 block with Roe edge band,
      LU-SGS-smoothed 3-level V-cycle, with the autotuner's launch shape of
      the fine level (two lanes per cell, four workgroups per block) and
      through the autotuner and the HIP graph: 3 V-cycles, every level.
* C3  64 blocks in the seeded 8x8 layout (33..257 x 33..129, 664,602 cells),
      Euler AUSM + MUSCL3, LU-SGS, tuned, graph: 3 passes, bit-exact.
* C4  all 256 seeded SA-RANS airfoil cases (2,801,760 cells) in one batch,
      tuned (the fused LU-SGS launch), graph: 2 passes.
* C5  a held-out 513x129 airfoil with the learned closure as the
      eddy-viscosity model: the residual given the network's mu_t, bit-exact.
* The reference's RAE2822 sample (3 blocks, reversed wake-cut interface, SST,
  Van Leer, characteristic MUSCL3, LU-SGS CFL 10) from its golden fixture,
  whose oracle run tests/test_golden.py pins to the reference build: 3 passes.

Tolerances: bit-exact where no transcendental function enters; SA and SST
(device exp / pow / tanh, not glibc) 1e-11 relative per pass, the bar the
per-iteration residual is held to; residual norms (tree sums on the device,
serial in the reference) 1e-12 relative.
"""
import os

import numpy as np
import pytest

import oracle as O
from ec2d import cases
from ec2d import formats as F
from ec2d import solver
from helpers import case_dir, first_diff

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _oracle_for(case):
    """The oracle on the host geometry the device case was built from."""
    oc = O.OracleCase(control=case.ctl, levels=case.levels)
    oc.init_flow()
    return oc


def _check(b, refs, cases_, exact, rtol, what, levels=(1,)):
    for ci, (case, ref) in enumerate(zip(cases_, refs)):
        for nm in levels:
            for mb in range(1, len(case.levels[nm - 1]) + 1):
                g = b.get("U", ci, nm, mb)
                r = ref.block(nm, mb).U
                if exact:
                    diff = first_diff(g, r)
                    assert diff is None, f"{what}: case {ci} level {nm} block {mb} U: {diff}"
                else:
                    np.testing.assert_allclose(g, r, rtol=rtol, atol=1e-300,
                                               err_msg=f"{what}: case {ci} level {nm} block {mb} U")
        nv = case.ctl.nvar
        grms = b.res_norms(ci, 1)[1]
        np.testing.assert_allclose(grms[:nv], ref.norms(1)[1][:nv], rtol=1e-12 if exact else 1e-10, atol=0,
                                   err_msg=f"{what}: case {ci} residual rms")


def test_c1_naca0012_257x65_euler():
    x, y, bc, _ = cases.naca_cgrid(257, 65, ds1=1e-3, ratio=1.15)
    d = cases.write_case(case_dir("cfg_c1"), cases.control_euler_naca(iters=2000), [x], [y], [bc])
    case = solver.Case.from_dir(d)
    ref = _oracle_for(case)
    with solver.Batch([case]) as b:
        b.init_flow()
        b.autotune("", reps=2)
        for k in range(5):
            ref.main_step()
            b.main_step()
            _check(b, [ref], [case], True, 0, f"pass {k}")
            for a in ("Res", "dt", "Lci", "Lcj"):
                g, r = b.get(a), getattr(ref.block(), a)
                if a in ("Lci", "Lcj"):
                    g, r = g[:-1, :-1], r[:-1, :-1]
                assert first_diff(g, r) is None, f"pass {k} {a}"


def _c2_dir():
    x, y, bc, _ = cases.naca_cgrid(513, 129, ds1=1e-6, ratio=1.15)
    ctl = cases.control_sa_naca(iters=1000, Ma=0.73, aoa=2.0, Re=6.5e6, cfl=5.0, num_mesh=3,
                                time_method=F.TIME_LU_SGS, scheme=F.SCHEME_MUSCL3, flux=F.FLUX_ROE)
    return cases.write_case(case_dir("cfg_c2"), ctl, [x], [y], [bc], fdm_blocks=(1,))


def test_c2_naca0012_513x129_sa_mg3_benched_shape():
    """The benchmark's launch shape set by hand: residual tiles of 4 x 64
    cells, LU-SGS two lanes per cell, four workgroups per block on the fine
    level (128 rows), one on the coarse levels; eager V-cycles."""
    case = solver.Case.from_dir(_c2_dir())
    ref = _oracle_for(case)
    with solver.Batch([case], tile_rows=4) as b:
        b.init_flow()
        b.lusgs_shape(1, 2, 4)
        b.lusgs_shape(2, 2, 1)
        b.lusgs_shape(3, 2, 1)
        for k in range(3):
            ref.main_step()
            b.main_step()
            _check(b, [ref], [case], False, 1e-11, f"V-cycle {k}", levels=(1, 2, 3))
            np.testing.assert_allclose(b.get("Amu_t"), ref.block().Amu_t, rtol=1e-11, atol=1e-300)


def test_c2_naca0012_513x129_sa_mg3_tuned_graph():
    """As the benchmark runs it: the autotuner picks the shapes, then the
    V-cycles replay from one HIP graph."""
    case = solver.Case.from_dir(_c2_dir())
    ref = _oracle_for(case)
    with solver.Batch([case], tile_rows=4) as b:
        b.init_flow()
        rep = b.autotune("", reps=2)
        b.run(3)
        for _ in range(3):
            ref.main_step()
        _check(b, [ref], [case], False, 1e-11, "graph, tuned: " + rep.replace("\n", "; "), levels=(1, 2, 3))


def test_c3_vortex_64_blocks_full_layout():
    d = case_dir("cfg_c3")
    cases.write_vortex_case(d, cases.control_vortex(iters=1000), seed=0)
    case = solver.Case.from_dir(d)
    assert sum(g.ncell for g in case.levels[0]) == 664602
    ref = O.OracleCase(d)
    ref.init_flow()
    with solver.Batch([case]) as b:
        b.init_flow()
        b.autotune("", reps=2)
        ref.main_step()
        b.main_step()
        _check(b, [ref], [case], True, 0, "pass 0")
        b.run(2)
        for _ in range(2):
            ref.main_step()
        _check(b, [ref], [case], True, 0, "graph passes 1-2")


def test_c4_ensemble_256_cases():
    prms = cases.ensemble_params(seed=0, n=256)
    root = case_dir("cfg_c4")
    batch = [solver.Case.from_dir(cases.write_ensemble_case(os.path.join(root, f"case{k:03d}"), p))
             for k, p in enumerate(prms)]
    assert sum(g.ncell for c in batch for g in c.levels[0]) == 2801760
    refs = [_oracle_for(c) for c in batch]
    with solver.Batch(batch) as b:
        b.init_flow()
        b.autotune("", reps=2)
        b.run(2)
        for r in refs:
            r.main_step()
            r.main_step()
        _check(b, refs, batch, False, 1e-11, "2 passes")
        for ci in range(0, 256, 17):
            np.testing.assert_allclose(b.get("Amu_t", ci), refs[ci].block().Amu_t, rtol=1e-11, atol=1e-300)


def test_c5_closure_model4_513x129_held_out():
    """C5: the learned closure as the eddy-viscosity model (turbulence model
    4) on a held-out 513 x 129 C-grid case (the first seed-1 ensemble
    parameter set, as the closure study's held-out airfoils).  After five
    passes, given the state and the network's mu_t field, the residual, time
    step and spectral radii equal the oracle's bit for bit (the oracle takes
    Amu_t from the caller and applies SA's wall-ghost rule to it)."""
    import dataclasses

    from test_gpu_closure import _trained_params

    theta, norm = _trained_params("cfg_c5_train")
    prm = dict(cases.ensemble_params(seed=1, n=32)[0], nx=513, ny=129)
    d = cases.write_ensemble_case(case_dir("cfg_c5"), prm, iters=10)
    path = os.path.join(d, "control.in")
    F.write_control(path, dataclasses.replace(F.read_control(path), Iflag_turbulence_model=F.TURB_CLOSURE))
    case = solver.Case.from_dir(d)
    assert case.ctl.nvar == 4 and case.levels[0][0].nx == 513 and case.levels[0][0].ny == 129
    with solver.Batch([case]) as b:
        b.closure_set_params(theta, norm)
        b.init_flow()
        b.run(5)
        b.comput_residual_one_mesh(1)
        U = b.get("U")
        mut = b.get("closure_mut")
        got = {a: b.get(a) for a in ("Res", "dt", "Lci", "Lcj", "Lvi", "Lvj")}
    assert np.all(np.isfinite(mut[1:-1, 1:-1]))
    ref = O.OracleCase(d)
    rb = ref.block(1, 1)
    rb.U[...] = U
    rb.Amu_t[...] = mut
    ref.residual(1)
    for a, g in got.items():
        r = getattr(rb, a)
        if a in ("Lci", "Lcj", "Lvi", "Lvj"):
            g, r = g[:-1, :-1], r[:-1, :-1]
        diff = first_diff(g, r)
        assert diff is None, f"{a}: {diff}"


def test_reference_sample_rae2822(tmp_path):
    """The reference's shipped case (isothermal wall, see the fixture)."""
    from test_golden import load_case

    load_case(os.path.join(GOLDEN, "rae2822_sample_sst_isothermal.npz"), str(tmp_path))
    d = str(tmp_path)
    case = solver.Case.from_dir(d)
    assert case.ctl.Iflag_turbulence_model == F.TURB_SST and len(case.levels[0]) == 3
    ref = O.OracleCase(d)
    ref.init_flow()
    with solver.Batch([case]) as b:
        b.init_flow()
        for k in range(3):
            ref.main_step()
            b.main_step()
            _check(b, [ref], [case], False, 1e-11, f"pass {k}")
            for mb in (1, 2, 3):
                np.testing.assert_allclose(b.get("Amu_t", 0, 1, mb)[2:-2, 2:-2], ref.block(1, mb).Amu_t[2:-2, 2:-2],
                                           rtol=1e-11, atol=1e-300)
