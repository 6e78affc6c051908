"""Per-case runs to a residual drop (ec2d_run_to_drop) and the error paths.

A batch of cases runs until each case's density Res_rms has dropped by the
requested factor from its first value; a case that has stops and stays frozen
while the others go on.  Each case's frozen state must equal the oracle's run
of that case alone for exactly the number of updates the device reports, and
the oracle must confirm that this is the first pass whose residual meets the
drop.  The NaN guard stops a case whose residual is not finite without
disturbing the others; a case that hits the pass limit stops with its own
status.  The reference's T < 1e-5 stop is reported with the reference's cell.
"""
import numpy as np
import pytest

import oracle as O
from ec2d import cases
from ec2d import formats as F
from ec2d import solver
from helpers import first_diff, naca_case, vortex_case

pytestmark = pytest.mark.gpu


def _ud1(Ma, aoa, cfl=20.0):
    c = cases.control_euler_naca(iters=10, Ma=Ma, aoa=aoa, cfl=cfl)
    c.Iflag_Scheme = F.SCHEME_UD1
    return c


def _oracle_history(d, n_checks):
    """Density Res_rms of the first n_checks residuals of the oracle's run and
    the oracle case after n_checks - 1 updates."""
    oc = O.OracleCase(d)
    oc.init_flow()
    rms = []
    for k in range(n_checks):
        oc.residual(1)
        rms.append(oc.norms(1)[1][0])
        if k < n_checks - 1:
            for mb in range(1, len(oc.meshes[0]["blocks"]) + 1):
                oc.lusgs(1, mb, 0.0)
            oc.update(1)
            oc.bc(1)
            oc.exchange(1)
    return oc, np.array(rms)


def test_run_to_drop_per_case_masks():
    drop = 1e-4
    ds = [naca_case(f"g_conv_{k}", _ud1(Ma, aoa)) for k, (Ma, aoa) in enumerate([(0.5, 1.25), (0.3, 3.0), (0.6, 0.0)])]
    batch = [solver.Case.from_dir(d) for d in ds]
    with solver.Batch(batch) as b:
        b.init_flow()
        iters, status, ms = b.run_to_drop(drop, max_iter=4000, check_every=50)
        states = [b.get("U", k) for k in range(3)]
        ksteps = [b.time(k)[0] for k in range(3)]
        rms = [b.res_norms(k)[1][0] for k in range(3)]
    assert status == [solver.Batch.CONVERGED] * 3, (iters, status)
    assert len(set(iters)) > 1, f"the cases should need different pass counts: {iters}"
    assert all(m > 0 for m in ms) and max(ms) == ms[iters.index(max(iters))]
    for k, d in enumerate(ds):
        oc, hist = _oracle_history(d, iters[k])
        # the device stopped at the first residual that meets the drop
        assert hist[-1] <= drop * hist[0], (k, hist[-1] / hist[0])
        assert (hist[:-1] > drop * hist[0]).all(), (k, np.argmax(hist[:-1] <= drop * hist[0]))
        diff = first_diff(states[k], oc.block().U)
        assert diff is None, f"case {k} frozen after {iters[k] - 1} updates: {diff}"
        assert ksteps[k] == iters[k] - 1
        np.testing.assert_allclose(rms[k], hist[-1], rtol=1e-12)


def test_run_to_drop_limit_and_nan_guard():
    """Case 0 cannot reach the drop within the pass limit (status 3), case 1
    starts with a NaN in one interior cell (status 2 at the first check, state
    untouched), case 2 converges; cases 0 and 2 equal their oracle runs."""
    ds = [naca_case(f"g_lim_{k}", _ud1(Ma, aoa)) for k, (Ma, aoa) in enumerate([(0.5, 1.25), (0.4, 2.0), (0.3, 3.0)])]
    batch = [solver.Case.from_dir(d) for d in ds]
    with solver.Batch(batch) as b:
        b.init_flow()
        U1 = b.get("U", 1)
        U1[5, 9, 0] = np.nan
        b.put("U", U1, 1)
        iters, status, ms = b.run_to_drop(1e-12, max_iter=300, check_every=64)
        assert status[0] == solver.Batch.LIMIT and iters[0] == 300, (iters, status)
        assert status[1] == solver.Batch.NOT_FINITE and iters[1] == 1, (iters, status)
        assert status[2] == solver.Batch.LIMIT
        assert first_diff(b.get("U", 1), U1) is None, "the NaN case must stay frozen"
        for k in (0, 2):
            oc, _ = _oracle_history(ds[k], iters[k])
            assert first_diff(b.get("U", k), oc.block().U) is None, k
        # after the run every case runs again: a plain pass moves case 0
        before = b.get("U", 0)
        b.run(1)
        assert not np.array_equal(before, b.get("U", 0))


def test_cold_cell_reports_reference_cell():
    """A cell with T < 1e-5 in block 2 (two cold cells, the first in the
    reference's loop order is reported): status 2 with the message of the
    reference's stop (sub_Residual.f90:143-146), as the oracle gives it."""
    d = vortex_case("g_cold", cases.control_vortex(iters=10))
    case = solver.Case.from_dir(d)
    ref = O.OracleCase(d)
    ref.init_flow()
    with solver.Batch([case]) as b:
        b.init_flow()
        U = b.get("U", 0, 1, 2)
        for (i, j) in ((7, 4), (3, 5)):
            q = U[j + 1, i + 1]
            q[3] = 0.5 * (q[1] ** 2 + q[2] ** 2) / q[0]  # internal energy 0: T = 0
        b.put("U", U, 0, 1, 2)
        ref.block(1, 2).U[...] = U
        b.comput_residual_one_mesh(1)
        from ec2d import capi

        with pytest.raises(capi.Ec2dError) as e:
            b.res_norms()
        with pytest.raises(O.SolverError) as eo:
            ref.residual(1)
    assert "T <1.d-5 ! 7 4 2" in str(e.value), str(e.value)
    assert "status 2" in str(e.value)
    assert str(eo.value).startswith("T <1.d-5 ! 7 4 2"), str(eo.value)


def test_run_to_drop_cold_cell_stops_only_its_case():
    """A cell with T < 1e-5 in case 1 (internal energy 0): in a per-case run
    the reference's stop ends that case alone (status 4 at the first check,
    state untouched) while cases 0 and 2 run on and equal their oracle runs;
    a plain pass afterwards, outside the per-case mode, stops as the
    reference does (status 2, the cell in the message)."""
    from ec2d import capi

    ds = [naca_case(f"g_cold_{k}", _ud1(Ma, aoa)) for k, (Ma, aoa) in enumerate([(0.5, 1.25), (0.4, 2.0), (0.3, 3.0)])]
    batch = [solver.Case.from_dir(d) for d in ds]
    with solver.Batch(batch) as b:
        b.init_flow()
        U1 = b.get("U", 1)
        q = U1[5, 9]
        q[3] = 0.5 * (q[1] ** 2 + q[2] ** 2) / q[0]
        b.put("U", U1, 1)
        iters, status, ms = b.run_to_drop(1e-12, max_iter=200, check_every=50)
        assert status[1] == solver.Batch.COLD and iters[1] == 1, (iters, status)
        assert status[0] == solver.Batch.LIMIT and status[2] == solver.Batch.LIMIT, (iters, status)
        assert first_diff(b.get("U", 1), U1) is None, "the cold case must stay frozen"
        for k in (0, 2):
            oc, _ = _oracle_history(ds[k], iters[k])
            assert first_diff(b.get("U", k), oc.block().U) is None, k
        with pytest.raises(capi.Ec2dError, match=r"T <1.d-5 ! 8 4 1 \(case 1\)"):
            b.run(1)


def test_sa_converged_cl_cd_match_oracle():
    """Converged forces to 1e-8: an SA-RANS C-grid case (65 x
    17, Re 1e5, M 0.5) run on the device to a 1e-8 density drop (about 12k
    passes) has Cl and Cd within 1e-8 relative of the oracle's run of the same
    number of updates (the device evaluates SA's exp / pow with its own math
    library, 1e-11 relative per step, so the two runs are not bit-identical:
    the converged fixed point is what they share)."""
    c = cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5)
    d = naca_case("g_conv_sa", c, ds1=1e-4)
    with solver.Batch([solver.Case.from_dir(d)]) as b:
        b.init_flow()
        iters, status, _ = b.run_to_drop(1e-8, max_iter=20000, check_every=200)
        cl_g, cd_g, _ = b.force_coefficients(0)
    assert status == [solver.Batch.CONVERGED], (iters, status)
    oc = O.OracleCase(d)
    oc.init_flow()
    for _ in range(iters[0] - 1):
        oc.main_step()
    cl_o, cd_o = oc.force_coefficients()[:2]
    assert abs(cl_g - cl_o) <= 1e-8 * abs(cl_o), (cl_g, cl_o)
    assert abs(cd_g - cd_o) <= 1e-8 * abs(cd_o), (cd_g, cd_o)


def test_nan_guard_sees_a_turbulence_equation(tmp_path):
    """The reference's RAE2822 sample (SST, 3 blocks; isothermal wall, as the
    fixture states) diverges in its own numerics: density stays near 4e-5
    while the omega residual grows from the third pass and overflows, step for
    step in the reference build and the oracle (tests/test_oracle_reference.py
    ::test_rae2822_sample_omega_overflow).  The NaN guard must stop it with
    status NOT_FINITE at the first check whose turbulence rms is not finite,
    the oracle's check, although its density rms is finite there, and freeze
    the case with the state that check saw: that of a plain run of
    iters - 1 passes (NaN cells included)."""
    import os

    from test_golden import load_case

    load_case(os.path.join(os.path.dirname(__file__), "golden", "rae2822_sample_sst_isothermal.npz"), str(tmp_path))
    d = str(tmp_path)
    oc = O.OracleCase(d)
    oc.init_flow()
    first_bad = None
    for k in range(1, 41):
        oc.main_step()
        rms = oc.norms(1)[1]
        if not np.all(np.isfinite(rms)):
            first_bad = k
            break
    assert first_bad is not None and np.isfinite(rms[0]), (first_bad, rms)
    with solver.Batch([solver.Case.from_dir(d)]) as b:
        b.init_flow()
        iters, status, _ = b.run_to_drop(1e-8, max_iter=200, check_every=5)
        assert status == [solver.Batch.NOT_FINITE] and iters == [first_bad], (iters, status, first_bad)
        U = [b.get("U", 0, 1, mb) for mb in (1, 2, 3)]
    with solver.Batch([solver.Case.from_dir(d)]) as b:
        b.init_flow()
        b.run(first_bad - 1)
        for mb in (1, 2, 3):
            assert first_diff(b.get("U", 0, 1, mb), U[mb - 1]) is None, f"block {mb}"
