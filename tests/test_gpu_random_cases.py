"""Seeded random multi-block cases on the GPU against the oracle
(tests/helpers.py random_case: layout, boundary type per subface, scheme,
flux, reconstruction, viscosity, wall temperature, time method, back
pressure).  The same seeds run oracle against the reference build in
tests/test_oracle_reference.py::test_random_cases.  Bar: bit-exact U, Res,
dt and spectral radii after every pass; norms as in test_gpu_parity.py."""
import pytest

from helpers import random_case
from test_gpu_parity import _compare_steps

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed", range(24))
def test_random_case(seed):
    d, n = random_case(seed, "g")
    _compare_steps(d, n)


@pytest.mark.parametrize("seed", range(12))
def test_random_turb_multiblock_case(seed):
    """Seeded random multi-block cases with Spalart-Allmaras, SST or
    Baldwin-Lomax (tests/helpers.py random_turb_multiblock_case; the same
    seeds run oracle against the reference build in
    tests/test_oracle_reference.py::test_random_turb_multiblock_cases): U,
    Res, dt and mu_t (its first ghost layer included) of every block within
    1e-11 relative per pass, the norms within 1e-10."""
    from helpers import random_turb_multiblock_case

    d, n = random_turb_multiblock_case(seed, "g")
    _compare_steps(d, n, ("U", "Res", "dt", "Amu_t"), exact=False, rtol=1e-11, norm_rtol=1e-10)


@pytest.mark.parametrize("seed", range(24))
def test_random_naca_case(seed):
    """Seeded random C-grid cases (tests/helpers.py random_naca_case),
    LU-SGS multigrid included.  Bit-exact without a turbulence model; with
    Baldwin-Lomax, SA or SST U, Res and mu_t (Amu_t, its first ghost layer
    included) within 1e-11 relative per pass and the norms within 1e-10 (the
    device evaluates exp, pow and tanh with its own math library)."""
    from helpers import random_naca_case
    from ec2d import formats as F
    import os

    d, n, exact, _ = random_naca_case(seed, "g")
    c = F.read_control(os.path.join(d, "control.in"))
    arrays = ("U", "Res") if c.Num_Mesh > 1 else ("U", "Res", "dt", "Lci", "Lcj")
    if c.Iflag_turbulence_model != F.TURB_NONE:
        arrays += ("Amu_t",)
    if exact:
        _compare_steps(d, n, arrays)
    else:
        _compare_steps(d, n, arrays, exact=False, rtol=1e-11, norm_rtol=1e-10)


@pytest.mark.parametrize("seed", range(12))
def test_random_dual_time_case(seed):
    """Seeded random dual-time cases (tests/helpers.py random_dual_case),
    early inner exits included: U, Res, dt, Un, Un1, the inner-loop index
    and the norms after every physical step, bit-exact without SA, 1e-11
    with it (as tests/test_gpu_dual.py)."""
    import oracle as O
    from ec2d import solver
    from helpers import random_dual_case
    from test_gpu_dual import _check

    d, exact = random_dual_case(seed, "g")
    # The device decides each inner-loop exit from tree-summed norms, the
    # oracle from serial sums (equal to about 1e-15 relative, 1e-11 with SA):
    # a seed whose exit test lands within 1e-8 of the limit could exit at
    # another inner iteration, so it is left out, with its values.
    probe = O.OracleCase(d)
    probe.init_flow()
    for _ in range(3):
        probe.main_step()
    limit = probe.ctl.Res_Inner_Limit
    close = [v for v in probe.inner_max_res if abs(v - limit) <= 1e-8 * limit]
    if close:
        pytest.skip(f"an inner-loop exit test lies within 1e-8 of the limit {limit:g}: {close}")
    ref = O.OracleCase(d)
    ref.init_flow()
    with solver.Batch([solver.Case.from_dir(d)]) as b:
        b.init_flow()
        for k in range(3):
            ref.main_step()
            b.main_step()
            _check(b, ref, k, exact)


@pytest.mark.parametrize("seed", range(6))
def test_random_batch(seed):
    """A batch of 2-4 seeded random C-grid cases that share the level options
    (turbulence model, scheme, flux, reconstruction, mesh levels, time method)
    and differ in grid size, Mach number, angle of attack, Reynolds number
    and wall distance: one launch per stage for all of them must equal each
    case run alone by the oracle (bit for bit without a turbulence model,
    1e-11 with one)."""
    import numpy as np
    import oracle as O
    from ec2d import cases, solver
    from ec2d import formats as F
    from helpers import first_diff, naca_case

    rng = np.random.default_rng(3000 + seed)
    turb = int(rng.choice([F.TURB_NONE, F.TURB_NONE, F.TURB_SA, F.TURB_SST, F.TURB_BL]))
    scheme, flux, recon = int(rng.integers(0, 7)), int(rng.integers(1, 7)), int(rng.integers(0, 3))
    nm = int(rng.integers(1, 4))
    tm = int(rng.choice([F.TIME_LU_SGS, F.TIME_RK3, F.TIME_EULER1]))
    ds = []
    for k in range(int(rng.integers(2, 5))):
        ma, aoa = float(rng.uniform(0.3, 0.7)), float(rng.uniform(-2.0, 4.0))
        if turb == F.TURB_NONE:
            c = cases.control_euler_naca(iters=10, Ma=ma, aoa=aoa)
        else:
            c = cases.control_sa_naca(iters=10, Re=float(10.0 ** rng.uniform(4.5, 5.5)), Ma=ma, aoa=aoa)
            c.Iflag_turbulence_model = turb
            if turb == F.TURB_SST:
                c.Twall = 300.0
        c.Iflag_Scheme, c.Iflag_Flux, c.IFlag_Reconstruction = scheme, flux, recon
        c.Time_Method, c.Num_Mesh = tm, nm
        c.CFL = {F.TIME_LU_SGS: 2.0, F.TIME_RK3: 0.3, F.TIME_EULER1: 0.2}[tm]
        nx, ny = int(rng.choice([49, 57, 65, 73])), int(rng.choice([13, 17, 21]))
        ds.append(naca_case(f"g_rbatch_{seed}_{k}", c, nx=nx, ny=ny, ds1=1e-4 if turb else 1e-3))
    cs = [solver.Case.from_dir(d) for d in ds]
    with solver.Batch(cs) as b:
        b.init_flow()
        for _ in range(3):
            b.main_step()
        Us = [b.get("U", k) for k in range(len(ds))]
    for k, d in enumerate(ds):
        ref = O.OracleCase(d)
        ref.init_flow()
        for _ in range(3):
            ref.main_step()
        if turb == F.TURB_NONE:
            assert first_diff(Us[k], ref.block().U) is None, f"case {k}"
        else:
            np.testing.assert_allclose(Us[k], ref.block().U, rtol=1e-11, atol=1e-300, err_msg=f"case {k}")


@pytest.mark.parametrize("seed", range(6))
def test_random_tuned_case(seed):
    """Seeded random C-grid cases tall enough for split sweeps (65..129
    points across), run with the launch shapes the autotuner picks on the live
    state (residual tile rows, lanes per cell, workgroups per block): the
    tuning leaves the state untouched and the tuned run equals the oracle
    (bit for bit without a turbulence model, 1e-11 with one)."""
    import numpy as np
    import oracle as O
    from ec2d import cases, solver
    from ec2d import formats as F
    from helpers import first_diff, naca_case

    rng = np.random.default_rng(4000 + seed)
    turb = int(rng.choice([F.TURB_NONE, F.TURB_NONE, F.TURB_SA, F.TURB_BL]))
    nm = int(rng.integers(1, 4))
    if turb == F.TURB_NONE:
        c = cases.control_euler_naca(iters=10)
    else:
        c = cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5)
        c.Iflag_turbulence_model = turb
    c.Iflag_Scheme, c.Iflag_Flux, c.IFlag_Reconstruction = int(rng.integers(0, 7)), int(rng.integers(1, 7)), 0
    c.Time_Method, c.Num_Mesh, c.CFL = F.TIME_LU_SGS, nm, 2.0
    nx, ny = int(rng.choice([97, 129, 161])), int(rng.choice([65, 97, 129]))
    d = naca_case(f"g_rtuned_{seed}", c, nx=nx, ny=ny, ds1=1e-4 if turb else 1e-3, fdm=(1,) if rng.random() < 0.3 else ())
    ref = O.OracleCase(d)
    ref.init_flow()
    with solver.Batch([solver.Case.from_dir(d)]) as b:
        b.init_flow()
        report = b.autotune("", reps=2)
        assert first_diff(b.get("U"), ref.block().U) is None, "state changed by tuning"
        for k in range(3):
            ref.main_step()
            b.main_step()
            if turb == F.TURB_NONE:
                assert first_diff(b.get("U"), ref.block().U) is None, f"step {k} ({report})"
            else:
                np.testing.assert_allclose(b.get("U"), ref.block().U, rtol=1e-11, atol=1e-300,
                                           err_msg=f"step {k} ({report})")
        b.res_norms()  # reports a timed-out cross-workgroup wait, if any
