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
    ref = O.OracleCase(d)
    ref.init_flow()
    with solver.Batch([solver.Case.from_dir(d)]) as b:
        b.init_flow()
        for k in range(3):
            ref.main_step()
            b.main_step()
            _check(b, ref, k, exact)
