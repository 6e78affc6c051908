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
