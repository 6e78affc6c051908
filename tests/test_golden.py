"""The oracle against golden vectors of the reference (CPU).

tests/golden/*.npz hold small generated cases (their input files) and the
state of the unmodified reference (flang build of its sources) after a few
main-loop passes, in full FP64 precision; tests/golden/make_golden.py wrote
them.  The oracle must reproduce every stored array and residual norm bit for
bit.
"""
import glob
import hashlib
import os

import numpy as np
import pytest

import oracle as O
from helpers import first_diff

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURES = sorted(f for f in glob.glob(os.path.join(HERE, "*.npz")) if not f.endswith("flang_list_directed.npz"))


def load_case(path, d):
    """Writes the fixture's input files into d (text, or bytes for the large
    fixtures) and returns the fixture."""
    z = np.load(path, allow_pickle=False)
    for k in z.files:
        if k.startswith("file:"):
            v = z[k]
            with open(os.path.join(d, k[5:]), "wb") as f:
                f.write(v.tobytes() if v.dtype == np.uint8 else str(v).encode())
    return z


def sha256(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p)[:-4] for p in FIXTURES])
def test_oracle_reproduces_reference(path, tmp_path):
    z = load_case(path, str(tmp_path))
    oc = O.OracleCase(str(tmp_path))
    oc.init_flow()
    steps = int(z["steps"])
    for k in range(steps):
        oc.main_step()
        rmax, rrms = oc.norms(1)
        assert np.array_equal(rrms, z["res_rms"][k]), f"step {k} rms {rrms} vs {z['res_rms'][k]}"
        assert np.array_equal(rmax, z["res_max"][k]), f"step {k} max"
    if "inner" in z.files:  # dual-time stepping: the reference's printed inner-loop indices
        assert oc.inner_iterations == [int(v) for v in z["inner"]]
    for key in z.files:
        if ":" not in key or key.startswith(("file:", "shape:")):
            continue
        if key.startswith("sha256:"):  # large fixtures: digest of the FP64 bytes
            _, name, mb = key.split(":")
            got = getattr(oc.block(1, int(mb)), name)
            assert list(got.shape) == list(z[f"shape:{name}:{mb}"]), key
            assert sha256(got) == str(z[key]), f"{key}: the oracle's array differs from the reference's"
            continue
        name, mb = key.split(":")
        got = getattr(oc.block(1, int(mb)), name)
        want = z[key]
        if name in ("Lvi", "Lvj"):
            got, want = got[:-1, :-1], want[:-1, :-1]
        diff = first_diff(got, want)
        assert diff is None, f"{key}: {diff}"


def test_fixtures_present():
    assert len(FIXTURES) >= 6
