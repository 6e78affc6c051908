"""ec2d_du_lu_sgs_2d_noalloc against ec2d_du_lu_sgs_2d (GPU).

The no-allocation variant keeps the LU-SGS records (and the rings of split
sweeps) in the library's static device array.  Its dU must equal the
allocating entry point's bit for bit, on two calls in a row of which the
second has a smaller input (so the array holds the first call's records when
the second starts), with split sweeps (rings in the array) and one selected
block of a multi-block case; a level whose records exceed the array is
refused with a message, and the state is never touched (U unchanged).
"""
import numpy as np
import pytest

from ec2d import capi
from ec2d import cases
from ec2d import solver
from helpers import first_diff, naca_case, vortex_case

pytestmark = pytest.mark.gpu


def _both(b, nmesh=1, mblock=0, alfa1=0.0, blocks=(1,)):
    U0 = [b.get("U", 0, nmesh, mb) for mb in blocks]
    b.du_lu_sgs_2d(nmesh, mblock, alfa1)
    ref = [b.get("dU", 0, nmesh, mb) for mb in blocks]
    # wipe the cells' dU (1:nx-1, 1:ny-1) so that the variant must write every
    # value itself; row ny and column nx of dU belong to no cell and stay
    for mb, r in zip(blocks, ref):
        w = r.copy()
        w[:-1, :-1] = np.nan
        b.put("dU", w, 0, nmesh, mb)
    b.du_lu_sgs_2d_noalloc(nmesh, mblock, alfa1)
    got = [b.get("dU", 0, nmesh, mb) for mb in blocks]
    for mb, g, r, u0 in zip(blocks, got, ref, U0):
        if mblock in (0, mb):
            assert np.isfinite(r[:-1, :-1]).all()
            diff = first_diff(g, r)
            assert diff is None, f"block {mb}: {diff}"
        assert first_diff(b.get("U", 0, nmesh, mb), u0) is None, "the LU-SGS call must not change U"


def test_noalloc_lusgs_equals_allocating_entry_point():
    # call 1: SA C-grid 257 x 65 (16,384 cells, records about 10 MiB), rows split over two workgroups
    c = cases.control_sa_naca(iters=10, Re=1e6, Ma=0.6)
    d1 = naca_case("na_big", c, nx=257, ny=65, ds1=1e-5)
    with solver.Batch([solver.Case.from_dir(d1)]) as b:
        b.init_flow()
        b.run(3)
        b.comput_residual_one_mesh(1)
        b.lusgs_shape(1, 2, 2)
        _both(b, alfa1=0.0)
        _both(b, alfa1=2.5)
    # call 2: a smaller input, 6 Euler blocks, one block selected and all of them
    d2 = vortex_case("na_small", cases.control_vortex(iters=10))
    with solver.Batch([solver.Case.from_dir(d2)]) as b:
        b.init_flow()
        b.run(2)
        b.comput_residual_one_mesh(1)
        _both(b, mblock=0, blocks=range(1, 7))
        _both(b, mblock=3, blocks=(3,))
        # the fused launch (its skewed input copy made in the scratch on every call)
        b.lusgs_shape(1, 1, 0)
        _both(b, mblock=0, blocks=range(1, 7))
        _both(b, mblock=2, blocks=(2,))


def test_noalloc_refuses_a_level_beyond_the_scratch():
    # 1025 x 129 SA C-grid: 131,072 cells, about 80 MiB of records
    c = cases.control_sa_naca(iters=10, Re=1e6, Ma=0.6)
    d = naca_case("na_huge", c, nx=1025, ny=129, ds1=1e-5)
    with solver.Batch([solver.Case.from_dir(d)]) as b:
        b.init_flow()
        b.comput_residual_one_mesh(1)
        with pytest.raises(capi.Ec2dError, match="more than the 64 MiB scratch"):
            b.du_lu_sgs_2d_noalloc(1, 0, 0.0)
        b.du_lu_sgs_2d(1, 0, 0.0)  # the allocating entry point still serves it
