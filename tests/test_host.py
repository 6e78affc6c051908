"""Host-side logic on the CPU: case files, case generators, batch set-up
checks that need no device."""
import dataclasses
import os

import numpy as np
import pytest

from ec2d import cases
from ec2d import formats as F
from ec2d import geometry as G


def test_control_round_trip(tmp_path):
    c = cases.control_sa_naca(iters=37, Ma=0.73, aoa=2.0, Re=6.5e6, cfl=5.0, num_mesh=3,
                              time_method=F.TIME_LU_SGS, scheme=F.SCHEME_MUSCL3, flux=F.FLUX_ROE)
    p = str(tmp_path / "control.in")
    F.write_control(p, c)
    r = F.read_control(p)
    for f in dataclasses.fields(F.Control):
        a, b = getattr(c, f.name), getattr(r, f.name)
        assert (list(a) == list(b)) if isinstance(a, (list, tuple)) else a == b, f.name
    assert r.iterations() == c.iterations() == 37


def test_mesh_and_bc_round_trip(tmp_path):
    x, y, bc, _ = cases.naca_cgrid(65, 17, ds1=1e-3)
    F.write_mesh2d(str(tmp_path / "Mesh2d.dat"), [x], [y])
    F.write_bcin(str(tmp_path / "bc2d_1.in"), [bc])
    xs, ys = F.read_mesh2d(str(tmp_path / "Mesh2d.dat"))
    assert np.array_equal(xs[0], x) and np.array_equal(ys[0], y)
    assert F.read_bcin(str(tmp_path / "bc2d_1.in")) == [bc]


@pytest.mark.parametrize("nx,ny,ds1", [(65, 17, 1e-3), (257, 65, 1e-3), (513, 129, 1e-6)])
def test_cgrid_is_valid_for_three_levels(nx, ny, ds1):
    x, y, bc, info = cases.naca_cgrid(nx, ny, ds1=ds1)
    assert x.shape == (ny, nx)
    # every subface limit admits two coarsenings
    for s in bc:
        for idx in s[2:6]:
            assert (idx - 1) % 4 == 0, s
    # wake cut: the two wake subfaces match point for point, reversed
    lw = [s for s in bc if s[6] == 1]
    assert len(lw) == 2 and lw[0][3] - lw[0][2] == lw[1][3] - lw[1][2]
    iw1, iw2 = lw[0][3], lw[1][2]
    assert np.array_equal(x[0, :iw1], x[0, iw2 - 1:][::-1]) and np.array_equal(y[0, :iw1], y[0, iw2 - 1:][::-1])
    lev = G.build_levels([x], [y], [bc], 3)
    for blocks in lev:
        B = blocks[0]
        assert np.all(B.vol[:-1, :-1] > 0)
    assert info["first_spacing"] == pytest.approx(ds1, rel=1e-12)


def test_vortex_multiblock_interfaces_match():
    xs, ys, bcs, field, _ = cases.vortex_multiblock(seed=3, nbx=3, nby=2, wmin=9, wmax=21, hmin=9, hmax=17, h=0.25)
    assert len(xs) == 6
    for b, subs in enumerate(bcs):
        for s in subs:
            if s[6] <= 0:
                continue
            nb = bcs[s[6] - 1][s[7] - 1]
            assert nb[6] == b + 1  # the neighbour points back
            n1 = s[3] - s[2] + s[5] - s[4]
            assert n1 == nb[3] - nb[2] + nb[5] - nb[4]

            def pts(k, sub):
                f, i0, i1, j0, j1 = sub[1:6]
                return np.stack([xs[k][j0 - 1:j1, i0 - 1:i1].ravel(), ys[k][j0 - 1:j1, i0 - 1:i1].ravel()])

            assert np.array_equal(pts(b, s), pts(s[6] - 1, nb))


def test_case_generators_are_deterministic(tmp_path):
    a = cases.ensemble_params(seed=5, n=8)
    b = cases.ensemble_params(seed=5, n=8)
    assert a == b
    x1, y1, _, _ = cases.naca_cgrid(65, 17)
    x2, y2, _, _ = cases.naca_cgrid(65, 17)
    assert np.array_equal(x1, x2) and np.array_equal(y1, y2)


def test_write_case_files(tmp_path):
    x, y, bc, _ = cases.naca_cgrid(65, 17)
    d = cases.write_case(str(tmp_path), cases.control_euler_naca(iters=3), [x], [y], [bc], fdm_blocks=(1,))
    assert set(os.listdir(d)) >= {"control.in", "Mesh2d.dat", "bc2d_1.in", "FDM.in"}
    assert F.read_fdm_in(os.path.join(d, "FDM.in")) == [1]


def test_fortran_pow_is_left_to_right_chain():
    x = 1.0000000000000002
    assert F.fortran_pow_int(x, 6) == ((((x * x) * x) * x) * x) * x


def test_fortran_edit_descriptors():
    """Fw.d and Ew.d as the reference's result files use them (the reference's
    own files are matched byte for byte in tests/test_oracle_reference.py)."""
    assert F.fortran_f(25.87895096612) == "       25.8789509661"
    assert F.fortran_f(-1e-12) == "       -0.0000000000"
    assert F.fortran_f(1e12) == "*" * 20
    assert F.fortran_e(0.006318817554, 20) == "    0.6318817554E-02"
    assert F.fortran_e(0.0, 24) == " " * 8 + "0.0000000000E+00"
    assert F.fortran_e(-123.456, 20) == "   -0.1234560000E+03"
    assert F.fortran_e(9.99999999996, 20) == "    0.1000000000E+02"  # rounding carries into the exponent
    assert F.fortran_e(1e-120, 20) == "    0.1000000000-119"


def test_list_directed_reals_match_flang():
    """formats.fortran_list_real against what flang's runtime printed for 3326
    doubles (tests/golden/flang_list_directed.npz, tests/golden/make_list_directed.py):
    random bit patterns over the whole range, subnormals, powers of two and
    ten with their neighbours, the edges of the F/E switch.  About one value
    in six of the random ones has digits other than the nearest shortest
    decimal (Python's repr), so a repr-based formatter fails here."""
    import struct

    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "flang_list_directed.npz"))
    bad = []
    n_repr = 0
    for b, t in zip(z["bits"], z["text"]):
        x = struct.unpack("<d", struct.pack("<q", int(b)))[0]
        want = t.decode()
        if F.fortran_list_real(x) != want:
            bad.append((repr(x), F.fortran_list_real(x), want))
        n_repr += float(want) == x and repr(abs(x)).lstrip("0").replace("e", "E") not in want
    assert not bad, bad[:5]
    assert n_repr > 100  # the fixture holds values where nearest-digit output would differ
    assert F.fortran_list_line("zone i=", 41) == " zone i= 41"
    assert F.fortran_list_line(0.5, -0.0, 12.0) == " .5 -0. 12."


def test_turbulence_restart_is_refused(tmp_path):
    """The reference defines no restart of turbulence fields
    (sub_init.f90:527-553): the host side refuses it instead of inventing one."""
    from ec2d import solver

    x, y, bc, _ = cases.naca_cgrid(33, 9, ds1=1e-3)
    c = cases.control_sa_naca(iters=1, Re=1e5, Ma=0.5)
    c.Iflag_init = 1
    d = cases.write_case(str(tmp_path / "sa"), c, [x], [y], [bc])
    with pytest.raises(ValueError, match="SST2d.dat"):
        solver.Case.from_dir(d).read_restart()
