"""The reference's result files (output / output_Res,
opencfd_ec2d_v1.5.4.f90:266-371) written from the device state.

tests/test_oracle_reference.py shows that formats.write_output and
append_residual reproduce the reference's files byte for byte from the
oracle's state.  Here Batch.output / output_res write them from the GPU's
state: for Euler cases, whose state matches the oracle bit for bit, the flow
files must be byte-identical to the oracle's; Residual.dat holds tree-summed
norms (1e-12 relative), compared as numbers to 1e-9.  Cp.dat, the wall-
pressure tool's file (read_flow-1.2.f90), follows from flow2d.dat and must be
byte-identical too.
"""
import os

import numpy as np
import pytest

import oracle as O
from ec2d import cases
from ec2d import formats as F
from ec2d import solver
from helpers import naca_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mg", [False, True])
def test_output_files_from_device(mg, tmp_path):
    c = cases.control_euler_naca(iters=10)
    if mg:
        c.Time_Method, c.Num_Mesh, c.CFL = F.TIME_RK3, 3, 0.3
    d = naca_case(f"g_out_{int(mg)}", c)
    ref = O.OracleCase(d)
    ref.init_flow()
    gdir, odir = tmp_path / "gpu", tmp_path / "oracle"
    gdir.mkdir()
    odir.mkdir()
    meshes = (1, 2) if mg else (1,)
    with solver.Batch([solver.Case.from_dir(d)]) as b:
        b.init_flow()
        for _ in range(2):
            ref.main_step()
            b.main_step()
        names = []
        for nm in meshes:
            names += [os.path.basename(p) for p in b.output(0, nm, str(gdir))]
            blocks = [{"U": x.U, "Amu_t": x.Amu_t, "x1": x.geo["x1"], "y1": x.geo["y1"]}
                      for x in ref.meshes[nm - 1]["blocks"]]
            F.write_output(str(odir), nm, ref.kstep[0], blocks, ref.meshes[nm - 1]["M"].nvar, c.Cv, c.gamma, c.Ma,
                           c.Re)
        b.output_res(0, 1, str(gdir))
        mx, rms = ref.norms(1)
        F.append_residual(str(odir), ref.kstep[0], mx, rms, c.nvar)
        names.append(os.path.basename(b.wall_cp(0, str(gdir))))
        F.write_wall_cp(str(odir), d)
    assert "flow2d.dat" in names and (not mg or "flow2d-2.dat" in names) and "Cp.dat" in names
    for n in names:
        assert (gdir / n).read_text() == (odir / n).read_text(), n
    g = (gdir / "Residual.dat").read_text().split()
    o = (odir / "Residual.dat").read_text().split()
    assert g[0] == o[0]  # Kstep
    np.testing.assert_allclose([float(v) for v in g[1:]], [float(v) for v in o[1:]], rtol=1e-9, atol=0)


@pytest.mark.parametrize("turb", [F.TURB_SA, F.TURB_SST, F.TURB_BL])
def test_output_files_from_device_sa(turb, tmp_path):
    """SA, SST (isothermal wall) and Baldwin-Lomax: flow2d.dat (with mu_t on
    the first ghost layer, as each model leaves it) and SA2d.dat / SST2d.dat
    from the device against the oracle's.  The device's turbulence paths agree
    to 1e-11 relative (exp, pow and tanh of its own math library), so the
    F20.10 columns are compared as numbers: equal within one unit of the last
    printed digit."""
    c = cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5)
    c.Iflag_turbulence_model = turb
    if turb == F.TURB_SST:
        c.Twall = 300.0
    d = naca_case(f"g_out_t{turb}", c, ds1=1e-4)
    ref = O.OracleCase(d)
    ref.init_flow()
    gdir, odir = tmp_path / "gpu", tmp_path / "oracle"
    gdir.mkdir()
    odir.mkdir()
    with solver.Batch([solver.Case.from_dir(d)]) as b:
        b.init_flow()
        for _ in range(2):
            ref.main_step()
            b.main_step()
        names = [os.path.basename(p) for p in b.output(0, 1, str(gdir))]
        blocks = [{"U": x.U, "Amu_t": x.Amu_t, "x1": x.geo["x1"], "y1": x.geo["y1"]} for x in ref.meshes[0]["blocks"]]
        F.write_output(str(odir), 1, ref.kstep[0], blocks, c.nvar, c.Cv, c.gamma, c.Ma, c.Re)
    assert "flow2d.dat" in names
    assert ("SA2d.dat" in names) == (turb == F.TURB_SA) and ("SST2d.dat" in names) == (turb == F.TURB_SST)
    for n in names:
        g = (gdir / n).read_text().splitlines()
        o = (odir / n).read_text().splitlines()
        assert len(g) == len(o), n
        for lg, lo in zip(g, o):
            if lg.lstrip()[:1].isalpha():
                assert lg == lo, n
                continue
            np.testing.assert_allclose([float(v) for v in lg.split()], [float(v) for v in lo.split()], rtol=1e-9,
                                       atol=1.5e-10, err_msg=n)
    if turb == F.TURB_SA:
        # the Amut column of the far-field ghost row j = ny (the last nx+1 lines) holds the free-stream
        # mu_t of SA (SST and Baldwin-Lomax leave zero there, as the oracle does)
        B = solver.Case.from_dir(d).levels[0][0]
        amut = np.array([float(v.split()[7]) for v in (gdir / "flow2d.dat").read_text().splitlines()[-(B.nx + 1):]])
        assert np.all(amut > 0)
