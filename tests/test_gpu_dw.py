"""Wall distance on the device (comput_dw, sub_turbulence_SST.f90:286-337,
computed by ec2d_commit when the caller gives no dw) against the host
restatement (ec2d/geometry.py), which tests/test_oracle_reference.py pins bit
for bit to the reference's geometry.  The dw planes must be identical, and so
must U after two steps.
"""
import pytest

from ec2d import cases
from ec2d import formats as F
from ec2d import solver
from helpers import first_diff, naca_case, vortex_case

pytestmark = pytest.mark.gpu


def _walls(bcs):
    for subs in bcs:
        for s in subs:
            if s[6] == F.BC_FARFIELD and s[1] in (1, 2):
                s[6] = F.BC_WALL


def _case(kind):
    if kind == "sa_naca":
        return naca_case("g_dw_sa", cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5), ds1=1e-4)
    c = cases.control_vortex(iters=10)
    c.If_viscous, c.Re, c.Iflag_turbulence_model, c.Twall = 1, 1e4, F.TURB_SST, 300.0
    c.Iflag_init = 0  # freestream start: the reference defines no restart of turbulence fields
    return vortex_case("g_dw_sst_mb", c, bc_edit=_walls)


def _run(d, device_dw):
    case = solver.Case.from_dir(d, device_dw=device_dw)
    assert (case.levels[0][0].dw is None) == device_dw
    with solver.Batch([case]) as b:
        nb = len(case.levels[0])
        dw = [b.get("dw", 0, 1, mb) for mb in range(1, nb + 1)]
        b.init_flow()
        b.main_step()
        b.main_step()
        U = [b.get("U", 0, 1, mb) for mb in range(1, nb + 1)]
    return dw, U


@pytest.mark.parametrize("kind", ["sa_naca", "sst_multiblock"])
def test_device_wall_distance_equals_host(kind):
    d = _case(kind)
    dw_h, U_h = _run(d, False)
    dw_d, U_d = _run(d, True)
    for mb, (a, b) in enumerate(zip(dw_d, dw_h), start=1):
        assert first_diff(a, b) is None, f"dw block {mb}: {first_diff(a, b)}"
        assert a[1:-1, 1:-1].max() < 1e9  # every cell found a wall
    for mb, (a, b) in enumerate(zip(U_d, U_h), start=1):
        assert first_diff(a, b) is None, f"U block {mb}: {first_diff(a, b)}"


GEOM = ("vol", "si", "sj", "ni1", "ni2", "nj1", "nj2", "x1", "y1")


@pytest.mark.parametrize("kind", ["cgrid_sa_mg2", "multiblock", "rae2822"])
def test_device_geometry_equals_host(kind, tmp_path):
    """Ghost nodes, cell centres, face metrics (and, for SA / SST, the wall
    distance) made on the device at ec2d_commit from the interior nodes
    equal the host restatement (pinned to the reference's own geometry) bit
    for bit on every level and block, and so do two passes run on them."""
    import os

    from ec2d import cases as C_

    if kind == "cgrid_sa_mg2":
        c = C_.control_sa_naca(iters=10, Re=1e5, Ma=0.5, num_mesh=2, time_method=F.TIME_RK3)
        c.CFL = 0.3
        d = naca_case("g_devgeo_c", c, nx=129, ny=33, ds1=1e-4)
    elif kind == "multiblock":
        from helpers import vortex_case
        d = vortex_case("g_devgeo_mb", C_.control_vortex(iters=10))
    else:
        from test_golden import load_case
        load_case(os.path.join(os.path.dirname(__file__), "golden", "rae2822_sample_sst_isothermal.npz"),
                  str(tmp_path))
        d = str(tmp_path)
    host = solver.Case.from_dir(d)
    dev = solver.Case.from_dir(d, device_geometry=True)
    U = {}
    for label, case in (("host", host), ("device", dev)):
        with solver.Batch([case]) as b:
            geo = {(lev, mb, n): b.get(n, 0, lev, mb) for lev in range(1, host.num_mesh + 1)
                   for mb in range(1, len(host.levels[lev - 1]) + 1) for n in GEOM}
            if host.ctl.Iflag_turbulence_model in (F.TURB_SA, F.TURB_SST):
                geo.update({(1, mb, "dw"): b.get("dw", 0, 1, mb) for mb in range(1, len(host.levels[0]) + 1)})
            b.init_flow()
            b.run(2)
            U[label] = [b.get("U", 0, 1, mb) for mb in range(1, len(host.levels[0]) + 1)]
            if label == "host":
                ref_geo = geo
    for key, v in geo.items():
        diff = first_diff(v, ref_geo[key])
        assert diff is None, f"{key}: {diff}"
    for mb, (a_, b_) in enumerate(zip(U["device"], U["host"]), start=1):
        assert first_diff(a_, b_) is None, f"U block {mb}"
