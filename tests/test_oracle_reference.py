"""The oracle against the reference itself (CPU, where oracle/_ref is built).

oracle/_ref/libec2d_ref.so is the flang build of the unmodified reference
sources plus a bind(C) harness (oracle/ref_harness.f90); __graft_entry__.build()
makes it wherever the reference sources are present.  Every test runs the same
generated case through both and requires bit-identical FP64 state.  Without
the reference build the tests are skipped; tests/test_golden.py then pins the
oracle against stored reference outputs.
"""
import itertools
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle as O
import refpy
from ec2d import cases
from ec2d import formats as F
from helpers import dual_case, first_diff, naca_case, vortex_case

pytestmark = pytest.mark.skipif(not refpy.available(), reason="reference build (oracle/_ref) not present")


def run_both(d, nsteps, arrays=("U", "Res", "dt")):
    R = refpy.RefCase(d)
    oc = O.OracleCase(d)
    oc.init_flow()
    nblock = R.mesh_info(1)["nblock"]
    for mb in range(1, nblock + 1):
        assert first_diff(oc.block(1, mb).U, R.get("U", 1, mb)) is None, f"initial U block {mb}"
    for k in range(nsteps):
        R.op("main_step")
        oc.main_step()
        for mb in range(1, nblock + 1):
            for a in arrays:
                got, want = getattr(oc.block(1, mb), a), R.get(a, 1, mb)
                if a in ("Lvi", "Lvj"):
                    got, want = got[:-1, :-1], want[:-1, :-1]
                diff = first_diff(got, want)
                assert diff is None, f"step {k} block {mb} {a}: {diff}"
        assert np.array_equal(oc.norms(1)[1], R.norms(1)[1]), f"step {k} rms"
    return R, oc


@pytest.mark.parametrize("scheme,flux,recon", [(s, 4, 0) for s in range(7)] + [(5, f, 0) for f in range(1, 7)] +
                         [(s, f, r) for s, f, r in itertools.product((1, 5), (2, 6), (1, 2))])
def test_schemes_fluxes_reconstructions(scheme, flux, recon):
    c = cases.control_euler_naca(iters=10)
    c.Iflag_Scheme, c.Iflag_Flux, c.IFlag_Reconstruction = scheme, flux, recon
    run_both(naca_case(f"r_sfr_{scheme}_{flux}_{recon}", c, nx=33, ny=9), 2)


def test_laminar_isothermal_wall():
    c = cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5)
    c.Iflag_turbulence_model, c.Twall = F.TURB_NONE, 350.0
    run_both(naca_case("r_lam", c, ds1=1e-4), 3, ("U", "Res", "dt", "Lvi", "Lvj", "Amu"))


def test_sa_adiabatic():
    c = cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5)
    run_both(naca_case("r_sa", c, ds1=1e-4), 3, ("U", "Res", "dt", "Amu_t"))


def test_sst_isothermal():
    c = cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5)
    c.Iflag_turbulence_model, c.Twall = F.TURB_SST, 300.0
    run_both(naca_case("r_sst", c, ds1=1e-4), 3, ("U", "Res", "dt", "Amu_t"))


def test_baldwin_lomax():
    """turbulence_model_BL (sub_turbulence_BL.f90): the C-grid's wall ends
    inside face 2, so its ghost copies also reach the first wake columns."""
    c = cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5)
    c.Iflag_turbulence_model = F.TURB_BL
    run_both(naca_case("r_bl", c, ds1=1e-4), 3, ("U", "Res", "dt", "Amu_t"))


def _walls(bcs):
    for subs in bcs:
        for s in subs:
            if s[6] == F.BC_FARFIELD and s[1] in (1, 2):
                s[6] = F.BC_WALL


def test_baldwin_lomax_multigrid():
    """Baldwin-Lomax on the fine level of a 2-level RK3 V-cycle (the coarse
    levels carry no turbulence model, opencfd_ec2d_v1.5.4.f90:252-259)."""
    c = cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5)
    c.Iflag_turbulence_model, c.Time_Method, c.Num_Mesh, c.CFL = F.TURB_BL, F.TIME_RK3, 2, 0.3
    run_both(naca_case("r_bl_mg", c, ds1=1e-4), 2, ("U", "Res", "dt", "Amu_t"))


def test_baldwin_lomax_two_walls():
    """A block with walls on two faces: each cell keeps the smaller mu_t of
    the two wall lines through it, and the second wall's ghost copies follow
    the first's."""
    c = cases.control_vortex(iters=10)
    c.If_viscous, c.Re, c.Iflag_turbulence_model = 1, 1e4, F.TURB_BL
    run_both(vortex_case("r_bl_mb", c, bc_edit=_walls), 3, ("U", "Res", "dt", "Amu_t"))


def _oracle_blocks(oc, nmesh):
    return [{"U": b.U, "Amu_t": b.Amu_t, "x1": b.geo["x1"], "y1": b.geo["y1"]} for b in oc.meshes[nmesh - 1]["blocks"]]


@pytest.mark.parametrize("kind", ["sa", "sst", "euler_mg3"])
def test_output_files_match_reference(kind, tmp_path):
    """output / output_Res (opencfd_ec2d_v1.5.4.f90:266-371): the files the
    reference writes from its state after two passes equal, byte for byte,
    the ones formats.write_output / append_residual write from the oracle's."""
    c = cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5) if kind != "euler_mg3" else cases.control_euler_naca(iters=10)
    if kind == "sst":
        c.Iflag_turbulence_model, c.Twall = F.TURB_SST, 300.0
    if kind == "euler_mg3":
        c.Time_Method, c.Num_Mesh, c.CFL = F.TIME_RK3, 3, 0.3
    d = naca_case(f"r_out_{kind}", c, ds1=1e-4 if kind != "euler_mg3" else 1e-3)
    if os.path.exists(os.path.join(d, "Residual.dat")):
        os.remove(os.path.join(d, "Residual.dat"))  # output_Res appends; case directories persist between runs
    R = refpy.RefCase(d)
    oc = O.OracleCase(d)
    oc.init_flow()
    for _ in range(2):
        R.op("main_step")
        oc.main_step()
    meshes = (1, 2) if kind == "euler_mg3" else (1,)
    mine = tmp_path / "mine"
    mine.mkdir()
    written = []
    for nm in meshes:
        R.op("output", nm)
        nv = oc.meshes[nm - 1]["M"].nvar
        written += F.write_output(str(mine), nm, oc.kstep[0], _oracle_blocks(oc, nm), nv, c.Cv, c.gamma, c.Ma, c.Re)
    for _ in range(2):
        R.op("output_res", 1)
        mx, rms = oc.norms(1)
        written.append(F.append_residual(str(mine), oc.kstep[0], mx, rms, c.nvar))
    for path in sorted(set(written)):
        name = os.path.basename(path)
        want = open(os.path.join(d, name)).read()
        got = open(path).read()
        assert got == want, f"{name}: first difference at {next(i for i, (x, y) in enumerate(zip(got, want)) if x != y) if got[:len(want)] != want[:len(got)] else min(len(got), len(want))}"


READ_FLOW = os.path.join(os.path.dirname(refpy.LIB_PATH), "read_flow")


@pytest.mark.skipif(not os.path.exists(READ_FLOW), reason="reference wall-Cp tool (oracle/_ref/read_flow) not built")
@pytest.mark.parametrize("kind", ["sa", "walls"])
def test_wall_cp_matches_reference_tool(kind, tmp_path):
    """formats.write_wall_cp against the reference's own post-processor
    (read_flow-1.2.f90, built unmodified into oracle/_ref/read_flow) on the
    flow2d.dat the reference writes: Cp.dat byte for byte.  "walls": a
    multi-block layout with wall subfaces on two faces of several blocks
    (several zones; the tool reads cell values at node indices)."""
    import shutil

    if kind == "sa":
        d = naca_case("r_cp_sa", cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5), ds1=1e-4)
    else:
        d = vortex_case("r_cp_walls", cases.control_vortex(iters=10), bc_edit=_walls)
    R = refpy.RefCase(d)
    for _ in range(2):
        R.op("main_step")
    R.op("output", 1)
    ref, mine = tmp_path / "ref", tmp_path / "mine"
    ref.mkdir()
    mine.mkdir()
    for n in ("control.in", "Mesh2d.dat", "bc2d_1.in", "flow2d.dat"):
        shutil.copy(os.path.join(d, n), ref / n)
    shutil.copy(os.path.join(d, "flow2d.dat"), mine / "flow2d.dat")
    subprocess.run([READ_FLOW], cwd=ref, check=True, capture_output=True, timeout=120)
    F.write_wall_cp(str(mine), d)
    want = (ref / "Cp.dat").read_text()
    got = (mine / "Cp.dat").read_text()
    assert want.count("zone") >= (1 if kind == "sa" else 2)
    assert got == want


def _symm(bcs):
    for subs in bcs:
        for s in subs:
            if s[6] == F.BC_FARFIELD and s[1] in (2, 4):
                s[6] = F.BC_SYMMETRY


def _inout(bcs):
    for subs in bcs:
        for s in subs:
            if s[6] == F.BC_FARFIELD and s[1] == 1:
                s[6] = F.BC_INLET
            if s[6] == F.BC_FARFIELD and s[1] == 3:
                s[6] = F.BC_OUTLET


@pytest.mark.parametrize("variant", ["farfield", "symmetry", "inlet_outlet"])
def test_multiblock(variant):
    c = cases.control_vortex(iters=10)
    edit = {"farfield": None, "symmetry": _symm, "inlet_outlet": _inout}[variant]
    if variant == "inlet_outlet":
        c.P_outlet = 1.0 / (1.4 * 0.5 * 0.5) * 0.98
    run_both(vortex_case(f"r_mb_{variant}", c, bc_edit=edit), 3)


@pytest.mark.parametrize("tm,nm,sa", [(F.TIME_RK3, 2, False), (F.TIME_RK3, 3, False), (F.TIME_EULER1, 2, False),
                                      (F.TIME_EULER1, 3, False), (F.TIME_RK3, 3, True)])
def test_multigrid(tm, nm, sa):
    c = cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5) if sa else cases.control_euler_naca(iters=10)
    c.Time_Method, c.Num_Mesh, c.CFL = tm, nm, 0.3
    run_both(naca_case(f"r_mg_{tm}_{nm}_{int(sa)}", c, ds1=1e-4 if sa else 1e-3), 2, ("U", "Res"))


def test_multigrid_pre_steps():
    c = cases.control_euler_naca(iters=10)
    c.Time_Method, c.Num_Mesh, c.CFL, c.Pre_Step_Mesh = F.TIME_RK3, 3, 0.3, [0, 2, 3]
    run_both(naca_case("r_mg_pre", c), 2, ("U", "Res"))


@pytest.mark.parametrize("sa", [False, True])
def test_fdm_weno5(sa):
    c = cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5) if sa else cases.control_euler_naca(iters=10)
    run_both(naca_case(f"r_fdm_{int(sa)}", c, ds1=1e-4 if sa else 1e-3, fdm=(1,)), 2)


def test_lusgs_with_dual_time_diagonal():
    """du_LU_SGS_2D alone with alfa1 != 0 (the dual-time diagonal term)."""
    c = cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5)
    d = naca_case("r_lusgs_alfa", c, ds1=1e-4)
    R = refpy.RefCase(d)
    oc = O.OracleCase(d)
    oc.init_flow()
    R.op("main_step")
    oc.main_step()
    R.op("residual", 1)
    oc.residual(1)
    for c8 in (0, 3, 40):  # the harness passes alfa1 = c8 / 8
        R.op("lusgs", 1, 1, c8)
        oc.lusgs(1, 1, c8 / 8.0)
        assert first_diff(oc.block().dU, R.get("dU")) is None, c8


def ref_dual(d, nsteps, tmp_path):
    """The reference's dual-time passes in a fresh process (ref_dual_run.py):
    saved arrays per pass and the inner-iteration counts it prints."""
    out = tmp_path / "ref.npz"
    r = subprocess.run([sys.executable, os.path.join(os.path.dirname(__file__), "ref_dual_run.py"), d, str(nsteps),
                        str(out)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    inner = [int(line.split()[-1]) for line in r.stdout.splitlines() if "Inner step" in line]
    return np.load(out), inner


@pytest.mark.parametrize("sa,limit,fdm", [(False, 1e-30, ()), (False, 3.5e-3, ()), (True, 1e-30, ()),
                                         (True, 1e-30, (1,))])
def test_dual_time_lusgs(sa, limit, fdm, tmp_path):
    """Dual_time_LU_SGS (sub_NS_singlegrid.f90:231-310) and its source term
    (sub_Residual.f90:28-34, 55-64), with both first-call flags: Sfac =
    1/(3 dt) on the process's first residual only, alfa1 = 1/dt on the first
    physical step only.  The 3.5e-3 inner limit makes inner loops exit early
    (the recorded max_m Res_rms values lie at least 7% from it); Un and Un1
    are compared on the interior, the cells the reference sets."""
    n_inner = 8 if limit > 1e-20 else 6
    d = dual_case(f"r_dual_{int(sa)}_{int(limit > 1e-20)}_{len(fdm)}", sa, limit, n_inner, fdm)
    z, inner = ref_dual(d, 3, tmp_path)
    oc = O.OracleCase(d)
    oc.init_flow()
    for k in range(3):
        oc.main_step()
        b = oc.block()
        for a in ("U", "Res", "dt"):
            diff = first_diff(getattr(b, a), z[f"{a}_{k}"])
            assert diff is None, f"step {k} {a}: {diff}"
        for a in ("Un", "Un1"):
            diff = first_diff(getattr(b, a)[1:-1, 1:-1], z[f"{a}_{k}"][1:-1, 1:-1])
            assert diff is None, f"step {k} {a}: {diff}"
        assert np.array_equal(oc.norms(1)[1], z[f"rms_{k}"]), f"step {k} rms"
    assert oc.inner_iterations == inner
    if limit > 1e-20:
        assert min(inner) <= n_inner and max(inner) >= 2, inner  # early exits, not on the first iteration
    else:
        assert inner == [n_inner + 1] * 3


@pytest.mark.parametrize("fdm", [False, True])
def test_geometry_matches_reference(fdm):
    """Cell geometry, ghost coordinates, wall distance and finite-difference
    metrics built in Python (ec2d.geometry) are the reference's
    (sub_init.f90:180-204, sub_boundary.f90:275-280, sub_turbulence_SST.f90:274-365,
    sub_Finite_Difference.f90 metrics) bit for bit, on every multigrid level."""
    c = cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5)
    c.Time_Method, c.Num_Mesh = F.TIME_RK3, 3
    d = naca_case(f"r_geom_{int(fdm)}", c, ds1=1e-4, fdm=(1,) if fdm else ())
    R = refpy.RefCase(d)
    oc = O.OracleCase(d)
    names = ["vol", "si", "sj", "ni1", "ni2", "nj1", "nj2", "x", "y", "x1", "y1"]
    for lev in range(1, 4):
        g = oc.block(lev, 1).g
        for n in names + (["dw"] if lev == 1 else []) + (["Akx", "Aky", "Aix", "Aiy", "Ajac"] if fdm and lev == 1
                                                           else []):
            want = R.get(n, lev, 1)
            got = np.asarray(g.fd[n] if n in g.fd else getattr(g, n))
            # only the entries the reference defines (sub_init.f90:184-204):
            # the rest of these arrays is never written there
            if n in ("si", "ni1", "ni2"):
                got, want = got[:-1, :], want[:-1, :]
            elif n in ("sj", "nj1", "nj2"):
                got, want = got[:, :-1], want[:, :-1]
            elif n == "vol":
                got, want = got[:-1, :-1], want[:-1, :-1]
            diff = first_diff(got, want)
            assert diff is None, f"level {lev} {n}: {diff}"


@pytest.mark.parametrize("seed", range(24))
def test_random_cases(seed):
    """Seeded random multi-block cases (layout, boundary type per subface,
    scheme, flux, reconstruction, viscosity, wall temperature, time method,
    back pressure; tests/helpers.py random_case): the oracle equals the
    reference bit for bit after every pass."""
    from helpers import random_case

    d, n = random_case(seed, "r")
    run_both(d, n, ("U", "Res", "dt"))


@pytest.mark.parametrize("seed", range(12))
def test_random_turb_multiblock_cases(seed):
    """Seeded random multi-block cases with Spalart-Allmaras, SST or
    Baldwin-Lomax (tests/helpers.py random_turb_multiblock_case): the oracle
    equals the reference bit for bit, U, Res, dt and mu_t (its first ghost
    layer included) of every block after every pass."""
    from helpers import random_turb_multiblock_case

    d, n = random_turb_multiblock_case(seed, "r")
    run_both(d, n, ("U", "Res", "dt", "Amu_t"))


@pytest.mark.parametrize("seed", range(24))
def test_random_naca_cases(seed):
    """Seeded random C-grid cases (turbulence model, scheme, flux,
    reconstruction, FD-WENO5 block, mesh levels, time method;
    tests/helpers.py random_naca_case): the oracle equals the reference bit
    for bit.  LU-SGS on several levels is this build's extension, which the
    reference refuses: those seeds are skipped here."""
    from helpers import random_naca_case

    d, n, _, reference = random_naca_case(seed, "r")
    if not reference:
        pytest.skip("LU-SGS multigrid: the reference refuses it (opencfd_ec2d_v1.5.4.f90:216-221)")
    c = F.read_control(os.path.join(d, "control.in"))
    arrays = ("U", "Res", "dt") + (("Amu_t",) if c.Iflag_turbulence_model != F.TURB_NONE else ())
    run_both(d, n, arrays)


@pytest.mark.parametrize("seed", range(12))
def test_random_dual_time_cases(seed, tmp_path):
    """Seeded random dual-time cases (tests/helpers.py random_dual_case):
    the oracle equals the reference, run in a fresh process, bit for bit,
    inner-loop exits included."""
    from helpers import random_dual_case

    d, _ = random_dual_case(seed, "r")
    z, inner = ref_dual(d, 3, tmp_path)
    oc = O.OracleCase(d)
    oc.init_flow()
    for k in range(3):
        oc.main_step()
        b = oc.block()
        for a in ("U", "Res", "dt"):
            diff = first_diff(getattr(b, a), z[f"{a}_{k}"])
            assert diff is None, f"step {k} {a}: {diff}"
        assert np.array_equal(oc.norms(1)[1], z[f"rms_{k}"]), f"step {k} rms"
    assert oc.inner_iterations == inner


def test_cold_cell_stop_matches_reference():
    """The reference's T < 1e-5 stop (sub_Residual.f90:143-146) ends its
    process, so it runs in a child; two cold cells in block 2, and the first
    in its loop order (j outer, i inner) is the one both report."""
    d = vortex_case("r_cold", cases.control_vortex(iters=10))
    script = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_cold_cell.py")
    r = subprocess.run([sys.executable, script, d, "2", "7", "4", "3", "5"], capture_output=True, text=True,
                       timeout=120)
    line = [ln for ln in r.stdout.splitlines() if "T <1.d-5" in ln]
    assert line, r.stdout[-2000:]
    i, j, mb = (int(t) for t in line[0].split("!")[1].split()[:3])
    assert (i, j, mb) == (7, 4, 2)
    oc = O.OracleCase(d)
    oc.init_flow()
    U = oc.block(1, 2).U
    for ci, cj in ((7, 4), (3, 5)):
        q = U[cj + 1, ci + 1]
        q[3] = 0.5 * (q[1] ** 2 + q[2] ** 2) / q[0]
    with pytest.raises(O.SolverError) as e:
        oc.residual(1)
    assert str(e.value).startswith(f"T <1.d-5 ! {i} {j} {mb}"), str(e.value)


def test_rae2822_sample_omega_overflow(tmp_path):
    """The reference's RAE2822 sample (SST, 3 blocks, LU-SGS CFL 10; the
    fixture's isothermal wall) does not converge in the reference's own
    numerics: the omega residual grows from the third pass and overflows at
    the thirteenth while density stays near 4e-5.  The oracle follows the
    reference build bit for bit (U, Res, dt, mu_t of every block and the
    norms) through the overflow, so the divergence is the reference's, which
    the device's NaN guard must then catch (tests/test_gpu_converge.py)."""
    from test_golden import load_case

    load_case(os.path.join(os.path.dirname(__file__), "golden", "rae2822_sample_sst_isothermal.npz"), str(tmp_path))
    R, oc = run_both(str(tmp_path), 13, ("U", "Res", "dt", "Amu_t"))
    rms = oc.norms(1)[1]
    assert np.all(np.isfinite(rms[:5])) and rms[5] == np.inf, rms
    assert 1e-5 < rms[0] < 1e-4, rms
