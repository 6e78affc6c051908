"""GPU path vs the C oracle, on identical inputs.

Bar: bit-exact FP64 state (U), residual, time step and spectral radii for the
Euler, laminar and multigrid paths; the residual norms (whose summation order
differs: a fixed tree on the GPU, a serial loop in the reference) within
1e-12 relative.  The SA path evaluates exp and pow with the GPU's math library
instead of glibc, so it is held to 1e-12 relative per iteration instead of
bit equality.
"""
import itertools

import numpy as np
import pytest

import oracle as O
from ec2d import cases
from ec2d import formats as F
from ec2d import solver
from helpers import first_diff, naca_case, vortex_case

pytestmark = pytest.mark.gpu


def _compare_steps(d, nsteps, arrays=("U", "Res", "dt", "Lci", "Lcj"), exact=True, rtol=1e-12, option=None,
                   tile_rows=4, norm_rtol=1e-12, tile_cols=64):
    ref = O.OracleCase(d)
    ref.init_flow()
    case = solver.Case.from_dir(d)
    with solver.Batch([case], tile_rows=tile_rows, tile_cols=tile_cols) as b:
        b.init_flow()
        for mb in range(1, len(case.levels[0]) + 1):
            diff = first_diff(b.get("U", 0, 1, mb), ref.block(1, mb).U)
            assert diff is None, f"initial U block {mb}: {diff}"
        for k in range(nsteps):
            ref.main_step()
            b.main_step()
            for mb in range(1, len(case.levels[0]) + 1):
                rb = ref.block(1, mb)
                for a in arrays:
                    g = b.get(a, 0, 1, mb)
                    r = getattr(rb, a)
                    if a in ("Lci", "Lcj", "Lvi", "Lvj"):
                        g, r = g[:-1, :-1], r[:-1, :-1]
                    if exact:
                        diff = first_diff(g, r)
                        assert diff is None, f"step {k} block {mb} {a}: {diff}"
                    else:
                        np.testing.assert_allclose(g, r, rtol=rtol, atol=1e-300, err_msg=f"step {k} block {mb} {a}")
            gmax, grms = b.res_norms()
            rmax, rrms = ref.norms()
            nv = case.ctl.nvar
            np.testing.assert_allclose(grms[:nv], rrms[:nv], rtol=norm_rtol, atol=0)
            np.testing.assert_allclose(gmax[:nv], rmax[:nv], rtol=norm_rtol if not exact else 0, atol=0)


def test_euler_cgrid_roe_muscl3():
    d = naca_case("g_euler", cases.control_euler_naca(iters=10))
    _compare_steps(d, 6)


@pytest.mark.parametrize("tile_rows", [1, 2, 8])
def test_euler_tile_shapes(tile_rows):
    d = naca_case("g_euler_t", cases.control_euler_naca(iters=10))
    _compare_steps(d, 2, tile_rows=tile_rows)


@pytest.mark.parametrize("rows", [8, 16])
@pytest.mark.parametrize("kind", ["euler", "laminar", "sa", "sst", "multiblock", "fdm_sa", "mg_lusgs_sa"])
def test_tile_32_columns(kind, rows):
    """The 32 x 8 and 32 x 16 residual tiles (viscous kernels rebuild p from
    staged d and T and read the cell centres from their planes; the j-face
    fluxes reuse the i-faces' LDS) against the oracle: tile edges fall on
    other cells than with 64-wide tiles."""
    arrays, exact, nsteps = ("U", "Res", "dt", "Lci", "Lcj"), True, 3
    if kind == "euler":
        d = naca_case("g_t32_e", cases.control_euler_naca(iters=10))
    elif kind == "laminar":
        c = cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5)
        c.Iflag_turbulence_model, c.Twall = F.TURB_NONE, 350.0
        d, arrays = naca_case("g_t32_lam", c, ds1=1e-4), arrays + ("Lvi", "Lvj")
    elif kind == "sa":
        c = cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5)
        d, arrays, exact = naca_case("g_t32_sa", c, ds1=1e-4, nx=129, ny=49), arrays + ("Lvi", "Lvj"), False
    elif kind == "sst":
        c = _sst()
        c.Twall = 300.0
        d, arrays, exact = naca_case("g_t32_sst", c, ds1=1e-4), arrays + ("Amu_t",), False
    elif kind == "multiblock":
        c = cases.control_vortex(iters=10)
        c.If_viscous, c.Re, c.Pr = 1, 1e4, 0.72
        d = vortex_case("g_t32_mb", c)
    elif kind == "fdm_sa":
        c = cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5)
        d, arrays, exact = naca_case("g_t32_fdm", c, ds1=1e-4, fdm=(1,)), ("U", "Res", "dt"), False
    else:
        c = cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5, num_mesh=2, time_method=F.TIME_LU_SGS)
        d, arrays, exact = naca_case("g_t32_mg", c, ds1=1e-4, fdm=(1,)), ("U", "Res"), False
    if exact:
        _compare_steps(d, nsteps, arrays, tile_rows=rows, tile_cols=32)
    else:
        _compare_steps(d, nsteps, arrays, exact=False, rtol=1e-11, tile_rows=rows, tile_cols=32)


SCHEMES = range(7)
FLUXES = range(1, 7)
RECONS = range(3)


@pytest.mark.parametrize("scheme,flux,recon", [(s, 4, 0) for s in SCHEMES] + [(5, f, 0) for f in FLUXES] +
                         [(s, f, r) for s, f, r in itertools.product((1, 5), (2, 3, 6), (1, 2))])
def test_scheme_flux_recon(scheme, flux, recon):
    c = cases.control_euler_naca(iters=10)
    c.Iflag_Scheme, c.Iflag_Flux, c.IFlag_Reconstruction = scheme, flux, recon
    d = naca_case(f"g_sfr_{scheme}_{flux}_{recon}", c)
    _compare_steps(d, 2)


def test_laminar_isothermal_wall():
    c = cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5)
    c.Iflag_turbulence_model = F.TURB_NONE
    c.Twall = 350.0
    d = naca_case("g_lam", c, ds1=1e-4)
    _compare_steps(d, 4, arrays=("U", "Res", "dt", "Lci", "Lcj", "Lvi", "Lvj"))


def test_sa_cgrid():
    c = cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5)
    d = naca_case("g_sa", c, ds1=1e-4)
    _compare_steps(d, 4, arrays=("U", "Res", "dt", "Lci", "Lcj", "Lvi", "Lvj"), exact=False, rtol=1e-11)


def test_multiblock_vortex_ausm():
    d = vortex_case("g_mb", cases.control_vortex(iters=10))
    _compare_steps(d, 4)


def _symm(bcs):
    for subs in bcs:
        for s in subs:
            if s[6] == F.BC_FARFIELD and s[1] in (2, 4):
                s[6] = F.BC_SYMMETRY


def test_multiblock_symmetry():
    d = vortex_case("g_mb_sym", cases.control_vortex(iters=10), bc_edit=_symm)
    _compare_steps(d, 3)


def test_multiblock_inlet_outlet():
    c = cases.control_vortex(iters=10)
    c.P_outlet = 1.0 / (1.4 * 0.5 * 0.5) * 0.98

    def edit(bcs):
        for subs in bcs:
            for s in subs:
                if s[6] == F.BC_FARFIELD and s[1] == 1:
                    s[6] = F.BC_INLET
                if s[6] == F.BC_FARFIELD and s[1] == 3:
                    s[6] = F.BC_OUTLET

    d = vortex_case("g_mb_io", c, bc_edit=edit)
    _compare_steps(d, 3)


@pytest.mark.parametrize("tm,nm", [(F.TIME_RK3, 2), (F.TIME_RK3, 3), (F.TIME_EULER1, 3)])
def test_multigrid_euler(tm, nm):
    c = cases.control_euler_naca(iters=10)
    c.Time_Method, c.Num_Mesh, c.CFL = tm, nm, 0.3
    d = naca_case(f"g_mg_{tm}_{nm}", c)
    _compare_steps(d, 3, arrays=("U", "Res"))


def test_graph_replay_matches_eager():
    d = naca_case("g_graph", cases.control_euler_naca(iters=10))
    with solver.Batch([solver.Case.from_dir(d)]) as b:
        b.init_flow()
        b.run(5)
        U_graph = b.get("U")
        k1, t1 = b.time()
    ref = O.OracleCase(d)
    ref.init_flow()
    for _ in range(5):
        ref.main_step()
    assert first_diff(U_graph, ref.block().U) is None
    assert k1 == 5 and t1 == ref.tt[0]


def test_batch_of_cases_equals_separate_runs():
    ds = []
    for k, (ma, aoa) in enumerate([(0.5, 1.25), (0.6, -1.0), (0.4, 3.0)]):
        c = cases.control_euler_naca(iters=10, Ma=ma, aoa=aoa)
        ds.append(naca_case(f"g_batch_{k}", c, nx=65 if k != 1 else 97, ny=17 if k != 2 else 25))
    cs = [solver.Case.from_dir(d) for d in ds]
    with solver.Batch(cs) as b:
        b.init_flow()
        b.run(3)
        Us = [b.get("U", k) for k in range(3)]
        norms = [b.res_norms(k)[1] for k in range(3)]
    for k, d in enumerate(ds):
        ref = O.OracleCase(d)
        ref.init_flow()
        for _ in range(3):
            ref.main_step()
        assert first_diff(Us[k], ref.block().U) is None, k
        np.testing.assert_allclose(norms[k][:4], ref.norms()[1][:4], rtol=1e-12)


def test_fdm_weno5_euler():
    d = naca_case("g_fdm_e", cases.control_euler_naca(iters=10), fdm=(1,))
    _compare_steps(d, 3, arrays=("U", "Res", "dt"))


def test_fdm_weno5_sa():
    c = cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5)
    d = naca_case("g_fdm_sa", c, ds1=1e-4, fdm=(1,))
    _compare_steps(d, 3, arrays=("U", "Res", "dt"), exact=False, rtol=1e-11)


@pytest.mark.parametrize("nm", [2, 3])
def test_multigrid_lusgs_sa(nm):
    """LU-SGS-smoothed V-cycle (extension), SA on the fine level, FD-WENO5 block."""
    c = cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5, num_mesh=nm, time_method=F.TIME_LU_SGS)
    d = naca_case(f"g_mg_lusgs_{nm}", c, ds1=1e-4, fdm=(1,))
    _compare_steps(d, 3, arrays=("U", "Res"), exact=False, rtol=1e-11)


def _sst(**kw):
    c = cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5, **kw)
    c.Iflag_turbulence_model = F.TURB_SST
    return c


def test_sst_isothermal_wall():
    """SST k-omega (device tanh in F1 / F2, hence the tolerance)."""
    c = _sst()
    c.Twall = 300.0
    d = naca_case("g_sst_iso", c, ds1=1e-4)
    _compare_steps(d, 4, arrays=("U", "Res", "dt", "Lci", "Lcj", "Lvi", "Lvj", "Amu_t"), exact=False, rtol=1e-11)


def test_sst_adiabatic_wall():
    """Adiabatic wall: the reference's omega ghost reads an unset density on
    this branch; device and oracle both take 0 (parity with the reference
    unpinned here, pinned on the isothermal wall)."""
    d = naca_case("g_sst_adi", _sst(), ds1=1e-4)
    _compare_steps(d, 3, arrays=("U", "Res", "dt", "Amu_t"), exact=False, rtol=1e-11)


def test_sst_fdm_weno5():
    c = _sst()
    c.Twall = 300.0
    d = naca_case("g_sst_fdm", c, ds1=1e-4, fdm=(1,))
    _compare_steps(d, 3, arrays=("U", "Res", "dt"), exact=False, rtol=1e-11)


def test_sst_multigrid_lusgs():
    c = _sst(num_mesh=2, time_method=F.TIME_LU_SGS)
    c.Twall = 300.0
    d = naca_case("g_sst_mg", c, ds1=1e-4)
    _compare_steps(d, 3, arrays=("U", "Res"), exact=False, rtol=1e-11)


@pytest.mark.parametrize("kind", ["euler", "laminar", "sa"])
def test_force_coefficients(kind):
    """ec2d_force_coefficients against the oracle's serial integral (the
    device sums the faces in a fixed tree, hence a relative tolerance)."""
    if kind == "euler":
        c, ds1, rtol = cases.control_euler_naca(iters=10, aoa=2.0), 1e-3, 1e-13
    else:
        c, ds1 = cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5), 1e-4
        if kind == "laminar":
            c.Iflag_turbulence_model, c.Twall, rtol = F.TURB_NONE, 350.0, 1e-12
        else:
            rtol = 1e-9  # the SA state itself agrees to 1e-11 per step
    d = naca_case(f"g_forces_{kind}", c, ds1=ds1)
    ref = O.OracleCase(d)
    ref.init_flow()
    with solver.Batch([solver.Case.from_dir(d)]) as b:
        b.init_flow()
        for _ in range(4):
            ref.main_step()
            b.main_step()
        cl, cd, f = b.force_coefficients(0)
    rf = ref.forces()
    rcl, rcd = ref.force_coefficients()
    np.testing.assert_allclose(f, rf, rtol=rtol, atol=1e-15 * np.abs(rf).max())
    np.testing.assert_allclose([cl, cd], [rcl, rcd], rtol=10 * rtol)
    assert abs(cl) > 1e-6  # a lifting state, not a trivial zero


def test_euler_tall_block_multiwave_sweep():
    """72 rows of cells: the LU-SGS sweep runs one lane per row over two waves
    (exchanging the boundary row through LDS every step)."""
    d = naca_case("g_tall", cases.control_euler_naca(iters=10), nx=129, ny=73)
    _compare_steps(d, 3)


def test_sa_tall_block_multiwave_sweep():
    c = cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5)
    d = naca_case("g_tall_sa", c, nx=129, ny=73, ds1=1e-4)
    _compare_steps(d, 2, arrays=("U", "Res", "dt"), exact=False, rtol=1e-11)


def test_autotune_keeps_results_and_caches(tmp_path):
    """ec2d_autotune picks launch shapes by timing; the results of the steps
    that follow stay bit-identical to the oracle, and a second call reads the
    cache instead of measuring."""
    c = cases.control_euler_naca(iters=10)
    c.Time_Method, c.Num_Mesh, c.CFL = F.TIME_LU_SGS, 2, 2.0
    d = naca_case("g_tune", c, nx=129, ny=73)
    cache = str(tmp_path / "tune.txt")
    ref = O.OracleCase(d)
    ref.init_flow()
    with solver.Batch([solver.Case.from_dir(d)]) as b:
        b.init_flow()
        rep = b.autotune(cache, reps=2)
        assert rep.count("measured") == 2, rep
        for _ in range(3):
            ref.main_step()
            b.main_step()
        assert first_diff(b.get("U"), ref.block().U) is None
        b.run(2)
        for _ in range(2):
            ref.main_step()
        assert first_diff(b.get("U"), ref.block().U) is None
        rep2 = b.autotune(cache, reps=2)
        assert rep2.count("cached") == 2, rep2


def test_autotune_cache_of_a_large_batch(tmp_path):
    """A batch of 140 small cases of several sizes (its list of block
    shapes runs to about 900 characters): a fresh batch of the same cases
    finds the first batch's choice in the cache file, which holds one short
    line per level configuration."""
    dirs = [naca_case(f"g_tune_many_{k}", cases.control_euler_naca(iters=10), nx=33 + 8 * (k % 3), ny=9)
            for k in range(140)]
    cache = str(tmp_path / "tune.txt")
    for want in ("measured", "cached"):
        with solver.Batch([solver.Case.from_dir(d) for d in dirs]) as b:
            b.init_flow()
            rep = b.autotune(cache, reps=1)
            assert rep.count(want) == 1, rep
    with open(cache) as f:
        lines = f.read().splitlines()
    assert len(lines) == 1 and len(lines[0]) < 120, lines


@pytest.mark.parametrize("lpc,split,ny", [(1, 2, 73), (2, 2, 73), (2, 4, 73), (2, 4, 129)])
def test_split_sweep_tall_block_multigrid(lpc, split, ny):
    """Rows of a block split between two or four workgroups (one CU each):
    neighbouring parts pass the boundary row through global memory; results
    stay bit-identical to the oracle (LU-SGS-smoothed 2-level V-cycle).  With
    72 rows split in four the last part holds padding rows only."""
    c = cases.control_euler_naca(iters=10)
    c.Time_Method, c.Num_Mesh, c.CFL = F.TIME_LU_SGS, 2, 2.0
    d = naca_case(f"g_split_{lpc}_{split}_{ny}", c, nx=129, ny=ny)
    ref = O.OracleCase(d)
    ref.init_flow()
    with solver.Batch([solver.Case.from_dir(d)]) as b:
        b.init_flow()
        b.lusgs_shape(1, lpc, split)
        b.lusgs_shape(2, lpc, 1)
        for k in range(3):
            ref.main_step()
            b.main_step()
            assert first_diff(b.get("U"), ref.block().U) is None, f"step {k}"
        b.run(3)
        for _ in range(3):
            ref.main_step()
        assert first_diff(b.get("U"), ref.block().U) is None
        b.res_norms()  # reports a timed-out cross-workgroup wait, if any


@pytest.mark.parametrize("split,ny", [(2, 65), (4, 97)])
def test_split_sweep_sa(split, ny):
    c = cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5)
    d = naca_case(f"g_split_sa_{split}", c, nx=129, ny=ny, ds1=1e-4)
    ref = O.OracleCase(d)
    ref.init_flow()
    with solver.Batch([solver.Case.from_dir(d)]) as b:
        b.init_flow()
        b.lusgs_shape(1, 2, split)
        for k in range(4):
            ref.main_step()
            b.main_step()
            np.testing.assert_allclose(b.get("U"), ref.block().U, rtol=1e-11, atol=1e-300, err_msg=f"step {k}")
        b.res_norms()


@pytest.mark.parametrize("split,hmax", [(2, 49), (4, 81)])
def test_split_sweep_uneven_multiblock(split, hmax):
    """64 uneven blocks, each split in two or four (256 workgroups): some
    upper parts hold padding rows only and still relay nothing but the
    protocol."""
    from helpers import case_dir

    d = case_dir(f"g_split_mb_{split}")
    cases.write_vortex_case(d, cases.control_vortex(iters=10), seed=5, nbx=8, nby=8, wmin=9, wmax=41, hmin=9,
                            hmax=hmax, h=0.25)
    ref = O.OracleCase(d)
    ref.init_flow()
    case = solver.Case.from_dir(d)
    with solver.Batch([case]) as b:
        b.init_flow()
        b.lusgs_shape(1, 2, split)
        for k in range(3):
            ref.main_step()
            b.main_step()
            for mb in range(1, 65):
                diff = first_diff(b.get("U", 0, 1, mb), ref.block(1, mb).U)
                assert diff is None, f"step {k} block {mb}: {diff}"
        b.res_norms()


@pytest.mark.parametrize("tm", [F.TIME_RK3, F.TIME_EULER1, F.TIME_LU_SGS])
def test_multigrid_scratch_arrays(tm):
    """The V-cycle fuses comput_force_function, Add_force_function and Set_Un
    into the coarse residual, both restrictions into one launch, and
    comput_delt_U, prolong_U and comput_new_U into one pass; every array the
    reference's cycle leaves behind (Un, QF, deltU on every level, U and Res)
    still equals the oracle's bit for bit."""
    c = cases.control_euler_naca(iters=10)
    c.Time_Method, c.Num_Mesh, c.CFL = tm, 3, (2.0 if tm == F.TIME_LU_SGS else 0.3)
    d = naca_case(f"g_mg_scratch_{tm}", c)
    ref = O.OracleCase(d)
    ref.init_flow()
    with solver.Batch([solver.Case.from_dir(d)]) as b:
        b.init_flow()
        for k in range(2):
            ref.main_step()
            b.main_step()
            for nm in (1, 2, 3):
                rb = ref.block(nm, 1)
                for a in ("U", "Un", "QF", "deltU") + (("Res",) if nm == 1 else ()):
                    if nm == 1 and a == "QF":
                        continue
                    diff = first_diff(b.get(a, 0, nm, 1), getattr(rb, a))
                    assert diff is None, f"pass {k} level {nm} {a}: {diff}"


@pytest.mark.parametrize("kind,wg", [("euler_64rows", 0), ("sa_48rows", 0), ("multiblock", 0), ("sst_48rows", 0),
                                     ("euler_64rows", -3), ("sa_48rows", -3), ("sst_48rows", -3)])
def test_fused_lusgs_preparation(kind, wg):
    """k_lusgs_fused (the records made in the sweep's workgroup, for blocks of
    at most 64 rows; 6 producer waves, or 3 with wg = -3) against the oracle:
    bit-exact for Euler, 1e-11 with the turbulence models; eager passes, then
    graph replays."""
    if kind == "euler_64rows":
        d, exact = naca_case("g_fused_e", cases.control_euler_naca(iters=10), nx=129, ny=65), True
    elif kind == "sa_48rows":
        d, exact = naca_case("g_fused_sa", cases.control_sa_naca(iters=10, Re=1e6, Ma=0.6), nx=129, ny=49,
                             ds1=1e-5), False
    elif kind == "sst_48rows":
        c = _sst()
        c.Twall = 300.0
        d, exact = naca_case("g_fused_sst", c, nx=97, ny=49, ds1=1e-4), False
    else:
        d, exact = vortex_case("g_fused_mb", cases.control_vortex(iters=10), nbx=3, nby=3, wmin=9, wmax=41, hmin=9,
                               hmax=65), True
    case = solver.Case.from_dir(d)
    ref = O.OracleCase(d)
    ref.init_flow()
    nb = len(case.levels[0])
    with solver.Batch([case]) as b:
        b.init_flow()
        b.lusgs_shape(1, 1, wg)
        for k in range(3):
            ref.main_step()
            b.main_step()
            for mb in range(1, nb + 1):
                g, r = b.get("U", 0, 1, mb), ref.block(1, mb).U
                if exact:
                    assert first_diff(g, r) is None, f"pass {k} block {mb}: {first_diff(g, r)}"
                else:
                    np.testing.assert_allclose(g, r, rtol=1e-11, atol=1e-300, err_msg=f"pass {k} block {mb}")
        b.run(3)
        for _ in range(3):
            ref.main_step()
        for mb in range(1, nb + 1):
            g, r = b.get("U", 0, 1, mb), ref.block(1, mb).U
            if exact:
                assert first_diff(g, r) is None, f"graph, block {mb}: {first_diff(g, r)}"
            else:
                np.testing.assert_allclose(g, r, rtol=1e-11, atol=1e-300, err_msg=f"graph, block {mb}")
        b.res_norms()  # reports a wait that ran out of its budget, if any


def test_fused_lusgs_refuses_tall_blocks():
    from ec2d import capi

    d = naca_case("g_fused_tall", cases.control_euler_naca(iters=10), nx=129, ny=73)
    with solver.Batch([solver.Case.from_dir(d)]) as b:
        with pytest.raises(capi.Ec2dError, match="at most 64 rows"):
            b.lusgs_shape(1, 1, 0)


def test_split_sweep_timeout_is_reported_and_cleared():
    """The error path of a split sweep's bounded waits, forced from the host:
    with a wait budget of 0 every receiving wave gives up at its first slot
    (the kernels still end normally).  ec2d_run then fails with the timeout,
    ec2d_get_array refuses the invalid arrays, and the error stays set; after
    ec2d_init_flow (rings back to the sentinel, error words cleared) and the
    default budget, the run is bit-identical to the oracle again."""
    from ec2d import capi

    c = cases.control_euler_naca(iters=10)
    c.Time_Method, c.CFL = F.TIME_LU_SGS, 2.0
    d = naca_case("g_split_timeout", c, nx=129, ny=97)
    ref = O.OracleCase(d)
    ref.init_flow()
    with solver.Batch([solver.Case.from_dir(d)]) as b:
        b.init_flow()
        b.lusgs_shape(1, 1, 2)
        b.set_wait_budget(0)
        with pytest.raises(capi.Ec2dError, match="timed out"):
            b.run(2)
        with pytest.raises(capi.Ec2dError, match="timed out"):
            b.get("U")
        with pytest.raises(capi.Ec2dError, match="timed out"):
            b.synchronize()
        b.set_wait_budget(solver.Batch.WAIT_BUDGET_DEFAULT)
        b.init_flow()
        b.run(3)
        for _ in range(3):
            ref.main_step()
        assert first_diff(b.get("U"), ref.block().U) is None
        b.res_norms()
