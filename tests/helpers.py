"""Shared helpers: small synthetic cases and GPU-vs-oracle comparison."""
import os
import tempfile

import numpy as np

from ec2d import cases
from ec2d import formats as F

CASE_ROOT = os.environ.get("EC2D_TEST_CASES", os.path.join(tempfile.gettempdir(), "ec2d_test_cases"))


def case_dir(name: str) -> str:
    d = os.path.join(CASE_ROOT, name)
    os.makedirs(d, exist_ok=True)
    return d


def naca_case(name, ctl: F.Control, nx=65, ny=17, ds1=1e-3, fdm=(), **kw):
    x, y, bc, _ = cases.naca_cgrid(nx, ny, ds1=ds1, **kw)
    return cases.write_case(case_dir(name), ctl, [x], [y], [bc], fdm_blocks=fdm)


def dual_case(name, sa, limit, n_inner, fdm=()):
    """A dual-time LU-SGS case (Time_Method -1) the reference keeps finite:
    Euler at CFL 5 with physical step 0.05, SA at CFL 0.5 with 0.1, local
    pseudo-time steps."""
    c = cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5) if sa else cases.control_euler_naca(iters=10)
    c.Time_Method = F.TIME_DUAL_LU_SGS
    c.Nstep_Inner_Limit, c.Res_Inner_Limit = n_inner, limit
    c.dt_global, c.CFL, c.Iflag_local_dt = (0.1, 0.5, 1) if sa else (0.05, 5.0, 1)
    return naca_case(name, c, ds1=1e-4 if sa else 1e-3, fdm=fdm)


def vortex_case(name, ctl: F.Control, seed=3, nbx=3, nby=2, wmin=9, wmax=21, hmin=9, hmax=17, h=0.25, bc_edit=None):
    from ec2d import geometry as G

    xs, ys, bcs, field, _ = cases.vortex_multiblock(seed=seed, nbx=nbx, nby=nby, wmin=wmin, wmax=wmax, hmin=hmin,
                                                   hmax=hmax, h=h)
    if bc_edit:
        bc_edit(bcs)
    lev = G.build_levels(xs, ys, bcs, 1)
    flow = ([b.x1[1:-1, 1:-1] for b in lev[0]], [b.y1[1:-1, 1:-1] for b in lev[0]],
            [field(b.x1[1:-1, 1:-1], b.y1[1:-1, 1:-1]) for b in lev[0]])
    return cases.write_case(case_dir(name), ctl, xs, ys, bcs, flow=flow)


def first_diff(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape:
        return f"shape {a.shape} vs {b.shape}"
    bad = np.argwhere(~((a == b) | (np.isnan(a) & np.isnan(b))))
    if bad.size == 0:
        return None
    k = tuple(bad[0])
    with np.errstate(divide="ignore", invalid="ignore"):
        rel = np.nanmax(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))
    return f"{len(bad)} differ, first at {k}: {a[k]!r} vs {b[k]!r}, max rel {rel:.3e}"


BC_CHOICES = (F.BC_FARFIELD, F.BC_WALL, F.BC_SYMMETRY, F.BC_INLET, F.BC_OUTLET)


def random_case(seed: int, prefix: str):
    """A seeded random multi-block case: 1..3 x 1..3 blocks of 9..21 x 9..17
    points, each outer subface a random boundary type (far field, wall,
    symmetry, inlet, outlet), random scheme / flux / reconstruction, inviscid
    or laminar with an adiabatic or isothermal wall, LU-SGS, RK3 or first-order
    Euler, with or without a back pressure.  Returns (case directory, steps)."""
    rng = np.random.default_rng(seed)
    c = cases.control_vortex(iters=10)
    c.Iflag_Scheme = int(rng.integers(0, 7))
    c.Iflag_Flux = int(rng.integers(1, 7))
    c.IFlag_Reconstruction = int(rng.integers(0, 3))
    if rng.random() < 0.5:
        c.If_viscous, c.Re = 1, float(10.0 ** rng.uniform(3, 5))
        c.Twall = 300.0 if rng.random() < 0.5 else -1.0
    tm = int(rng.choice([F.TIME_LU_SGS, F.TIME_RK3, F.TIME_EULER1]))
    c.Time_Method = tm
    c.CFL = {F.TIME_LU_SGS: 2.0, F.TIME_RK3: 0.5, F.TIME_EULER1: 0.3}[tm]
    if rng.random() < 0.5:
        c.P_outlet = 1.0 / (1.4 * 0.5 * 0.5) * float(rng.uniform(0.95, 1.0))
    kinds = [int(v) for v in rng.integers(0, len(BC_CHOICES), size=64)]

    def edit(bcs):
        k = 0
        for subs in bcs:
            for s in subs:
                if s[6] == F.BC_FARFIELD:
                    s[6] = BC_CHOICES[kinds[k % len(kinds)]]
                    k += 1

    d = vortex_case(f"{prefix}_rand_{seed}", c, seed=seed, nbx=int(rng.integers(1, 4)), nby=int(rng.integers(1, 4)),
                    bc_edit=edit)
    return d, 3


def random_turb_multiblock_case(seed: int, prefix: str):
    """A seeded random multi-block case with a turbulence model: 2..3 x 1..3
    blocks of 9..21 x 9..17 points, every outer subface a random boundary
    type and the first one a wall (so the wall distance, the model's ghost
    rules and mu_t on the first ghost layer act across block interfaces,
    the 3-block RAE2822 file's reversed orientation included when the layout
    has it), Spalart-Allmaras, SST (isothermal wall) or Baldwin-Lomax, random
    scheme / flux / reconstruction, LU-SGS or RK3, freestream start.  Returns
    (case directory, steps)."""
    rng = np.random.default_rng(5000 + seed)
    turb = int(rng.choice([F.TURB_SA, F.TURB_SST, F.TURB_BL]))
    c = cases.control_vortex(iters=10)
    c.Iflag_init = 0
    c.If_viscous, c.Re, c.Pr = 1, float(10.0 ** rng.uniform(3.5, 5.0)), 0.72
    c.Iflag_turbulence_model = turb
    c.vt_inf = 3.0
    c.Twall = 300.0 if (turb == F.TURB_SST or rng.random() < 0.5) else -1.0
    c.Iflag_Scheme = int(rng.integers(0, 7))
    c.Iflag_Flux = int(rng.integers(1, 7))
    c.IFlag_Reconstruction = int(rng.integers(0, 3))
    tm = int(rng.choice([F.TIME_LU_SGS, F.TIME_RK3]))
    c.Time_Method = tm
    c.CFL = {F.TIME_LU_SGS: 1.0, F.TIME_RK3: 0.3}[tm]
    kinds = [int(v) for v in rng.integers(0, len(BC_CHOICES), size=64)]

    def edit(bcs):
        k = 0
        for subs in bcs:
            for s in subs:
                if s[6] == F.BC_FARFIELD:
                    s[6] = F.BC_WALL if k == 0 else BC_CHOICES[kinds[k % len(kinds)]]
                    k += 1

    d = vortex_case(f"{prefix}_rturb_{seed}", c, seed=seed, nbx=int(rng.integers(2, 4)), nby=int(rng.integers(1, 4)),
                    bc_edit=edit)
    return d, 3


def random_naca_case(seed: int, prefix: str):
    """A seeded random single-block C-grid case: grid 49..73 x 13..21 points,
    inviscid, laminar, Baldwin-Lomax, SA or SST (isothermal wall), random
    scheme / flux / reconstruction, optionally an FD-WENO5 block, one to three
    mesh levels with LU-SGS, RK3 or first-order Euler.  Returns (case
    directory, steps, exact, reference) - exact: whether the GPU must match the
    oracle bit for bit (no exp / pow / tanh of Baldwin-Lomax, SA or SST); reference: whether
    the reference runs the case (it refuses LU-SGS on more than one level)."""
    rng = np.random.default_rng(1000 + seed)
    turb = int(rng.choice([F.TURB_NONE, F.TURB_NONE, F.TURB_BL, F.TURB_SA, F.TURB_SST]))
    nm = int(rng.integers(1, 4))
    tm = int(rng.choice([F.TIME_LU_SGS, F.TIME_RK3, F.TIME_EULER1]))
    if turb == F.TURB_NONE and rng.random() < 0.5:
        c = cases.control_euler_naca(iters=10)
    else:
        c = cases.control_sa_naca(iters=10, Re=float(10.0 ** rng.uniform(4.5, 5.5)), Ma=0.5)
        c.Iflag_turbulence_model = turb
        if turb == F.TURB_SST:
            c.Twall = 300.0
    c.Iflag_Scheme = int(rng.integers(0, 7))
    c.Iflag_Flux = int(rng.integers(1, 7))
    c.IFlag_Reconstruction = int(rng.integers(0, 3))
    c.Time_Method, c.Num_Mesh = tm, nm
    c.CFL = {F.TIME_LU_SGS: 2.0, F.TIME_RK3: 0.3, F.TIME_EULER1: 0.2}[tm]
    nx = int(rng.choice([49, 57, 65, 73]))
    ny = int(rng.choice([13, 17, 21]))
    fdm = (1,) if rng.random() < 0.3 else ()
    d = naca_case(f"{prefix}_rnaca_{seed}", c, nx=nx, ny=ny, ds1=1e-4 if c.If_viscous else 1e-3, fdm=fdm)
    exact = turb == F.TURB_NONE
    return d, 3, exact, not (tm == F.TIME_LU_SGS and nm > 1)


def random_dual_case(seed: int, prefix: str):
    """A seeded random dual-time LU-SGS case (Time_Method -1) on the C-grid:
    Euler, laminar or SA, random scheme / flux / reconstruction, optionally an
    FD-WENO5 block, 3..8 inner iterations, an inner limit that ends some inner
    loops early or none, physical steps of 0.03..0.08 (Euler, laminar) or 0.1
    (SA).  Returns (case directory, exact): exact without SA."""
    rng = np.random.default_rng(2000 + seed)
    kind = int(rng.integers(0, 3))  # 0 Euler, 1 laminar, 2 SA
    c = cases.control_sa_naca(iters=10, Re=1e5, Ma=0.5) if kind else cases.control_euler_naca(iters=10)
    if kind == 1:
        c.Iflag_turbulence_model = F.TURB_NONE
    c.Time_Method = F.TIME_DUAL_LU_SGS
    c.Iflag_Scheme = int(rng.integers(0, 7))
    c.Iflag_Flux = int(rng.integers(1, 7))
    c.IFlag_Reconstruction = int(rng.integers(0, 3))
    c.Nstep_Inner_Limit = int(rng.integers(3, 9))
    c.Res_Inner_Limit = float(10.0 ** rng.uniform(-3.0, -1.5)) if rng.random() < 0.6 else 1e-30
    c.Iflag_local_dt = 1
    if kind == 2:
        c.dt_global, c.CFL = 0.1, 0.5
    else:
        c.dt_global, c.CFL = float(rng.uniform(0.03, 0.08)), float(rng.uniform(1.0, 5.0))
    fdm = (1,) if rng.random() < 0.3 else ()
    d = naca_case(f"{prefix}_rdual_{seed}", c, ds1=1e-4 if kind else 1e-3, fdm=fdm)
    return d, kind != 2
