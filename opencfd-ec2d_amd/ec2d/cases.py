"""Synthetic cases: NACA 4-digit C-grids, a multi-block vortex box, seeded
ensembles.  Each generator writes a case directory in the reference's own file
formats (control.in, Mesh2d.dat, bc2d_1.in, optionally flow2d.dat / FDM.in),
so the same directory runs on the reference Fortran and on this package.

Benchmark workloads (sizes in points; a block has (nx-1)(ny-1) cells):

* naca_cgrid(257, 65, ...)  NACA 0012, Euler, Roe/MUSCL3/LU-SGS (16,384 cells)
* naca_cgrid(513, 129, ...) NACA 0012, SA-RANS, 3-level multigrid (65,536 cells)
* vortex_multiblock(seed)   64 sheared/perturbed Cartesian blocks, AUSM/MUSCL3
* ensemble_params(seed)     256 NACA mpxx SA-RANS cases of uneven sizes
"""
from __future__ import annotations

import math
import os
from typing import List, Optional, Tuple

import numpy as np

from . import formats as F


# --------------------------------------------------------------------------
# NACA 4-digit section
# --------------------------------------------------------------------------
def naca4(m: float, p: float, t: float, xc: np.ndarray):
    """Upper and lower surface of a NACA mpxx section at chord stations xc.

    Thickness uses the closed-trailing-edge coefficient (-0.1036) so that the
    wake cut of the C-grid starts from a single trailing-edge point.
    """
    yt = 5.0 * t * (0.2969 * np.sqrt(xc) - 0.1260 * xc - 0.3516 * xc ** 2 + 0.2843 * xc ** 3 - 0.1036 * xc ** 4)
    if m > 0.0:
        yc = np.where(xc < p, m / p ** 2 * (2 * p * xc - xc ** 2), m / (1 - p) ** 2 * ((1 - 2 * p) + 2 * p * xc - xc ** 2))
        dyc = np.where(xc < p, 2 * m / p ** 2 * (p - xc), 2 * m / (1 - p) ** 2 * (p - xc))
    else:
        yc = np.zeros_like(xc)
        dyc = np.zeros_like(xc)
    th = np.arctan(dyc)
    xu, yu = xc - yt * np.sin(th), yc + yt * np.cos(th)
    xl, yl = xc + yt * np.sin(th), yc - yt * np.cos(th)
    return xu, yu, xl, yl


def _layer_distances(ny: int, ds1: float, ratio: float, far: float) -> Tuple[np.ndarray, dict]:
    """Wall-normal distances d[0..ny-1] with d[0]=0, d[1]=ds1, d[-1]=far.

    Layers grow by `ratio` from the wall.  When a constant `ratio` overshoots
    `far` within ny-1 layers, the outer layers switch at layer K to a smaller
    constant ratio q solved to land exactly on `far` (K is the largest layer
    count for which q stays >= 1).  When `ratio` undershoots, a constant
    ratio is solved for all layers instead.  The returned info records what
    was used.
    """
    n = ny - 1

    def geo_sum(s0, q, k):
        return s0 * k if abs(q - 1.0) < 1e-15 else s0 * (q ** k - 1.0) / (q - 1.0)

    def solve_q(s0, k, target):
        lo, hi = 1.0, 3.0
        if geo_sum(s0, 1.0, k) > target:
            lo, hi = 0.3, 1.0
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            if geo_sum(s0, mid, k) > target:
                hi = mid
            else:
                lo = mid
        return 0.5 * (lo + hi)

    if geo_sum(ds1, ratio, n) <= far:
        q = solve_q(ds1, n, far)
        sp = ds1 * q ** np.arange(n)
        info = {"inner_ratio": q, "switch_layer": n, "outer_ratio": q}
    else:
        best = None
        for K in range(n - 1, 0, -1):
            inner = geo_sum(ds1, ratio, K)
            if inner >= far:
                continue
            s0 = ds1 * ratio ** K
            q = solve_q(s0, n - K, far - inner)
            if 1.0 <= q <= ratio:
                best = (K, q)
                break
        K, q = best
        sp = np.concatenate([ds1 * ratio ** np.arange(K), ds1 * ratio ** K * q ** np.arange(n - K)])
        info = {"inner_ratio": ratio, "switch_layer": K, "outer_ratio": q}
    d = np.concatenate([[0.0], np.cumsum(sp)])
    d *= far / d[-1]
    return d, info


def naca_cgrid(nx: int, ny: int, m: float = 0.0, p: float = 0.4, t: float = 0.12, ds1: float = 1e-3,
               ratio: float = 1.15, far: float = 20.0, n_wake: Optional[int] = None, mg_align: int = 4,
               wake_ds_growth: float = 0.0):
    """Single-block C-grid around a NACA 4-digit section.

    i runs along the C (lower wake from the outflow, lower surface, leading
    edge, upper surface, upper wake); j runs from the surface/wake cut (j=1)
    to the far field (j=ny).  The grid is right-handed (positive Jacobian).
    Wall-normal distribution from _layer_distances; surface points use cosine
    clustering; the wake stretches geometrically from the trailing-edge
    spacing to x = 1 + far.  Subface limits satisfy (idx-1) % mg_align == 0 so
    the grid admits log2(mg_align)+1 multigrid levels.

    wake_ds_growth > 0 lets the first wall-normal spacing of the wake columns
    grow linearly with the distance x - 1 behind the trailing edge, ds1 +
    wake_ds_growth * (x - 1) (at most half the uniform spacing far / (ny - 1)),
    each column's layers redistributed to reach the far field (0: every column
    uses the surface distribution).

    Returns (x, y, bc, info) with x, y of shape (ny, nx).
    """
    if n_wake is None:
        n_wake = (nx - 1) // 5
        n_wake -= n_wake % mg_align
    n_af = nx - 2 * n_wake            # airfoil nodes incl. both trailing-edge copies
    assert n_af % 2 == 1 and n_af >= 9, (nx, n_wake)
    half = (n_af + 1) // 2            # nodes from TE to LE on one side
    beta = np.linspace(0.0, math.pi, half)
    xc = 0.5 * (1.0 - np.cos(beta))   # 0 .. 1, clustered at both ends
    xu, yu, xl, yl = naca4(m, p, t, xc)
    # lower surface TE -> LE, then upper LE -> TE (LE shared)
    sx = np.concatenate([xl[::-1], xu[1:]])
    sy = np.concatenate([yl[::-1], yu[1:]])
    sy[0] = 0.0
    sy[-1] = 0.0
    sx[0] = 1.0
    sx[-1] = 1.0
    te_ds = math.hypot(sx[-1] - sx[-2], sy[-1] - sy[-2])
    # wake stations x = 1 .. 1+far
    d_w, _ = _layer_distances(n_wake + 1, te_ds, 1.25, far)
    wx = 1.0 + d_w                      # n_wake+1 points, wx[0] = 1 (TE)
    # inner curve along i
    ix = np.concatenate([wx[::-1], sx[1:-1], wx])
    iy = np.concatenate([np.zeros(n_wake + 1), sy[1:-1], np.zeros(n_wake + 1)])
    assert ix.size == nx
    # unit outward normals
    nxv = np.zeros(nx)
    nyv = np.zeros(nx)
    nyv[: n_wake + 1] = -1.0
    nyv[nx - n_wake - 1:] = 1.0
    for k in range(n_wake + 1, nx - n_wake - 1):
        tx = ix[k + 1] - ix[k - 1]
        ty = iy[k + 1] - iy[k - 1]
        L = math.hypot(tx, ty)
        nxv[k], nyv[k] = -ty / L, tx / L
    # near the trailing edge turn the surface normals towards the wake normal
    # (straight down / up) so that grid lines leaving the trailing-edge wedge
    # do not cross; the turn fades out over ~0.05 chord of arc length
    for k in range(n_wake + 1, nx - n_wake - 1):
        lower = k < nx // 2
        s_te = abs(1.0 - ix[k]) + abs(iy[k])
        w = math.exp(-s_te / 0.05)
        ax = (1.0 - w) * nxv[k]
        ay = (1.0 - w) * nyv[k] + w * (-1.0 if lower else 1.0)
        L = math.hypot(ax, ay)
        nxv[k], nyv[k] = ax / L, ay / L
    # outer C: straight lines y = -far / +far above the wake and the aft part
    # of the section (x >= xs), a semicircle of radius `far` about (xs, 0)
    # ahead of it; grid lines run from each inner point to its outer point
    xs_ = 0.5
    ox = np.empty(nx)
    oy = np.empty(nx)
    ox[: n_wake + 1] = wx[::-1]
    oy[: n_wake + 1] = -far
    ox[nx - n_wake - 1:] = wx
    oy[nx - n_wake - 1:] = far
    af = np.arange(n_wake + 1, nx - n_wake - 1)
    aft = ix[af] >= xs_
    ox[af[aft]] = ix[af[aft]]
    oy[af[aft]] = np.where(af[aft] < nx // 2, -far, far)
    fwd = af[~aft]
    k0, k1 = fwd[0] - 1, fwd[-1] + 1
    seg = np.arange(k0, k1 + 1)
    arc = np.concatenate([[0.0], np.cumsum(np.hypot(np.diff(ix[seg]), np.diff(iy[seg])))])
    sfr = arc / arc[-1]
    # end points of the forward segment map to the ends of the semicircle
    th = -0.5 * math.pi - math.pi * sfr
    ox[seg[1:-1]] = xs_ + far * np.cos(th[1:-1])
    oy[seg[1:-1]] = far * np.sin(th[1:-1])
    ox[k0], oy[k0] = ix[k0], -far
    ox[k1], oy[k1] = ix[k1], far
    d, dinfo = _layer_distances(ny, ds1, ratio, far)
    D = np.repeat(d[:, None], nx, axis=1)   # layer distances per column
    if wake_ds_growth > 0.0:
        for k in list(range(n_wake)) + list(range(nx - n_wake, nx)):
            ds_k = min(ds1 + wake_ds_growth * (ix[k] - 1.0), 0.5 * far / (ny - 1))
            D[:, k] = _layer_distances(ny, ds_k, ratio, far)[0]
    F_ = D / far
    BL = np.minimum(1.0, D / 0.25) ** 0.5   # 0 at the wall (surface-normal lines), 1 beyond 0.25 chord
    x = ix + (1.0 - BL) * D * nxv + BL * (ox - ix) * F_
    y = iy + (1.0 - BL) * D * nyv + BL * (oy - iy) * F_
    # C-grid wake cut is a 1:1 self-interface on face 2 (reversed index order)
    iw1 = n_wake + 1                    # last node of the lower wake (TE)
    iw2 = nx - n_wake                   # first node of the upper wake (TE)
    bc = [
        [1, 1, 1, 1, 1, ny, F.BC_FARFIELD, 0, 0],
        [2, 2, 1, iw1, 1, 1, 1, 4, 1],
        [3, 2, iw1, iw2, 1, 1, F.BC_WALL, 0, 0],
        [4, 2, iw2, nx, 1, 1, 1, 2, 1],
        [5, 3, nx, nx, 1, ny, F.BC_FARFIELD, 0, 0],
        [6, 4, 1, nx, ny, ny, F.BC_FARFIELD, 0, 0],
    ]
    _check_right_handed(x, y)
    info = {"n_wake": n_wake, "n_airfoil": n_af, "first_spacing": d[1], **dinfo}
    return x, y, bc, info


def _check_right_handed(x, y):
    # cross product of the cell diagonals P(i+1,j+1)-P(i,j) and P(i,j+1)-P(i+1,j)
    d1x, d1y = x[1:, 1:] - x[:-1, :-1], y[1:, 1:] - y[:-1, :-1]
    d2x, d2y = x[1:, :-1] - x[:-1, 1:], y[1:, :-1] - y[:-1, 1:]
    a = d1x * d2y - d1y * d2x
    if not np.all(a > 0):
        bad = np.argwhere(a <= 0)
        raise ValueError(f"grid not right-handed at {len(bad)} cells, first (j,i)={tuple(bad[0])}")


# --------------------------------------------------------------------------
# multi-block Cartesian box with an isentropic vortex
# --------------------------------------------------------------------------
def vortex_multiblock(seed: int = 0, nbx: int = 8, nby: int = 8, wmin: int = 33, wmax: int = 257,
                      hmin: int = 33, hmax: int = 129, h: float = 1.0 / 64.0, shear: float = 0.2,
                      amp: float = 0.05, Ma: float = 0.5, gamma: float = 1.4, beta: float = 5.0,
                      radius: float = 1.0, noise: float = 0.01):
    """Blocks of a sheared, sinusoidally perturbed Cartesian grid with 1:1
    point-matched interfaces, plus the initial (d,u,v,T) field on each block's
    cells and ghost ring (j=0..ny, i=0..nx).

    Column widths (points) are drawn from [wmin, wmax], row heights from
    [hmin, hmax] with numpy.random.default_rng(seed); node spacing is h before
    the shear x += shear*y and the perturbation of amplitude amp*h.
    """
    rng = np.random.default_rng(seed)
    W = rng.integers(wmin, wmax + 1, size=nbx)
    H = rng.integers(hmin, hmax + 1, size=nby)
    NX = int(np.sum(W - 1)) + 1
    NY = int(np.sum(H - 1)) + 1
    gi, gj = np.meshgrid(np.arange(NX), np.arange(NY))
    X0 = gi * h
    Y0 = gj * h
    Lx, Ly = X0.max(), Y0.max()
    X = X0 + shear * Y0 + amp * h * np.sin(2 * math.pi * Y0 / (Ly / 3.0)) * np.sin(2 * math.pi * X0 / (Lx / 5.0))
    Y = Y0 + amp * h * np.sin(2 * math.pi * X0 / (Lx / 4.0)) * np.cos(2 * math.pi * Y0 / (Ly / 2.0))
    xs, ys, bcs, dims = [], [], [], []
    i0s = np.concatenate([[0], np.cumsum(W - 1)])
    j0s = np.concatenate([[0], np.cumsum(H - 1)])
    for r in range(nby):
        for c in range(nbx):
            nx, ny = int(W[c]), int(H[r])
            i0, j0 = int(i0s[c]), int(j0s[r])
            xs.append(X[j0:j0 + ny, i0:i0 + nx].copy())
            ys.append(Y[j0:j0 + ny, i0:i0 + nx].copy())
            dims.append((nx, ny))

    def bid(cc, rr):
        return rr * nbx + cc + 1

    for r in range(nby):
        for c in range(nbx):
            nx, ny = int(W[c]), int(H[r])
            subs = []
            subs.append([1, 1, 1, 1, 1, ny] + ([bid(c - 1, r), 3, 2] if c > 0 else [F.BC_FARFIELD, 0, 0]))
            subs.append([2, 2, 1, nx, 1, 1] + ([bid(c, r - 1), 4, 2] if r > 0 else [F.BC_FARFIELD, 0, 0]))
            subs.append([3, 3, nx, nx, 1, ny] + ([bid(c + 1, r), 1, 2] if c < nbx - 1 else [F.BC_FARFIELD, 0, 0]))
            subs.append([4, 4, 1, nx, ny, ny] + ([bid(c, r + 1), 2, 2] if r < nby - 1 else [F.BC_FARFIELD, 0, 0]))
            bcs.append(subs)

    xc, yc = 0.5 * (X.min() + X.max()), 0.5 * (Y.min() + Y.max())
    modes = [(rng.uniform(0.5, 3.0), rng.uniform(0.5, 3.0), rng.uniform(0, 2 * math.pi)) for _ in range(4)]

    def field(x, y):
        cinf = 1.0 / Ma
        rx, ry = (x - xc) / radius, (y - yc) / radius
        r2 = rx * rx + ry * ry
        e = np.exp(0.5 * (1.0 - r2))
        du = -beta / (2 * math.pi) * cinf * ry * e
        dv = beta / (2 * math.pi) * cinf * rx * e
        # temperature ratio of the isentropic vortex (speed-of-sound units)
        Tr = 1.0 - (gamma - 1.0) * beta ** 2 / (8.0 * gamma * math.pi ** 2) * e * e
        d = Tr ** (1.0 / (gamma - 1.0))
        p = d * Tr / (gamma * Ma * Ma)
        pert_d = sum(np.sin(2 * math.pi * (a * x / Lx + b * y / Ly) + ph) for a, b, ph in modes[:2]) / 2
        pert_p = sum(np.cos(2 * math.pi * (a * x / Lx + b * y / Ly) + ph) for a, b, ph in modes[2:]) / 2
        d = d * (1.0 + noise * pert_d)
        p = p * (1.0 + noise * pert_p)
        T = p * gamma * Ma * Ma / d
        return np.stack([d, 1.0 + du, dv, T], axis=-1)

    return xs, ys, bcs, field, {"W": W.tolist(), "H": H.tolist(), "cells": int(np.sum(W - 1) * np.sum(H - 1))}


# --------------------------------------------------------------------------
# case directories
# --------------------------------------------------------------------------
def write_case(dirpath: str, control: F.Control, xs, ys, bcs, fdm_blocks=(), flow=None) -> str:
    os.makedirs(dirpath, exist_ok=True)
    F.write_control(os.path.join(dirpath, "control.in"), control)
    F.write_mesh2d(os.path.join(dirpath, "Mesh2d.dat"), xs, ys)
    F.write_bcin(os.path.join(dirpath, "bc2d_1.in"), bcs)
    fdm = os.path.join(dirpath, "FDM.in")
    if fdm_blocks:
        F.write_fdm_in(fdm, list(fdm_blocks))
    elif os.path.exists(fdm):
        os.remove(fdm)
    for stale in ("dist_wall.dat",):
        p = os.path.join(dirpath, stale)
        if os.path.exists(p):
            os.remove(p)
    if flow is not None:
        x1s, y1s, prims = flow
        F.write_flow2d(os.path.join(dirpath, "flow2d.dat"), x1s, y1s, prims)
    return dirpath


def control_euler_naca(iters: int = 2000, Ma: float = 0.5, aoa: float = 1.25, cfl: float = 5.0) -> F.Control:
    """Inviscid NACA 0012 at Mach 0.5, AoA 1.25 deg, Roe + MUSCL3 (van Albada), LU-SGS CFL 5."""
    c = F.Control(Ma=Ma, Re=1.0e6, AoA=aoa, Pr=0.72, If_viscous=0, Iflag_turbulence_model=F.TURB_NONE,
                  Iflag_local_dt=1, dt_global=2.0 ** -10, CFL=cfl, dtmax=1.0, dtmin=1e-10,
                  Time_Method=F.TIME_LU_SGS, Iflag_Scheme=F.SCHEME_MUSCL3, Iflag_Flux=F.FLUX_ROE,
                  IFlag_Reconstruction=F.RECON_ORIGINAL, Num_Mesh=1)
    c.t_end = iters * c.dt_global
    return c


def control_sa_naca(iters: int = 100, Ma: float = 0.73, aoa: float = 2.0, Re: float = 6.5e6, cfl: float = 5.0,
                    num_mesh: int = 1, time_method: int = F.TIME_LU_SGS, scheme: int = F.SCHEME_MUSCL3,
                    flux: int = F.FLUX_ROE) -> F.Control:
    c = F.Control(Ma=Ma, Re=Re, AoA=aoa, Pr=0.72, If_viscous=1, Iflag_turbulence_model=F.TURB_SA,
                  Iflag_local_dt=1, dt_global=2.0 ** -10, CFL=cfl, dtmax=1.0, dtmin=1e-10,
                  Time_Method=time_method, Iflag_Scheme=scheme, Iflag_Flux=flux,
                  IFlag_Reconstruction=F.RECON_ORIGINAL, Num_Mesh=num_mesh, T_inf=288.15, vt_inf=3.0)
    c.t_end = iters * c.dt_global
    return c


def control_vortex(iters: int = 100, Ma: float = 0.5, cfl: float = 5.0) -> F.Control:
    c = F.Control(Ma=Ma, Re=1.0e6, AoA=0.0, If_viscous=0, Iflag_turbulence_model=F.TURB_NONE, Iflag_init=1,
                  Iflag_local_dt=1, dt_global=2.0 ** -10, CFL=cfl, dtmax=1.0, dtmin=1e-10,
                  Time_Method=F.TIME_LU_SGS, Iflag_Scheme=F.SCHEME_MUSCL3, Iflag_Flux=F.FLUX_AUSM,
                  IFlag_Reconstruction=F.RECON_ORIGINAL, Num_Mesh=1)
    c.t_end = iters * c.dt_global
    return c


def ensemble_params(seed: int = 0, n: int = 256):
    """Seeded NACA mpxx SA-RANS ensemble: thickness 8-16%, camber 0-4% at
    30-50% chord, Mach 0.3-0.8, AoA -2..6 deg, Re log-uniform 1e6..1e7, grid
    from 129x49 to 257x65 (nx in steps of 8 so subfaces stay aligned)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        nx = int(rng.integers(0, 17)) * 8 + 129
        ny = int(rng.integers(0, 5)) * 4 + 49
        out.append(dict(t=rng.uniform(0.08, 0.16), m=rng.uniform(0.0, 0.04), p=rng.uniform(0.3, 0.5),
                        Ma=rng.uniform(0.3, 0.8), aoa=rng.uniform(-2.0, 6.0),
                        Re=10 ** rng.uniform(6.0, 7.0), nx=nx, ny=ny))
    return out


def write_vortex_case(dirpath: str, control: F.Control, seed: int = 0, nbx: int = 8, nby: int = 8, wmin: int = 33,
                      wmax: int = 257, hmin: int = 33, hmax: int = 129, h: float = 1.0 / 64.0, bc_edit=None) -> dict:
    """Multi-block vortex case (vortex_multiblock) with its initial field as
    flow2d.dat (the control must have Iflag_init = 1); returns the generator's
    info (column widths, row heights, cells)."""
    from . import geometry as G

    xs, ys, bcs, field, info = vortex_multiblock(seed=seed, nbx=nbx, nby=nby, wmin=wmin, wmax=wmax, hmin=hmin,
                                                 hmax=hmax, h=h)
    if bc_edit:
        bc_edit(bcs)
    lev = G.build_levels(xs, ys, bcs, 1)
    cx = [b.x1[1:-1, 1:-1] for b in lev[0]]
    cy = [b.y1[1:-1, 1:-1] for b in lev[0]]
    write_case(dirpath, control, xs, ys, bcs, flow=(cx, cy, [field(x, y) for x, y in zip(cx, cy)]))
    return info


def write_ensemble_case(dirpath: str, prm: dict, iters: int = 1000, cfl: float = 5.0, ds1: float = 5e-6,
                        wake_ds_growth: float = 1e-3) -> str:
    """One SA-RANS airfoil case of the ensemble (a dict of ensemble_params):
    NACA mpxx C-grid of nx x ny points, Roe + MUSCL3, LU-SGS on one level.
    The wake's first spacing grows by wake_ds_growth per chord behind the
    trailing edge: with the surface's 5e-6 chord in the whole wake, the
    cells beside the wake cut reach aspect ratios near 1e5, and the residual
    of those cells, which LU-SGS couples across the cut only through the
    ghost cells, then decays about ten times more slowly."""
    x, y, bc, _ = naca_cgrid(prm["nx"], prm["ny"], m=prm["m"], p=prm["p"], t=prm["t"], ds1=ds1,
                             wake_ds_growth=wake_ds_growth)
    ctl = control_sa_naca(iters=iters, Ma=prm["Ma"], aoa=prm["aoa"], Re=prm["Re"], cfl=cfl)
    return write_case(dirpath, ctl, [x], [y], [bc])

