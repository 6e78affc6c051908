"""Host-side set-up of one OpenCFD-EC2D case: mesh levels, geometry, ghost
coordinates, wall distance and finite-difference metrics.

This is the set-up the reference performs once before its main loop.  It
feeds the device, which can also make all of it but the finite-difference
metrics itself from the nodes (node_levels; ec2d_commit then runs the ghost
nodes, centres, metrics and wall distance on the GPU).
Every quantity is computed with the same floating-point operations, in the
same order, as the reference, so that the arrays handed to the device are
bit-identical to the reference's own:

* volumes, face lengths and normals: sub_init.f90:180-204 (fine mesh) and
  sub_init.f90:321-353 (coarse meshes made by point skipping, :321-327);
* ghost node coordinates by 1:1 exchange or linear extrapolation, ghost corner
  nodes and cell centres: sub_boundary.f90:174-281;
* wall distance: sub_turbulence_SST.f90:286-337;
* finite-difference metrics: sub_Finite_Difference.f90:314-439.

Array layout: every 2-D array is a C-ordered numpy array whose first index is
j and second is i, i.e. the memory order of the reference's column-major
Fortran arrays.  ``lo`` gives the Fortran lower bound of both indices.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from .formats import BC_WALL, LAP


@dataclass
class BlockGeom:
    nx: int
    ny: int
    bc: List[List[int]]
    fvm_fdm: int = 0
    x: np.ndarray = None      # nodes, bounds (1-LAP:nx+LAP, 1-LAP:ny+LAP)
    y: np.ndarray = None
    x1: np.ndarray = None     # cell centres, bounds (1-LAP:nx+LAP-1, 1-LAP:ny+LAP-1)
    y1: np.ndarray = None
    vol: np.ndarray = None    # bounds (1:nx, 1:ny)
    si: np.ndarray = None
    sj: np.ndarray = None
    ni1: np.ndarray = None
    ni2: np.ndarray = None
    nj1: np.ndarray = None
    nj2: np.ndarray = None
    dw: Optional[np.ndarray] = None   # bounds (0:nx, 0:ny)
    fd: Dict[str, np.ndarray] = field(default_factory=dict)  # Akx..Ajac, bounds (1:nx-1, 1:ny-1)

    # Fortran-index helpers -------------------------------------------------
    def node(self, arr, i, j):
        return arr[j + LAP - 1, i + LAP - 1]

    @property
    def ncell(self) -> int:
        return (self.nx - 1) * (self.ny - 1)


def _fine_geometry(B: BlockGeom) -> None:
    """vol, si, sj and face normals from the nodes (sub_init.f90:180-204)."""
    nx, ny = B.nx, B.ny
    o = LAP - 1  # offset of Fortran index 1 in the node arrays
    X = B.x
    Y = B.y

    def xs(i0, i1, j0, j1, A):  # Fortran slice A(i0:i1, j0:j1) -> numpy [j, i]
        return A[j0 + o:j1 + o + 1, i0 + o:i1 + o + 1]

    vol = np.zeros((ny, nx))
    a = xs(1, nx - 1, 1, ny - 1, X) - xs(2, nx, 2, ny, X)
    b = xs(2, nx, 1, ny - 1, Y) - xs(1, nx - 1, 2, ny, Y)
    c = xs(2, nx, 1, ny - 1, X) - xs(1, nx - 1, 2, ny, X)
    d = xs(1, nx - 1, 1, ny - 1, Y) - xs(2, nx, 2, ny, Y)
    vol[: ny - 1, : nx - 1] = np.abs(a * b - c * d) * 0.5

    si = np.zeros((ny, nx))
    ni1 = np.zeros((ny, nx))
    ni2 = np.zeros((ny, nx))
    dx = xs(1, nx, 2, ny, X) - xs(1, nx, 1, ny - 1, X)
    dy = xs(1, nx, 2, ny, Y) - xs(1, nx, 1, ny - 1, Y)
    s = np.sqrt(dx * dx + dy * dy)
    si[: ny - 1, :] = s
    ni1[: ny - 1, :] = dy / s
    ni2[: ny - 1, :] = -dx / s

    sj = np.zeros((ny, nx))
    nj1 = np.zeros((ny, nx))
    nj2 = np.zeros((ny, nx))
    dx = xs(2, nx, 1, ny, X) - xs(1, nx - 1, 1, ny, X)
    dy = xs(2, nx, 1, ny, Y) - xs(1, nx - 1, 1, ny, Y)
    s = np.sqrt(dx * dx + dy * dy)
    sj[:, : nx - 1] = s
    nj1[:, : nx - 1] = -dy / s
    nj2[:, : nx - 1] = dx / s
    B.vol, B.si, B.sj, B.ni1, B.ni2, B.nj1, B.nj2 = vol, si, sj, ni1, ni2, nj1, nj2


def _new_block(nx, ny, bc, xn, yn) -> BlockGeom:
    B = BlockGeom(nx=nx, ny=ny, bc=[list(s) for s in bc])
    B.x = np.full((ny + 2 * LAP, nx + 2 * LAP), np.nan)
    B.y = np.full((ny + 2 * LAP, nx + 2 * LAP), np.nan)
    B.x[LAP:LAP + ny, LAP:LAP + nx] = xn
    B.y[LAP:LAP + ny, LAP:LAP + nx] = yn
    _fine_geometry(B)
    return B


def coarsen(fine: List[BlockGeom]) -> List[BlockGeom]:
    """Next-coarser mesh by taking every second node (sub_init.f90:276-373)."""
    out = []
    for B1 in fine:
        nx = (B1.nx - 1) // 2 + 1
        ny = (B1.ny - 1) // 2 + 1
        xn = B1.x[LAP:LAP + B1.ny:2, LAP:LAP + B1.nx:2]
        yn = B1.y[LAP:LAP + B1.ny:2, LAP:LAP + B1.nx:2]
        bc = []
        for s in B1.bc:
            f_no, face, ist, iend, jst, jend, nb, sub, ori = s
            bc.append([f_no, face, (ist - 1) // 2 + 1, (iend - 1) // 2 + 1, (jst - 1) // 2 + 1,
                       (jend - 1) // 2 + 1, nb, sub, ori])
        out.append(_new_block(nx, ny, bc, xn.copy(), yn.copy()))
    return out


def _fill_ghost_coords(blocks: List[BlockGeom]) -> None:
    """Ghost nodes and cell centres (sub_boundary.f90:188-281)."""
    o = LAP - 1

    for B in blocks:
        for Bc in B.bc:
            _, face, ib2, ie2, jb2, je2, nb, msub, orient = Bc
            for A_name in ("x", "y"):
                A = getattr(B, A_name)
                if nb > 0:
                    B1 = blocks[nb - 1]
                    A1 = getattr(B1, A_name)
                    _, face1, ib1, ie1, jb1, je1, _, _, _ = B1.bc[msub - 1]
                    n1 = ie2 - ib2 + je2 - jb2 + 1
                    Ux = np.empty((LAP, n1))
                    for i in range(1, LAP + 1):
                        if face1 == 1:
                            Ux[i - 1] = A1[jb1 + o:je1 + o + 1, ib1 + i + o]
                        elif face1 == 2:
                            Ux[i - 1] = A1[jb1 + i + o, ib1 + o:ie1 + o + 1]
                        elif face1 == 3:
                            Ux[i - 1] = A1[jb1 + o:je1 + o + 1, ie1 - LAP - 1 + i + o]
                        else:
                            Ux[i - 1] = A1[je1 - LAP - 1 + i + o, ib1 + o:ie1 + o + 1]
                    if orient != 2:
                        Ux = Ux[:, ::-1]
                    for i in range(1, LAP + 1):
                        if face == 1:
                            A[jb2 + o:jb2 + o + n1, ib2 - LAP - 1 + i + o] = Ux[i - 1]
                        elif face == 2:
                            A[jb2 - LAP - 1 + i + o, ib2 + o:ib2 + o + n1] = Ux[i - 1]
                        elif face == 3:
                            A[jb2 + o:jb2 + o + n1, ie2 + i + o] = Ux[i - 1]
                        else:
                            A[je2 + i + o, ib2 + o:ib2 + o + n1] = Ux[i - 1]
                else:
                    for i in range(1, LAP + 1):
                        if face == 1:
                            A[jb2 + o:je2 + o + 1, ib2 - i + o] = (
                                2.0 * A[jb2 + o:je2 + o + 1, ib2 + o] - A[jb2 + o:je2 + o + 1, ib2 + i + o])
                        elif face == 2:
                            A[jb2 - i + o, ib2 + o:ie2 + o + 1] = (
                                2.0 * A[jb2 + o, ib2 + o:ie2 + o + 1] - A[jb2 + i + o, ib2 + o:ie2 + o + 1])
                        elif face == 3:
                            A[jb2 + o:je2 + o + 1, ie2 + i + o] = (
                                2.0 * A[jb2 + o:je2 + o + 1, ie2 + o] - A[jb2 + o:je2 + o + 1, ie2 - i + o])
                        else:
                            A[je2 + i + o, ib2 + o:ie2 + o + 1] = (
                                2.0 * A[je2 + o, ib2 + o:ie2 + o + 1] - A[je2 - i + o, ib2 + o:ie2 + o + 1])
        nx, ny = B.nx, B.ny
        for A in (B.x, B.y):
            def g(i, j):
                return A[j + o, i + o]

            for j in range(1, LAP + 1):
                for i in range(1, LAP + 1):
                    i1, j1, i2, j2 = 1 - i, 1 - j, nx + i, ny + j
                    A[j1 + o, i1 + o] = g(i1, 1) + g(1, j1) - g(1, 1)
                    A[j2 + o, i2 + o] = g(i2, ny) + g(nx, j2) - g(nx, ny)
                    A[j2 + o, i1 + o] = g(i1, ny) + g(1, j2) - g(1, ny)
                    A[j1 + o, i2 + o] = g(i2, 1) + g(nx, j1) - g(nx, 1)
        # cell centres over 1-LAP .. n+LAP-1 (the products by the default-real
        # literal 0.25 are exact)
        for name, A in (("x1", B.x), ("y1", B.y)):
            c = (A[:-1, :-1] + A[:-1, 1:] + A[1:, :-1] + A[1:, 1:]) * 0.25
            setattr(B, name, c)


def wall_distance(blocks: List[BlockGeom]) -> None:
    """Distance of cell centres to the nearest wall node (sub_turbulence_SST.f90:286-337)."""
    o = LAP - 1
    wx, wy = [], []
    for B1 in blocks:
        for s in B1.bc:
            if s[6] != BC_WALL:
                continue
            _, _, ist, iend, jst, jend = s[:6]
            for j1 in range(jst, jend + 1):
                for i1 in range(ist, iend + 1):
                    wx.append(B1.x[j1 + o, i1 + o])
                    wy.append(B1.y[j1 + o, i1 + o])
    wx = np.array(wx)
    wy = np.array(wy)
    for B in blocks:
        nx, ny = B.nx, B.ny
        xn = B.x[1 + o:ny + o + 1, 1 + o:nx + o + 1]
        yn = B.y[1 + o:ny + o + 1, 1 + o:nx + o + 1]
        dwn = np.full((ny, nx), 1.0e10)
        for k in range(wx.size):  # same sequence of strict-less updates as the reference
            ddx = xn - wx[k]
            ddy = yn - wy[k]
            d1 = np.sqrt(ddx * ddx + ddy * ddy)
            np.copyto(dwn, d1, where=d1 < dwn)
        dw = np.ones((ny + 1, nx + 1))
        dw[1:ny, 1:nx] = (dwn[:-1, :-1] + dwn[:-1, 1:] + dwn[1:, :-1] + dwn[1:, 1:]) * 0.25
        B.dw = dw


def _dx0(f: np.ndarray) -> np.ndarray:
    """6th-order i-derivative with one-sided closures (sub_Finite_Difference.f90:387-412), f[j, i]."""
    hx = 1.0
    b1 = 8.0 / (12.0 * hx)
    b2 = 1.0 / (12.0 * hx)
    a1 = 1.0 / (60.0 * hx)
    a2 = -3.0 / (20.0 * hx)
    a3 = 3.0 / (4.0 * hx)
    nx = f.shape[1]
    F = lambda i: f[:, i - 1]  # noqa: E731  (Fortran column i)
    fx = np.zeros_like(f)
    for i in range(4, nx - 2):
        fx[:, i - 1] = a1 * (F(i + 3) - F(i - 3)) + a2 * (F(i + 2) - F(i - 2)) + a3 * (F(i + 1) - F(i - 1))
    fx[:, 0] = (-3.0 * F(1) + 4.0 * F(2) - F(3)) / (2.0 * hx)
    fx[:, 1] = (-2.0 * F(1) - 3.0 * F(2) + 6.0 * F(3) - F(4)) / (6.0 * hx)
    fx[:, 2] = b1 * (F(4) - F(2)) - b2 * (F(5) - F(1))
    fx[:, nx - 3] = b1 * (F(nx - 1) - F(nx - 3)) - b2 * (F(nx) - F(nx - 4))
    fx[:, nx - 2] = (F(nx - 3) - 6.0 * F(nx - 2) + 3.0 * F(nx - 1) + 2.0 * F(nx)) / (6.0 * hx)
    fx[:, nx - 1] = (F(nx - 2) - 4.0 * F(nx - 1) + 3.0 * F(nx)) / (2.0 * hx)
    return fx


def fd_metrics(B: BlockGeom) -> None:
    """Finite-difference metrics on cell centres (sub_Finite_Difference.f90:314-362)."""
    nx, ny = B.nx - 1, B.ny - 1
    o = LAP - 1
    xx = B.x1[1 + o:ny + o + 1, 1 + o:nx + o + 1].copy()
    yy = B.y1[1 + o:ny + o + 1, 1 + o:nx + o + 1].copy()
    xk = _dx0(xx)
    yk = _dx0(yy)
    xi = _dx0(xx.T.copy()).T   # dy0 is dx0 along j (identical arithmetic)
    yi = _dx0(yy.T.copy()).T
    ajac = xk * yi - xi * yk
    ajac = 1.0 / ajac
    B.fd = {
        "Akx": ajac * yi,
        "Aky": -ajac * xi,
        "Aix": -ajac * yk,
        "Aiy": ajac * xk,
        "Ajac": ajac,
    }


def mesh_depth(blocks_dims, bcs) -> int:
    """Number of multigrid levels the mesh allows, capped at 3 (sub_init.f90:39-103)."""
    km_grid = blocks_dims[0][0]
    for ni, nj in blocks_dims:
        km, nn = 1, 2
        while (ni - 1) % nn == 0 and (ni - 1) // nn >= 2 and (nj - 1) % nn == 0 and (nj - 1) // nn >= 2:
            km += 1
            nn *= 2
        km_grid = min(km_grid, km)
    for subs in bcs:
        for s in subs:
            ist, iend, jst, jend = s[2:6]
            nn, km = 1, 1
            while (ist - 1) % nn == 0 and (iend - 1) % nn == 0 and (jst - 1) % nn == 0 and (jend - 1) % nn == 0:
                nn *= 2
                km += 1
                if nn > 1 << 20:
                    break
            km_grid = min(km_grid, km)
    return km_grid


def node_levels(xs, ys, bcs, num_mesh: int) -> List[List[BlockGeom]]:
    """All mesh levels of one case with the interior nodes only (every second
    node for each coarser level, sub_init.f90:276-373): ghost nodes, cell
    centres, metrics and wall distance are left to the device (ec2d_commit)."""
    depth = mesh_depth([(x.shape[1], x.shape[0]) for x in xs], bcs)
    num_mesh = min(3, depth, num_mesh)
    levels = []
    for lev in range(num_mesh):
        blocks = []
        for k, (x, y, bc) in enumerate(zip(xs, ys, bcs)):
            if lev == 0:
                nx, ny, b, xn, yn = x.shape[1], x.shape[0], [list(s) for s in bc], x, y
            else:
                F_ = levels[-1][k]
                nx, ny = (F_.nx - 1) // 2 + 1, (F_.ny - 1) // 2 + 1
                xn = F_.x[LAP:LAP + F_.ny:2, LAP:LAP + F_.nx:2]
                yn = F_.y[LAP:LAP + F_.ny:2, LAP:LAP + F_.nx:2]
                b = [[s[0], s[1], (s[2] - 1) // 2 + 1, (s[3] - 1) // 2 + 1, (s[4] - 1) // 2 + 1, (s[5] - 1) // 2 + 1,
                      s[6], s[7], s[8]] for s in F_.bc]
            B = BlockGeom(nx=nx, ny=ny, bc=b)
            B.x = np.full((ny + 2 * LAP, nx + 2 * LAP), np.nan)
            B.y = np.full((ny + 2 * LAP, nx + 2 * LAP), np.nan)
            B.x[LAP:LAP + ny, LAP:LAP + nx] = xn
            B.y[LAP:LAP + ny, LAP:LAP + nx] = yn
            blocks.append(B)
        levels.append(blocks)
    return levels


def build_levels(xs, ys, bcs, num_mesh: int, fdm_blocks=(), need_dw: bool = False) -> List[List[BlockGeom]]:
    """All mesh levels of one case, finest first."""
    depth = mesh_depth([(x.shape[1], x.shape[0]) for x in xs], bcs)
    num_mesh = min(3, depth, num_mesh)
    fine = [_new_block(x.shape[1], x.shape[0], bc, x, y) for x, y, bc in zip(xs, ys, bcs)]
    for m in fdm_blocks:
        if 1 <= m <= len(fine):
            fine[m - 1].fvm_fdm = 1
    levels = [fine]
    for _ in range(1, num_mesh):
        levels.append(coarsen(levels[-1]))
    for lev in levels:
        _fill_ghost_coords(lev)
    for B in fine:
        if B.fvm_fdm == 1:
            fd_metrics(B)
    if need_dw:
        wall_distance(fine)
    return levels
