"""Mesh-level invariants of the oracle (CPU), fixed by the mathematics rather
than by the reference's outputs:

* freestream preservation: a uniform state on a closed, sheared and perturbed
  multi-block mesh with far-field boundaries gives Res = 0 to round-off, for
  every scheme / flux / reconstruction (the face normals close each cell);
* LU-SGS with Res = 0 returns dU = 0 exactly;
* multigrid transfers (sub_NS_multigrid.f90:271-398): the residual
  restriction conserves the sum of Res; on an affine mesh the volume-weighted
  restriction and the 9/16-3/16-3/16-1/16 prolongation reproduce linear
  fields exactly.
"""
import itertools
import math

import numpy as np
import pytest

import oracle as O
from ec2d import cases
from ec2d import formats as F
from ec2d import geometry as G


def vortex_geometry():
    xs, ys, bcs, _, _ = cases.vortex_multiblock(seed=4, nbx=3, nby=2, wmin=9, wmax=21, hmin=9, hmax=17, h=0.25)
    return xs, ys, bcs


def uniform_state(oc, ctl):
    d0, u0, v0 = 1.0, math.cos(ctl.aoa_rad), math.sin(ctl.aoa_rad)
    p0 = 1.0 / (ctl.gamma * ctl.Ma * ctl.Ma)
    for b in oc.meshes[0]["blocks"]:
        b.U[..., 0] = d0
        b.U[..., 1] = d0 * u0
        b.U[..., 2] = d0 * v0
        b.U[..., 3] = p0 / (ctl.gamma - 1.0) + 0.5 * d0 * (u0 * u0 + v0 * v0)
    return p0


@pytest.mark.parametrize("scheme,flux,recon", [(s, f, r) for s, f, r in itertools.product(
    (0, 1, 2, 3, 4, 5, 6), (4,), (0,))] + [(5, f, 0) for f in (1, 2, 3, 5, 6)] + [(5, 4, 1), (5, 4, 2), (1, 6, 2)])
def test_freestream_preservation(scheme, flux, recon):
    xs, ys, bcs = vortex_geometry()
    c = cases.control_vortex(iters=1)
    c.AoA, c.Iflag_init = 3.0, 0
    c.Iflag_Scheme, c.Iflag_Flux, c.IFlag_Reconstruction = scheme, flux, recon
    oc = O.OracleCase(control=c, levels=G.build_levels(xs, ys, bcs, 1))
    p0 = uniform_state(oc, c)
    oc.bc(1)
    oc.exchange(1)
    oc.residual(1)
    for b in oc.meshes[0]["blocks"]:
        scale = p0 * b.g.si.max()
        assert np.abs(b.Res).max() < 1e-13 * scale, (scheme, flux, recon, np.abs(b.Res).max())


def test_lusgs_zero_residual_gives_zero_update():
    x, y, bc, _ = cases.naca_cgrid(65, 17, ds1=1e-4)
    c = cases.control_sa_naca(iters=3, Re=1e5, Ma=0.5)
    oc = O.OracleCase(control=c, levels=G.build_levels([x], [y], [bc], 1, need_dw=True))
    oc.init_flow()
    oc.main_step()
    oc.residual(1)
    b = oc.block()
    b.Res[...] = 0.0
    b.dU[...] = 7.0
    oc.lusgs(1, 1, 0.0)
    assert np.all(b.dU[:-1, :-1] == 0.0)


def affine_block(nx=17, ny=9, h=0.25, shear=0.3):
    j, i = np.mgrid[0:ny, 0:nx].astype(float)
    x = h * i + shear * h * j
    y = h * j
    bc = [[1, 1, 1, 1, 1, ny, F.BC_FARFIELD, 0, 0], [2, 2, 1, nx, 1, 1, F.BC_FARFIELD, 0, 0],
          [3, 3, nx, nx, 1, ny, F.BC_FARFIELD, 0, 0], [4, 4, 1, nx, ny, ny, F.BC_FARFIELD, 0, 0]]
    return x, y, bc


def two_level_case():
    x, y, bc = affine_block()
    c = cases.control_euler_naca(iters=1)
    c.Time_Method, c.Num_Mesh = F.TIME_RK3, 2
    return O.OracleCase(control=c, levels=G.build_levels([x], [y], [bc], 2))


def test_residual_restriction_conserves_sum():
    oc = two_level_case()
    rng = np.random.default_rng(1)
    fine = oc.block(1, 1)
    fine.Res[...] = rng.normal(size=fine.Res.shape)
    coarse = oc.block(2, 1)
    coarse.QF[...] = 0.0
    O.lib().o_interp2h(oc.M(1), oc.M(2), 2)
    np.testing.assert_allclose(coarse.QF.sum(axis=(0, 1)), fine.Res[..., :4].sum(axis=(0, 1)), rtol=1e-12)


def test_state_restriction_of_linear_field():
    oc = two_level_case()
    fine, coarse = oc.block(1, 1), oc.block(2, 1)
    # cell-centre coordinates (x1, y1 carry LAP ghost layers; U has LAP too)
    for m, (a, bx, by) in enumerate([(1.0, 0.1, -0.2), (0.5, 0.0, 0.3), (-0.25, 0.2, 0.0), (2.0, 0.05, 0.05)]):
        fine.U[..., m] = a + bx * fine.g.x1 + by * fine.g.y1  # both span Fortran -1..n+1
    O.lib().o_interp2h(oc.M(1), oc.M(2), 1)
    for m, (a, bx, by) in enumerate([(1.0, 0.1, -0.2), (0.5, 0.0, 0.3), (-0.25, 0.2, 0.0), (2.0, 0.05, 0.05)]):
        want = a + bx * coarse.g.x1 + by * coarse.g.y1
        # interior cells, Fortran 1..n-1 <-> numpy 2..n
        np.testing.assert_allclose(coarse.U[2:-2, 2:-2, m], want[2:-2, 2:-2], rtol=1e-13, atol=1e-14)


def test_prolongation_of_linear_correction():
    oc = two_level_case()
    fine, coarse = oc.block(1, 1), oc.block(2, 1)
    # deltU bounds (-1:nx+1, -1:ny+1): numpy index k <-> Fortran index k - 1
    J, I = np.mgrid[0:coarse.deltU.shape[0], 0:coarse.deltU.shape[1]] - 1
    for m, (a, b, c) in enumerate([(0.75, 0.125, -0.25), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), (-1.5, 0.5, 0.5)]):
        coarse.deltU[..., m] = a + b * I + c * J
    O.lib().o_prolong(oc.M(2), oc.M(1), 2)
    Jf, If = np.mgrid[0:fine.deltU.shape[0], 0:fine.deltU.shape[1]] - 1
    sel = (slice(1, fine.ny + 2), slice(1, fine.nx + 2))  # Fortran 0..nx, 0..ny are written
    for m, (a, b, c) in enumerate([(0.75, 0.125, -0.25), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), (-1.5, 0.5, 0.5)]):
        want = a + b * (If + 0.5) / 2.0 + c * (Jf + 0.5) / 2.0
        np.testing.assert_allclose(fine.deltU[sel + (m,)], want[sel], rtol=1e-14, atol=1e-14)
