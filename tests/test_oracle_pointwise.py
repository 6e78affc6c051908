"""Closed-form checks of the oracle's pointwise pieces (CPU).

These pin the C restatement independently of the reference build: each
property follows from the mathematics of the published method, so a dropped
term, a wrong sign or a transposed index fails one of them.

* interface fluxes (sub_flux_split.f90): consistency F(Q, Q) = F_exact(Q) for
  all six; exact one-sided upwinding for supersonic states (for Roe only where
  the reference's averaging keeps the Roe property); HLLC keeps a contact /
  shear wave exactly; the Roe averaging quirk itself by a hand-computed value;
* reconstruction schemes (sub_Residual.f90:591-659): constants for all,
  linear data for the 2nd/3rd-order ones, quadratic cell averages for UD3,
  hand-computed values of MUSCL2's asymmetric parameters (fP k=1, b=2;
  fm k=1/3, b=1);
* reconstruction variants (:505-586): characteristic and conservative
  reconstruction of a uniform state return that state;
* FD-WENO5 (sub_Finite_Difference.f90:207-275): quadratic cell averages give
  the exact interface value, and random data agree with an independent
  vectorised Jiang-Shu WENO5;
* Steger-Warming splitting (:154-202): F+ + F- = F(Q).A, F- = 0 when every
  eigenvalue is positive;
* comput_DFn (sub_LU_SGS.f90:111-138): the Euler normal-flux difference.
"""
import numpy as np
import pytest

import oracle as O

G = 1.4


def exact_flux(Q, g=G):
    d, u, v, p = Q
    E = p / (g - 1.0) + 0.5 * d * (u * u + v * v)
    return np.array([d * u, d * u * u + p, d * u * v, u * (E + p)])


def states(rng, n, mach_lo, mach_hi):
    out = []
    for _ in range(n):
        d = rng.uniform(0.2, 3.0)
        p = rng.uniform(0.2, 3.0)
        a = np.sqrt(G * p / d)
        M = rng.uniform(mach_lo, mach_hi)
        out.append(np.array([d, M * a, rng.uniform(-1, 1), p]))
    return out


FLUXES = {1: "steger-warming", 2: "hll", 3: "hllc", 4: "roe", 5: "van leer", 6: "ausm+"}


@pytest.mark.parametrize("iflux", sorted(FLUXES))
def test_flux_consistency(iflux):
    rng = np.random.default_rng(10 + iflux)
    for Q in states(rng, 40, -2.5, 2.5):
        F = O.pt_flux(iflux, Q, Q)
        Fe = exact_flux(Q)
        np.testing.assert_allclose(F, Fe, rtol=1e-12, atol=1e-12 * np.abs(Fe).max(), err_msg=f"{FLUXES[iflux]} {Q}")


@pytest.mark.parametrize("iflux", sorted(FLUXES))
def test_flux_supersonic_upwinding(iflux):
    # right states within a few per cent of the left ones, Mach >= 1.5, so the
    # wave-speed estimates of HLL/HLLC and Roe's eigenvalues (u - a > 0.1)
    # are all positive
    # Roe: the reference averages the density as (sqrt(dl) + sqrt(dr))^2 / 4
    # and divides the shear-wave strength by the sound speed
    # (sub_flux_split.f90:288, 316), so the Roe property, and with it exact
    # upwinding, holds only for equal densities and tangential velocities
    rng = np.random.default_rng(20 + iflux)
    L = states(rng, 20, 1.5, 3.0)
    R = [Q * rng.uniform(0.97, 1.03, 4) for Q in L]
    if iflux == 4:
        R = [np.array([QL[0], QR[1], QL[2], QR[3]]) for QL, QR in zip(L, R)]
    for QL, QR in zip(L, R):
        F = O.pt_flux(iflux, QL, QR)
        Fe = exact_flux(QL)
        np.testing.assert_allclose(F, Fe, rtol=1e-11, atol=1e-11 * np.abs(Fe).max(), err_msg=FLUXES[iflux])
    # and from the right for flow to the left
    for QL, QR in zip(L, R):
        QL = QL * np.array([1, -1, 1, 1])
        QR = QR * np.array([1, -1, 1, 1])
        F = O.pt_flux(iflux, QL, QR)
        Fe = exact_flux(QR)
        np.testing.assert_allclose(F, Fe, rtol=1e-11, atol=1e-11 * np.abs(Fe).max(), err_msg=FLUXES[iflux])


def test_hllc_contact_and_shear():
    rng = np.random.default_rng(5)
    for _ in range(20):
        p = rng.uniform(0.3, 2.0)
        QL = np.array([rng.uniform(0.2, 3), 0.0, rng.uniform(-1, 1), p])
        QR = np.array([rng.uniform(0.2, 3), 0.0, rng.uniform(-1, 1), p])
        F = O.pt_flux(3, QL, QR)
        np.testing.assert_allclose(F, [0.0, p, 0.0, 0.0], atol=1e-14)


def test_roe_density_average_quirk():
    # dl = 1, dr = 4, equal velocities and pressures: only the density wave
    # (lambda_2 = |u|, strength dr - dl = 3) is active, so
    # F_mass = (dl u + dr u)/2 - |u| * 3 / 2; with u = 0.5: 1.25 - 0.75 = 0.5,
    # independent of the density average; the momentum flux adds the same
    # wave times u-tilde, which is the arithmetic mean of u for equal u:
    # F_mom = (dl u^2 + p + dr u^2 + p)/2 - 0.5 * 0.5 * 3 * 0.5 = 0.625 + p - 0.375
    F = O.pt_flux(4, [1.0, 0.5, 0.0, 1.0], [4.0, 0.5, 0.0, 1.0])
    assert F[0] == pytest.approx(0.5, abs=1e-15)
    assert F[1] == pytest.approx(1.25, abs=1e-15)
    # the shear-wave strength carries 1/a-tilde: dl = dr = 1, p = 1, u = 0,
    # v jumps by 1: h = g/(g-1) + v^2/2 on both sides with v = 0 and 1,
    # a-tilde^2 = (g-1)(h-tilde - v-tilde^2/2) with v-tilde = 1/2,
    # h-tilde = g/(g-1) + 1/4 -> a-tilde^2 = g + (g-1)(1/4 - 1/8)
    a2 = G + (G - 1.0) * 0.125
    lam = (0.0 + 0.01) / 0.2  # Harten-fixed |u| = 0 with delta = 0.1
    F = O.pt_flux(4, [1.0, 0.0, 0.0, 1.0], [1.0, 0.0, 1.0, 1.0])
    assert F[2] == pytest.approx(-0.5 * lam * 1.0 / np.sqrt(a2), rel=1e-14)


def test_unknown_flux_is_zero():
    assert np.all(O.pt_flux(9, [1, 0.5, 0, 1], [1, 0.5, 0, 1]) == 0.0)


SCHEMES = range(7)


@pytest.mark.parametrize("flag", SCHEMES)
def test_scheme_constant(flag):
    for c in (0.0, 1.0, -3.25, 7.5e3):
        assert O.pt_scheme(1, [c] * 4, flag) == pytest.approx(c, rel=1e-15, abs=1e-15)
        assert O.pt_scheme(-1, [c] * 4, flag) == pytest.approx(c, rel=1e-15, abs=1e-15)


@pytest.mark.parametrize("flag", range(1, 7))
def test_scheme_linear(flag):
    # points k = 1..4, the face between points 2 and 3 sits at k = 2.5
    rng = np.random.default_rng(flag)
    for _ in range(20):
        a, b = rng.uniform(-2, 2, 2)
        u = a + b * np.arange(1, 5)
        exact = a + 2.5 * b
        assert O.pt_scheme(1, u, flag) == pytest.approx(exact, rel=1e-12, abs=1e-12)
        assert O.pt_scheme(-1, u, flag) == pytest.approx(exact, rel=1e-12, abs=1e-12)


def test_scheme_ud1_is_upwind_value():
    u = [1.0, 2.0, 4.0, 8.0]
    assert O.pt_scheme(1, u, 0) == 2.0 and O.pt_scheme(-1, u, 0) == 4.0


def test_scheme_ud3_quadratic_cell_averages():
    # cell averages of x^2 over unit cells centred at k: k^2 + 1/12
    rng = np.random.default_rng(3)
    for _ in range(10):
        c0, c1, c2 = rng.uniform(-1, 1, 3)
        k = np.arange(1, 5.0)
        avg = c0 + c1 * k + c2 * (k * k + 1.0 / 12.0)
        exact = c0 + c1 * 2.5 + c2 * 6.25
        assert O.pt_scheme(1, avg, 2) == pytest.approx(exact, rel=1e-13, abs=1e-13)
        assert O.pt_scheme(-1, avg, 2) == pytest.approx(exact, rel=1e-13, abs=1e-13)


def test_scheme_muscl2_asymmetric_parameters():
    # fP: k = 1, b = 2 -> u2 + 1/4 * 2 * minmod(u3-u2, 2(u2-u1)) = 1 + 1/2 * minmod(2, 2) = 2
    # fm: k = 1/3, b = 1 -> u3 - 1/4 * (2/3 minmod(1, 2) + 4/3 minmod(2, 1)) = 3 - 1/2 = 2.5
    u = [0.0, 1.0, 3.0, 4.0]
    assert O.pt_scheme(1, u, 4) == pytest.approx(2.0, abs=1e-15)
    assert O.pt_scheme(-1, u, 4) == pytest.approx(2.5, abs=1e-15)
    # minmod limits to zero at an extremum
    assert O.pt_scheme(1, [0.0, 1.0, 0.0, 1.0], 1) == 1.0


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_reconstruction_uniform_state(kind):
    d, u, v, p = 1.3, 0.4, -0.2, 0.9
    E = p / (G - 1.0) + 0.5 * d * (u * u + v * v)
    cons = np.array([d, d * u, d * v, E])
    prim = np.array([d, u, v, p])
    U0 = np.tile(prim if kind == 0 else cons, (4, 1))
    for flag in SCHEMES:
        UL, UR = O.pt_recon(kind, U0, G, flag)
        np.testing.assert_allclose(UL, prim, rtol=1e-13)
        np.testing.assert_allclose(UR, prim, rtol=1e-13)


def weno5_textbook(v, side):
    """Jiang & Shu (1996) 5th-order WENO, written from the paper: for the
    positive-going flux the face value h_{k+1/2} uses cells k-2..k+2; the
    negative-going one mirrors the stencil about the face."""
    n = len(v)
    h = np.full(n, np.nan)
    eps = 1e-6
    d = np.array([0.1, 0.6, 0.3])  # stencils {k-2..k}, {k-1..k+1}, {k..k+2}
    for k in range(2, n - 2):
        if side > 0:
            a, b, c, e, f = v[k - 2], v[k - 1], v[k], v[k + 1], v[k + 2]
            dst = k
        else:
            # mirrored: the face k-1/2 seen from the right
            a, b, c, e, f = v[k + 2], v[k + 1], v[k], v[k - 1], v[k - 2]
            dst = k - 1
        q = np.array([a / 3 - 7 * b / 6 + 11 * c / 6, -b / 6 + 5 * c / 6 + e / 3, c / 3 + 5 * e / 6 - f / 6])
        beta = np.array([13 / 12 * (a - 2 * b + c) ** 2 + 0.25 * (a - 4 * b + 3 * c) ** 2,
                         13 / 12 * (b - 2 * c + e) ** 2 + 0.25 * (b - e) ** 2,
                         13 / 12 * (c - 2 * e + f) ** 2 + 0.25 * (3 * c - 4 * e + f) ** 2])
        alpha = d / (eps + beta) ** 2
        h[dst] = np.dot(alpha / alpha.sum(), q)
    return h


@pytest.mark.parametrize("side", [1, -1])
def test_weno5_quadratic_exact(side):
    rng = np.random.default_rng(7)
    x = np.arange(12.0)
    for _ in range(5):
        c0, c1, c2 = rng.uniform(-1, 1, 3)
        avg = c0 + c1 * x + c2 * (x * x + 1.0 / 12.0)
        h = O.pt_weno5(side, avg)
        faces = x + 0.5
        exact = c0 + c1 * faces + c2 * faces * faces
        ok = ~np.isnan(h)
        assert ok.sum() == len(x) - 4
        np.testing.assert_allclose(h[ok], exact[ok], rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("side", [1, -1])
def test_weno5_vs_textbook(side):
    rng = np.random.default_rng(11 + side)
    for _ in range(10):
        v = np.cumsum(rng.normal(size=16))
        v[rng.integers(4, 12)] += 5.0  # a jump
        h = O.pt_weno5(side, v)
        t = weno5_textbook(v, side)
        ok = ~np.isnan(h)
        assert np.array_equal(ok, ~np.isnan(t))
        np.testing.assert_allclose(h[ok], t[ok], rtol=1e-11, atol=1e-12)


def test_steger_warming_split():
    rng = np.random.default_rng(13)
    for _ in range(30):
        d, p = rng.uniform(0.2, 3, 2)
        u, v = rng.uniform(-2, 2, 2)
        A1, A2 = rng.uniform(-2, 2, 2)
        c = np.sqrt(G * p / d)
        fp, fm = O.pt_split_sw(d, u, v, c, A1, A2, G)
        V = A1 * u + A2 * v
        E = p / (G - 1.0) + 0.5 * d * (u * u + v * v)
        exact = np.array([d * V, d * u * V + p * A1, d * v * V + p * A2, (E + p) * V])
        np.testing.assert_allclose(fp + fm, exact, rtol=1e-12, atol=1e-12 * np.abs(exact).max())
    # all eigenvalues positive -> no negative part
    fp, fm = O.pt_split_sw(1.0, 3.0, 0.0, 1.0, 1.0, 0.0, G)
    assert np.all(fm == 0.0)


def test_dfn_is_normal_flux_difference():
    rng = np.random.default_rng(17)

    def fn(U, n1, n2):
        d, mu, mv, E = U[:4]
        u, v = mu / d, mv / d
        p = (G - 1.0) * (E - 0.5 * d * (u * u + v * v))
        un = u * n1 + v * n2
        f = [d * un, mu * un + p * n1, mv * un + p * n2, (E + p) * un] + [x * un for x in U[4:]]
        return np.array(f)

    for nv in (4, 5, 6):
        for _ in range(10):
            d, p = rng.uniform(0.3, 2, 2)
            u, v = rng.uniform(-1, 1, 2)
            U = np.array([d, d * u, d * v, p / (G - 1) + 0.5 * d * (u * u + v * v)] + list(rng.uniform(0, 1, nv - 4)))
            dU = rng.normal(scale=1e-2, size=nv)
            th = rng.uniform(0, 2 * np.pi)
            n1, n2 = np.cos(th), np.sin(th)
            DF = O.pt_dfn(U, dU, n1, n2, G)
            exact = fn(U + dU, n1, n2) - fn(U, n1, n2)
            np.testing.assert_allclose(DF, exact, rtol=1e-9, atol=1e-13)
            assert np.all(O.pt_dfn(U, np.zeros(nv), n1, n2, G) == 0.0)
