"""Pins the hexadecapole and gate oracle (oracle/multipole.py) with what the mathematics
fixes, independently of the Cassan (2017) expressions it restates
(src/caustics/multipole.py:8-244, src/caustics/lightcurve.py:23-85).

- W_k: W_1 = conj(z - lens_eq(z)) is the deflection, and W_{k+1} = dW_k/dz
  (central differences of the oracle's own lens equation and W terms).
- mu2, mu4 per image: the average of the point-source magnification A over a uniform
  disk of radius rho is A + rho^2 Lap(A)/8 + rho^4 Lap^2(A)/192 + O(rho^6), and the
  hexadecapole value is sum_images |mu0 + mu2 rho^2/2 + mu4 rho^4/24|.  So
  sum sign(mu0) mu2 = Lap(A)/4 and sum sign(mu0) mu4 = Lap^2(A)/8, with the Laplacians
  taken by finite-difference stencils of the 40-digit point-source magnification.
  A dropped term, wrong coefficient, sign or index in any a_pq recurrence, in mu2 or
  in mu4 fails this at once.
- Limb-darkening factors (1 - Gamma/5), (1 - 11 Gamma/35): ratios of the second and
  fourth radial moments of the linear profile to the uniform disk's, by quadrature.
- Error order: for one lens the hexadecapole error against the disk average (closed-
  form Paczynski magnification, polar Gauss-Legendre quadrature) falls as rho^6,
  uniform and limb-darkened.
- Gate: the lens-map derivatives f', f'' are derivatives of f, and w = z + f(conj z);
  wherever the binary gate accepts a point, the hexadecapole value is within the
  reference's rtol 5e-4 of the disk average (tests/test_lightcurve.py:30-69, with the
  source-plane quadrature standing in for VBBinaryLensing)."""
import mpmath as mp
import numpy as np
import pytest

from oracle import lens
from oracle import mp_lens
from oracle import multipole as M

BINARY = [dict(s=0.9, q=0.2), dict(s=1.3, q=0.05)]
TRIPLE = [dict(s=1.2, q=0.3, q3=0.01, r3=0.8, psi=1.0), dict(s=0.7, q=0.5, q3=0.05, r3=1.1,
                                                             psi=2.0)]
CASES = [(2, BINARY[0], 0.3 + 0.4j), (2, BINARY[1], -0.2 + 0.5j),
         (3, TRIPLE[0], 0.5 - 0.6j), (3, TRIPLE[1], -0.4 + 0.3j)]


def _images(w, nlenses, prm):
    lp, x_cm = lens.lens_params(nlenses, **prm)
    z, mask = lens.images_point_source(np.atleast_1d(np.asarray(w, complex)) + x_cm, nlenses,
                                       **lp)
    return z, mask, lp


@pytest.mark.parametrize("nlenses,prm,w", CASES)
def test_W_terms_are_derivatives_of_the_deflection(nlenses, prm, w):
    z, mask, lp = _images(w, nlenses, prm)
    z = z[mask]
    rs, _ = lens.lens_positions_masses(nlenses, **lp)
    W1 = lambda x: np.conj(x - lens.lens_eq(x, nlenses, **lp))
    W = lambda x: [W1(x)] + M.W_terms(x, nlenses, **lp)
    # steps scaled by each image's distance to the nearest lens (the derivatives grow
    # as that distance shrinks); central differences at h and h/2, Richardson-combined
    h = 1e-3 * np.abs(z[:, None] - rs[None, :]).min(1)
    W0 = W(z)
    for u in (1.0, 1j):   # along both axes: W_k is analytic, the two must agree
        def D(hh):
            Wp, Wm = W(z + u * hh), W(z - u * hh)
            return [(Wp[k] - Wm[k]) / (2 * u * hh) for k in range(5)]
        D1, D2 = D(h), D(h / 2)
        for k in range(5):
            np.testing.assert_allclose((4 * D2[k] - D1[k]) / 3, W0[k + 1], rtol=1e-8, atol=0)


def _laplacians(nlenses, prm, w, h=1e-4):
    """Lap(A) and Lap^2(A) at w by the 5- and 13-point stencils at steps h and h/2,
    Richardson-combined, on the 40-digit point-source magnification."""
    with mp.workdps(mp_lens.DPS):
        w0, h = mp.mpc(w), mp.mpf(h)
        f = lambda x: mp_lens.mag_point_source_mpf(x, nlenses, **prm)

        def stencils(hh):
            cache = {}

            def F(i, j):
                if (i, j) not in cache:
                    cache[(i, j)] = f(w0 + hh * mp.mpc(i, j))
                return cache[(i, j)]
            ring = F(1, 0) + F(-1, 0) + F(0, 1) + F(0, -1)
            lap = (ring - 4 * F(0, 0)) / hh**2
            bil = (20 * F(0, 0) - 8 * ring
                   + 2 * (F(1, 1) + F(1, -1) + F(-1, 1) + F(-1, -1))
                   + F(2, 0) + F(-2, 0) + F(0, 2) + F(0, -2)) / hh**4
            return lap, bil
        l1, b1 = stencils(h)
        l2, b2 = stencils(h / 2)
        return float((4 * l2 - l1) / 3), float((4 * b2 - b1) / 3)


@pytest.mark.parametrize("nlenses,prm,w", CASES)
def test_hexadecapole_terms_are_taylor_coefficients_of_the_disk_average(nlenses, prm, w):
    z, mask, lp = _images(w, nlenses, prm)
    mu0, dq, dh = M.hexadecapole_cassan(M.W_terms(z, nlenses, **lp), 1.0, 0.0)
    sg = np.sign(mu0.real)
    mu2 = float(np.sum(mask * sg * 2.0 * dq))
    mu4 = float(np.sum(mask * sg * 24.0 * dh))
    lap, bil = _laplacians(nlenses, prm, w)
    print(f"nl={nlenses} w={w}: mu2 {mu2:.12g} vs Lap/4 {lap / 4:.12g}; "
          f"mu4 {mu4:.12g} vs Lap^2/8 {bil / 8:.12g}")
    assert abs(mu2 - lap / 4) <= 1e-9 * abs(mu2)
    assert abs(mu4 - bil / 8) <= 1e-8 * abs(mu4)


def _radial_moment(k, u1, n=64):
    """<r^k> over the unit disk weighted by the normalised linear limb-darkening profile
    3/(3-u1) (1 - u1 + u1 sqrt(1 - r^2)) (integrate.py:29-44), r = sin(phi)."""
    x, wx = np.polynomial.legendre.leggauss(n)
    phi = 0.25 * np.pi * (x + 1)
    r = np.sin(phi)
    B = 3.0 / (3.0 - u1) * (1.0 - u1 + u1 * np.cos(phi))
    return float(np.sum(0.25 * np.pi * wx * B * r**k * 2 * r * np.cos(phi)))


@pytest.mark.parametrize("u1", [0.2, 0.5, 0.9])
def test_limb_darkening_factors_are_moment_ratios(u1):
    z, mask, lp = _images(0.3 + 0.4j, 2, BINARY[0])
    W = M.W_terms(z, 2, **lp)
    _, dq0, dh0 = M.hexadecapole_cassan(W, 0.05, 0.0)
    _, dq, dh = M.hexadecapole_cassan(W, 0.05, u1)
    np.testing.assert_allclose(dq / dq0, _radial_moment(2, u1) / _radial_moment(2, 0.0),
                               rtol=1e-13)
    np.testing.assert_allclose(dh / dh0, _radial_moment(4, u1) / _radial_moment(4, 0.0),
                               rtol=1e-13)


def disk_average(f, w, rho, u1=0.0, nr=48, nt=128):
    """(1/pi rho^2) int_disk B(|x|/rho) f(w + x) dA: Gauss-Legendre in phi with
    r = rho sin(phi) (smooth in phi for the limb-darkened profile) and the trapezoid
    rule in the angle (periodic).  w may be an array of centres."""
    x, wx = np.polynomial.legendre.leggauss(nr)
    phi = 0.25 * np.pi * (x + 1)
    wphi = 0.25 * np.pi * wx
    th = 2 * np.pi * np.arange(nt) / nt
    r = rho * np.sin(phi)
    jac = rho * rho * np.sin(phi) * np.cos(phi) * wphi
    B = 3.0 / (3.0 - u1) * (1.0 - u1 + u1 * np.cos(phi))
    w = np.atleast_1d(np.asarray(w, complex))
    pts = w[:, None, None] + r[None, :, None] * np.exp(1j * th)[None, None, :]
    A = f(pts.reshape(-1)).reshape(pts.shape)
    return (A * (jac * B)[None, :, None]).sum((1, 2)) * (2 * np.pi / nt) / (np.pi * rho * rho)


def paczynski(w):
    u = np.abs(w)
    return (u * u + 2) / (u * np.sqrt(u * u + 4))


@pytest.mark.parametrize("u,u1", [(0.3, 0.0), (0.3, 0.5), (0.5, 0.0), (0.5, 0.5)])
def test_hexadecapole_error_is_sixth_order_single_lens(u, u1):
    errs, qerrs = [], []
    for rho in (0.05, 0.025, 0.0125):
        exact = disk_average(paczynski, u, rho, u1)[0]
        assert abs(exact - disk_average(paczynski, u, rho, u1, 64, 192)[0]) < 1e-13 * exact
        z, mask = lens.images_point_source(np.array([u + 0j]), 1)
        mu, _ = M.mag_hexadecapole(z, mask, rho, u1, nlenses=1)
        mu0, dq, _ = M.hexadecapole_cassan(M.W_terms(z, 1), rho, u1)
        errs.append(abs(mu[0] - exact))
        qerrs.append(abs(np.sum(np.abs(mu0 + dq)) - exact))
    slope = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    qslope = np.log2(np.array(qerrs[:-1]) / np.array(qerrs[1:]))
    print(f"u={u} u1={u1}: hexadecapole error slope {slope}, quadrupole-only {qslope}")
    assert np.all(np.abs(slope - 6.0) < 0.05)
    assert np.all(np.abs(qslope - 4.0) < 0.05)


@pytest.mark.parametrize("nlenses,prm,w", CASES[:2])
def test_lens_map_derivatives(nlenses, prm, w):
    z, mask, lp = _images(w, nlenses, prm)
    a, e1 = lp["a"], lp["e1"]
    x = np.concatenate([z.reshape(-1), np.conj(z.reshape(-1))])
    # the lens equation is w = z + f(conj z)
    np.testing.assert_allclose(lens.lens_eq(x, 2, **lp), x + M.lens_map_f(np.conj(x), a, e1),
                               rtol=1e-14, atol=1e-14)
    h = 1e-5
    for f, fp in ((M.lens_map_f, M.lens_map_fp), (M.lens_map_fp, M.lens_map_fpp)):
        d = (f(x + h, a, e1) - f(x - h, a, e1)) / (2 * h)
        np.testing.assert_allclose(d, fp(x, a, e1), rtol=1e-7)


def _gate_systems(g, n, correlated):
    """(q, s, rho) as the reference's test draws them (one uniform variate for all three,
    tests/test_lightcurve.py:36-39) or independently over the same ranges."""
    for _ in range(n):
        u = g.uniform(size=3)
        if correlated:
            u[:] = u[0]
        yield (10 ** (-6 + 6 * u[0]), 10 ** (-1 + (np.log10(4) + 1) * u[1]),
               10 ** (-3 + 2 * u[2]))


@pytest.mark.parametrize("correlated", [True, False])
def test_gate_accepts_only_accurate_hexadecapole(correlated):
    """Reference criterion (tests/test_lightcurve.py:30-69): 250-point caustics, test
    points within 15 rho of them; wherever the gate accepts, |mu_hex - A| / A <= 5e-4."""
    g = np.random.default_rng(11 if correlated else 12)
    n_acc = 0
    for q, s, rho in _gate_systems(g, 15, correlated):
        _, ca = lens.critical_and_caustic_curves(npts=250, nlenses=2, s=s, q=q)
        ca = ca.reshape(-1)
        w = ca + g.uniform(0, 15 * rho, ca.shape) * np.exp(1j * g.uniform(-np.pi, np.pi,
                                                                         ca.shape))
        mu, use = M.hexadecapole_gate(w, rho, nlenses=2, s=s, q=q)
        if not use.any():
            continue
        ps = lambda x: lens.mag_point_source(x, 2, s=s, q=q)
        A = disk_average(ps, w[use], rho, 0.0, 24, 48)
        A2 = disk_average(ps, w[use], rho, 0.0, 32, 64)
        assert np.all(np.abs(A - A2) <= 1e-7 * A2)      # the quadrature has converged
        err = np.abs(mu[use] - A2) / A2
        n_acc += int(use.sum())
        print(f"q={q:.2e} s={s:.3f} rho={rho:.2e}: accepted {use.sum()}/{len(w)}, "
              f"max rel err {err.max():.2e}")
        assert err.max() <= 5e-4
    assert n_acc >= 1000
