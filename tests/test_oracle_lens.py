"""Pins the oracle's point-source magnification and gradient (oracle/lens.py).

Closed forms: the single-lens Paczynski magnification; the binary lens at q -> 0
tends to the single lens; magnification of a far source -> 1.  The gradient of the
polynomial route (the reference's implicit root derivative chained through the
coefficients) is checked against central differences and against an independent
lens-equation derivative evaluated here."""
import numpy as np
import pytest

from oracle import lens
from caustics_amd import workloads as W


def test_single_lens_paczynski():
    u = np.linspace(0.05, 4, 100)
    w = u * np.exp(1j * np.linspace(0, 3, 100))
    m = lens.mag_point_source(w, nlenses=1)
    np.testing.assert_allclose(m, (u**2 + 2) / (u * np.sqrt(u**2 + 4)), rtol=1e-13)


def test_binary_small_q_tends_to_single_lens():
    # a tiny second mass far away perturbs the single-lens magnification by O(q)
    w = np.array([0.3 + 0.2j, -0.5 + 0.1j, 0.05 - 0.4j])
    s = 10.0
    q = 1e-9
    a = 0.5 * s
    x1 = -s * q / (1 + q)   # heavy lens position in the centre-of-mass frame
    m = lens.mag_point_source(w + x1, nlenses=2, s=s, q=q)
    u = np.abs(w)
    np.testing.assert_allclose(m, (u**2 + 2) / (u * np.sqrt(u**2 + 4)), rtol=1e-6)


@pytest.mark.parametrize("nlenses", [2, 3])
def test_polynomial_route_gradient_vs_fd(nlenses):
    g = W.rng(70 + nlenses)
    n = 60
    p = W.lens_configs(g, n, nlenses)
    w = W.disk(g, n, radius=1.2)
    names = ["s", "q"] if nlenses == 2 else ["s", "q", "q3", "r3", "psi"]
    h = 1e-6
    checked = 0
    for i in range(n):
        pi = {k: p[k][i] for k in names}
        if not filter_unambiguous(w[i], nlenses, pi):
            continue   # a root's residual is near the 1e-6 filter: finite differences jump
        checked += 1
        m, gr = lens.mag_point_source_and_grad(np.array([w[i]]), nlenses=nlenses, **pi)
        for k, key in enumerate(["wre", "wim"] + names):
            pp, pm = dict(pi), dict(pi)
            wp, wm = w[i], w[i]
            if key == "wre":
                wp, wm = w[i] + h, w[i] - h
            elif key == "wim":
                wp, wm = w[i] + 1j * h, w[i] - 1j * h
            else:
                pp[key] += h
                pm[key] -= h
            fd = (lens.mag_point_source(wp, nlenses=nlenses, **pp)
                  - lens.mag_point_source(wm, nlenses=nlenses, **pm)) / (2 * h)
            assert abs(fd - gr[0, k]) <= 1e-4 * max(abs(fd), 1e-2 * m[0]), (i, key, fd, gr[0, k])
    assert checked >= n // 2


def filter_unambiguous(w, nlenses, params, lo=1e-8, hi=1e-4):
    """True when no root's lens-equation residual lies within [lo, hi] (far from the
    1e-6 image filter threshold, so a 1e-6 perturbation cannot flip the mask)."""
    lp, x_cm = lens.lens_params(nlenses, **params)
    z, _ = lens.images_point_source(np.array([w + x_cm]), nlenses, **lp)
    r = np.abs(lens.lens_eq(z[:, 0], nlenses, **lp) - (w + x_cm))
    return not np.any((r > lo) & (r < hi))


def test_lens_equation_route_agrees_with_polynomial_route():
    # dz = (u - K conj(u)) / det for the image z of a source move dw = u
    g = W.rng(9)
    a, e1 = 0.45, 0.7
    for _ in range(20):
        w = complex(*g.uniform(-0.8, 0.8, 2))
        z, mask = lens.images_point_source(np.array([w]), 2, a=a, e1=e1)
        z = z[:, 0][mask[:, 0]]
        zb = np.conj(z)
        K = e1 / (zb - a) ** 2 + (1 - e1) / (zb + a) ** 2
        det = 1 - np.abs(K) ** 2
        for u in (1.0, 1j):
            dz = (u - K * np.conj(u)) / det
            h = 1e-7
            zp, mp = lens.images_point_source(np.array([w + h * u]), 2, a=a, e1=e1)
            zp = zp[:, 0]
            near = zp[np.argmin(np.abs(zp[:, None] - z[None, :]), axis=0)]
            np.testing.assert_allclose((near - z) / h, dz, rtol=1e-5, atol=1e-6)


def test_sequential_images_are_the_same_roots_followed_continuously():
    # images_point_source_sequential (point_source.py:1711-1759): the same root set as
    # independent solves at every position, with each row the continuation of itself
    lp = dict(a=0.6, e1=0.5, e2=0.3, r3=0.2 + 0.7j)
    t = np.linspace(-1.5, 1.5, 300)
    w = (0.2j + t) * np.exp(0.3j)
    zs, ms = lens.images_point_source_sequential(w, nlenses=3, init_mode=1, **lp)
    zi, mi = lens.images_point_source(w, nlenses=3, init_mode=1, **lp)
    d = np.abs(np.sort_complex(zs.T) - np.sort_complex(zi.T))
    assert np.max(d) < 1e-9
    step = np.abs(np.diff(zs, axis=1))
    alt = np.abs(zs[None, :, 1:] - zs[:, None, :-1]).min(axis=1)
    assert np.mean(step <= alt + 1e-12) > 0.99
    assert ms.sum() == mi.sum()
