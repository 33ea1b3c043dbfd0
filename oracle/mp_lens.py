"""High-precision (mpmath) evaluation of the point-source route, for condition-aware
parity.  TEST INFRASTRUCTURE ONLY.

The FP64 polynomial route (coefficients, roots, image filter, sum of 1/|det J|) loses
digits near caustics, where images are nearly double roots: both the GPU kernels and
the FP64 oracle then carry errors far above 1e-10 that differ from each other.  What
the mathematics fixes is the value for the exact inputs; this module computes it in
40-digit arithmetic from the FP64 lens-frame inputs (w + x_cm, lens positions and
masses exactly as FP64 numbers), so that a test can ask whether a result is as close
to it as the FP64 oracle is.

Same formulas as the oracle: lens polynomial in product form (oracle/lens.py:
poly_coeffs_product), image filter |lens_eq(z) - w| < 1e-6
(src/caustics/point_source.py:1703-1707), magnification sum 1/|det J| over real images
(point_source.py:1828-1830), hexadecapole terms (oracle/multipole.py).
"""
import mpmath as mp
import numpy as np

from . import lens

DPS = 40


def _pmul(a, b):
    out = [mp.mpc(0)] * (len(a) + len(b) - 1)
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            out[i + j] += x * y
    return out


def _padd(a, b):
    n = max(len(a), len(b))
    a = [mp.mpc(0)] * (n - len(a)) + list(a)
    b = [mp.mpc(0)] * (n - len(b)) + list(b)
    return [x + y for x, y in zip(a, b)]


def _frame(nlenses, lp):
    r, e = lens.lens_positions_masses(nlenses, **lp)
    return [mp.mpc(complex(x)) for x in r], [mp.mpf(float(x)) for x in e]


def poly_mp(wl, r, e):
    """Highest-first product-form lens polynomial (oracle/lens.py poly_coeffs_product)."""
    m = len(r)
    wb = mp.conj(wl)
    lin = [[mp.mpc(1), -rj] for rj in r]
    D = [mp.mpc(1)]
    for f in lin:
        D = _pmul(D, f)
    N = [mp.mpc(0)]
    for j in range(m):
        t = [e[j]]
        for k in range(m):
            if k != j:
                t = _pmul(t, lin[k])
        N = _padd(N, t)
    M = [_padd([(wb - mp.conj(r[j])) * c for c in D], N) for j in range(m)]
    P = [mp.mpc(1), -wl]
    for j in range(m):
        P = _pmul(P, M[j])
    S = [mp.mpc(0)]
    for j in range(m):
        t = [e[j]]
        for k in range(m):
            if k != j:
                t = _pmul(t, M[k])
        S = _padd(S, t)
    return _padd(P, [-c for c in _pmul(D, S)])


def images_mp(wl, nlenses, lp):
    """(roots, lens-equation residuals) at DPS digits for one lens-frame source point."""
    with mp.workdps(DPS):
        r, e = _frame(nlenses, lp)
        w = mp.mpc(complex(wl))
        c = poly_mp(w, r, e)
        z = mp.polyroots(c, maxsteps=400, extraprec=4 * DPS)
        res = []
        for zi in z:
            zb = mp.conj(zi)
            wz = zi - sum(ej / (zb - mp.conj(rj)) for ej, rj in zip(e, r))
            res.append(abs(wz - w))
        return z, res, r, e


def mag_point_source_mp(w, nlenses, **params):
    """(magnification, ambiguous) for each source position w (user frame).
    ambiguous is True where a root's residual lies within a factor 10 of the 1e-6
    filter, so that rounding can legitimately flip the image count."""
    lp, x_cm = lens.lens_params(nlenses, **params)
    w = np.atleast_1d(np.asarray(w, complex))
    out = np.zeros(w.shape)
    amb = np.zeros(w.shape, bool)
    for i, wi in enumerate(w):
        wl = complex(wi + x_cm)
        with mp.workdps(DPS):
            z, res, r, e = images_mp(wl, nlenses, lp)
            m = mp.mpf(0)
            for zi, ri in zip(z, res):
                if 1e-7 < ri < 1e-5:
                    amb[i] = True
                if ri < 1e-6:
                    zb = mp.conj(zi)
                    K = sum(ej / (zb - mp.conj(rj)) ** 2 for ej, rj in zip(e, r))
                    m += 1 / abs(1 - abs(K) ** 2)
            out[i] = float(m)
    return out, amb


def roots_mp(coeffs_asc):
    """Roots (DPS digits, rounded to complex128) of each row of ascending FP64
    coefficients taken as exact."""
    c = np.atleast_2d(coeffs_asc)
    out = np.zeros((c.shape[0], c.shape[1] - 1), complex)
    with mp.workdps(DPS):
        for i, row in enumerate(c):
            z = mp.polyroots([mp.mpc(complex(x)) for x in row[::-1]], maxsteps=400,
                             extraprec=4 * DPS)
            out[i] = [complex(x) for x in z]
    return out
