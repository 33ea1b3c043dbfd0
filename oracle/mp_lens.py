"""High-precision (mpmath) evaluation of the point-source route, for condition-aware
parity.  TEST INFRASTRUCTURE ONLY.

The FP64 polynomial route (lens frame, coefficients, roots, image filter, sum of
1/|det J|) loses digits near caustics, where images are nearly double roots: both the
GPU kernels and the FP64 oracle then carry errors far above 1e-10 that differ from
each other.  What the mathematics fixes is the value for the exact inputs; this module
computes it in 40-digit arithmetic with every input (source position and user lens
parameters, as given in FP64) taken as exact, so that a test can ask whether a result
is as close to it as FP64 evaluations of the oracle are.

Same formulas as the oracle: parameter mapping of point_source.py:1798-1817, lens
polynomial in product form (oracle/lens.py: poly_coeffs_product), image filter
|lens_eq(z) - w| < 1e-6 (point_source.py:1703-1707), magnification sum 1/|det J| over
real images (point_source.py:1828-1830).
"""
import mpmath as mp
import numpy as np

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


def frame_mp(nlenses, s=None, q=None, q3=None, r3=None, psi=None):
    """Lens positions, masses and x_cm at DPS digits from the user parameters taken as
    exact (the mapping of point_source.py:1798-1817)."""
    s, q = mp.mpf(float(s)), mp.mpf(float(q))
    a = s / 2
    x_cm = a * (1 - q) / (1 + q)
    if nlenses == 2:
        e1 = 1 / (1 + q)
        return [mp.mpc(a), mp.mpc(-a)], [e1, 1 - e1], x_cm
    q3 = mp.mpf(float(q3))
    e1 = q / (1 + q + q3)
    e2 = q * e1
    r3c = mp.mpf(float(r3)) * mp.expj(mp.mpf(float(psi)))
    return [mp.mpc(a), mp.mpc(-a), r3c], [e1, e2, 1 - e1 - e2], x_cm


def mag_point_source_mp(w, nlenses, **params):
    """(magnification, undecided) for each source position w (user frame), with every
    input (w and the user lens parameters) taken as exact.

    undecided is the part of the magnification that an FP64 evaluation cannot decide:
    inf where a root's residual lies within a factor 10 of the 1e-6 filter, else the
    sum of the contributions of real images that sit in a root cluster (another root
    closer than 1e-3 |z|): FP64 roots there carry errors up to eps^(1/m) of the cluster
    size, enough to push their residual past the filter."""
    w = np.atleast_1d(np.asarray(w, complex))
    out = np.zeros(w.shape)
    amb = np.zeros(w.shape)
    for i, wi in enumerate(w):
        with mp.workdps(DPS):
            r, e, x_cm = frame_mp(nlenses, **params)
            wl = mp.mpc(complex(wi)) + x_cm
            c = poly_mp(wl, r, e)
            z = mp.polyroots(c, maxsteps=400, extraprec=4 * DPS)
            res = []
            for zi in z:
                zb = mp.conj(zi)
                res.append(abs(zi - sum(ej / (zb - mp.conj(rj)) for ej, rj in zip(e, r)) - wl))
            m = mp.mpf(0)
            for k, (zi, ri) in enumerate(zip(z, res)):
                if 1e-7 < ri < 1e-5:
                    amb[i] = np.inf
                if ri < 1e-6:
                    zb = mp.conj(zi)
                    K = sum(ej / (zb - mp.conj(rj)) ** 2 for ej, rj in zip(e, r))
                    c = 1 / abs(1 - abs(K) ** 2)
                    m += c
                    dmin = min(abs(zi - zj) for j, zj in enumerate(z) if j != k)
                    if dmin < 1e-3 * max(1, abs(zi)):
                        amb[i] += float(c)
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


def mag_point_source_mpf(w, nlenses, **params):
    """Point-source magnification at one source position w (user frame; an mpmath
    number or a complex taken as exact), returned as an mpmath number at the working
    precision (no rounding to FP64), so that finite-difference stencils of it can use
    small steps.  Call inside mp.workdps(DPS)."""
    r, e, x_cm = frame_mp(nlenses, **params)
    wl = mp.mpc(w) + x_cm
    z = mp.polyroots(poly_mp(wl, r, e), maxsteps=400, extraprec=4 * mp.mp.dps)
    m = mp.mpf(0)
    for zi in z:
        zb = mp.conj(zi)
        if abs(zi - sum(ej / (zb - mp.conj(rj)) for ej, rj in zip(e, r)) - wl) < 1e-6:
            K = sum(ej / (zb - mp.conj(rj)) ** 2 for ej, rj in zip(e, r))
            m += 1 / abs(1 - abs(K) ** 2)
    return m
