"""NumPy FP64 restatement of the hexadecapole approximation and the contour gate.
TEST INFRASTRUCTURE ONLY.

Follows (fbartolic/caustics):
  _mag_hexadecapole_cassan ... src/caustics/multipole.py:8-205 (Cassan 2017 recurrences,
                               the explicit order-2..5 expressions kept term for term)
  _mag_hexadecapole .......... src/caustics/multipole.py:208-244
  _caustics_proximity_test ... src/caustics/lightcurve.py:23-75
  _planetary_caustic_test .... src/caustics/lightcurve.py:78-85
  gate in mag() .............. src/caustics/lightcurve.py:200-227 (hexadecapole without u1;
                               planetary test only when q < 0.01)
"""
from math import factorial

import numpy as np

from . import lens

cj = np.conjugate


def hexadecapole_cassan(W, rho, u1=0.0):
    W2, W3, W4, W5, W6 = W
    Gamma = 2 * u1 / (3.0 - u1)

    def akl(mu, Q):
        return mu * (cj(Q) + cj(W2) * Q)

    mu0 = 1 / (1.0 - np.abs(W2) ** 2)
    a10 = mu0 * (1.0 + cj(W2))
    a01 = mu0 * 1j * (1.0 - cj(W2))
    # p = 2
    a20 = akl(mu0, W3 * a10 * a10)
    a11 = akl(mu0, W3 * a10 * a01)
    a02 = akl(mu0, W3 * a01 * a01)
    # p = 3
    a30 = akl(mu0, W3 * 3.0 * a20 * a10 + W4 * a10**3)
    a21 = akl(mu0, W3 * (2.0 * a11 * a10 + a01 * a20) + W4 * a01 * a10 * a10)
    a12 = akl(mu0, W3 * (2.0 * a11 * a01 + a10 * a02) + W4 * a10 * a01 * a01)
    a03 = akl(mu0, W3 * 3.0 * a02 * a01 + W4 * a01**3)
    # p = 4
    a40 = akl(mu0, W3 * (4.0 * a30 * a10 + 3.0 * a20 * a20) + W4 * 6.0 * a20 * a10 * a10
              + W5 * a10**4)
    a31 = akl(mu0, W3 * (3.0 * a21 * a10 + 3.0 * a11 * a20 + a01 * a30)
              + W4 * 3.0 * (a11 * a10 * a10 + a01 * a20 * a10) + W5 * a01 * a10**3)
    a22 = akl(mu0, W3 * (2.0 * a12 * a10 + 2.0 * a11 * a11 + a02 * a20 + 2.0 * a21 * a01)
              + W4 * (a02 * a10 * a10 + 4.0 * a11 * a01 * a10 + a01 * a01 * a20)
              + W5 * a01 * a01 * a10 * a10)
    a13 = akl(mu0, W3 * (3.0 * a12 * a01 + 3.0 * a11 * a02 + a10 * a03)
              + W4 * 3.0 * (a11 * a01 * a01 + a10 * a02 * a01) + W5 * a10 * a01**3)
    a04 = akl(mu0, W3 * (4.0 * a03 * a01 + 3.0 * a02 * a02) + W4 * 6.0 * a02 * a01 * a01
              + W5 * a01**4)
    # p = 5
    a50 = akl(mu0, W3 * (5.0 * a40 * a10 + 10.0 * a20 * a30)
              + W4 * (10.0 * a30 * a10 * a10 + 15.0 * a20 * a20 * a10)
              + W5 * 10.0 * a20 * a10**3 + W6 * a10**5)
    a41 = akl(mu0, W3 * (4.0 * a31 * a10 + 4.0 * a11 * a30 + 6.0 * a21 * a20 + a01 * a40)
              + W4 * (6.0 * a21 * a10 * a10 + 12.0 * a11 * a20 * a10 + 3.0 * a01 * a20 * a20
                      + 4.0 * a01 * a10 * a30)
              + W5 * (4.0 * a11 * a10**3 + 6.0 * a01 * a10 * a10 * a20)
              + W6 * a01 * a10**4)
    a32 = akl(mu0, W3 * (3.0 * a22 * a10 + 6.0 * a11 * a21 + 3.0 * a12 * a20 + a02 * a30
                         + 2.0 * a31 * a01)
              + W4 * (3.0 * a12 * a10 * a10 + 6.0 * a11 * a11 * a10 + 3.0 * a02 * a20 * a10
                      + 6.0 * a01 * a11 * a20 + 6.0 * a01 * a21 * a10 + a01 * a01 * a30)
              + W5 * (a02 * a10**3 + 6.0 * a11 * a01 * a10 * a10 + 3.0 * a01 * a01 * a10 * a20)
              + W6 * a01 * a01 * a10**3)
    a23 = akl(mu0, W3 * (3.0 * a22 * a01 + 6.0 * a11 * a12 + 3.0 * a21 * a02 + a20 * a03
                         + 2.0 * a13 * a10)
              + W4 * (3.0 * a21 * a01 * a01 + 6.0 * a11 * a11 * a01 + 3.0 * a20 * a02 * a01
                      + 6.0 * a10 * a11 * a02 + 6.0 * a10 * a12 * a01 + a10 * a10 * a03)
              + W5 * (a20 * a01**3 + 6.0 * a11 * a10 * a01 * a01 + 3.0 * a10 * a10 * a01 * a02)
              + W6 * a10 * a10 * a01**3)
    a14 = akl(mu0, W3 * (4.0 * a13 * a01 + 4.0 * a11 * a03 + 6.0 * a12 * a02 + a10 * a04)
              + W4 * (6.0 * a12 * a01 * a01 + 12.0 * a11 * a02 * a01 + 3.0 * a10 * a02 * a02
                      + 4.0 * a10 * a01 * a03)
              + W5 * (4.0 * a11 * a01**3 + 6.0 * a10 * a01 * a01 * a02)
              + W6 * a10 * a01**4)
    a05 = akl(mu0, W3 * (5.0 * a04 * a01 + 10.0 * a02 * a03)
              + W4 * (10.0 * a03 * a01 * a01 + 15.0 * a02 * a02 * a01)
              + W5 * 10.0 * a02 * a01**3 + W6 * a01**5)
    mu2 = 0.25 * np.imag(a01 * cj(a12 + a30) + cj(a10) * (a03 + a21) + 2.0 * a02 * cj(a11)
                         + 2.0 * a11 * cj(a20))
    mu4 = 0.125 * np.imag((a05 + 2.0 * a23 + a41) * cj(a10) + 4.0 * a04 * cj(a11)
                          + 4 * (a13 + a31) * cj(a20) + 6.0 * a12 * cj(a21)
                          + 6.0 * a21 * cj(a30) + a03 * (6.0 * cj(a12) + 2.0 * cj(a30))
                          + 4.0 * a02 * cj(a13 + a31) + 4.0 * a11 * cj(a40)
                          + a01 * cj(a14 + 2.0 * a32 + a50))
    dq = 0.5 * mu2 * (1.0 - 0.2 * Gamma) * rho**2
    dh = mu4 * (1.0 - 11.0 / 35.0 * Gamma) * rho**4 / 24.0
    return mu0, dq, dh


def W_terms(z, nlenses, **lp):
    """W_k = (-1)^(k-1) (k-1)! sum_j e_j / (z - r_j)^k, k = 2..6 (multipole.py:212-231)."""
    if nlenses == 1:
        rs, es = [0.0], [1.0]
    elif nlenses == 2:
        rs, es = [lp["a"], -lp["a"]], [lp["e1"], 1.0 - lp["e1"]]
    else:
        rs = [lp["a"], -lp["a"], lp["r3"]]
        es = [lp["e1"], lp["e2"], 1.0 - lp["e1"] - lp["e2"]]
    out = []
    for k in range(2, 7):
        s = sum(e / (z - r) ** k for e, r in zip(es, rs))
        out.append((-1) ** (k - 1) * factorial(k - 1) * s)
    return out


def mag_hexadecapole(z, mask, rho, u1=0.0, nlenses=2, **lp):
    """(mu_multi, |dq|+|dh| summed over real images), multipole.py:208-244."""
    mu0, dq, dh = hexadecapole_cassan(W_terms(z, nlenses, **lp), rho, u1)
    mu_multi = np.sum(mask * np.abs(mu0 + dq + dh), axis=0)
    return mu_multi, np.sum(mask * np.abs(dq), axis=0) + np.sum(mask * np.abs(dh), axis=0)


def lens_map_f(x, a, e1):
    """f of the binary lens map w = z + f(conj z) (lightcurve.py:30)."""
    return -e1 / (x - a) - (1 - e1) / (x + a)


def lens_map_fp(x, a, e1):
    """f' (lightcurve.py:31)."""
    return e1 / (x - a) ** 2 + (1 - e1) / (x + a) ** 2


def lens_map_fpp(x, a, e1):
    """f'' (lightcurve.py:32)."""
    return 2 * (e1 / (a - x) ** 3 - (1 - e1) / (a + x) ** 3)


def caustics_proximity_test(w, z, mask, rho, dmu, nlenses=2, c_m=1e-2, gamma=0.02, c_f=4.0,
                            rho_min=1e-3, **lp):
    """lightcurve.py:23-75 (binary branch)."""
    a, e1 = lp["a"], lp["e1"]
    f = lambda x: lens_map_f(x, a, e1)
    fp = lambda x: lens_map_fp(x, a, e1)
    fpp = lambda x: lens_map_fpp(x, a, e1)
    zb = cj(z)
    zhat = cj(w) - f(z)
    fp_z, fpp_z, fp_zb, fp_zh, fpp_zb = fp(z), fpp(z), fp(zb), fp(zhat), fpp(zb)
    J = 1.0 - np.abs(fp_z * fp_zb)
    mu_cusp = 6 * np.imag(3 * fp_zb**3 * fpp_z**2) / J**5 * (rho + rho_min) ** 2
    mu_cusp = np.sum(np.abs(mu_cusp) * mask, axis=0)
    t_mc = gamma * mu_cusp + dmu < c_m
    Jh = 1 - np.abs(fp_z * fp_zh)
    factor = np.abs(J * Jh**2 / (Jh * fpp_zb * fp_z - cj(Jh) * fpp_z * fp_zb * fp_zh))
    t_fi = 0.5 * np.sum(~mask * factor, axis=0) > c_f * (rho + rho_min)
    t_fi = np.where(np.sum(~mask, axis=0) == 0, True, t_fi)
    return t_fi & t_mc


def planetary_caustic_test(w, rho, c_p=2.0, **lp):
    """lightcurve.py:78-85 (q here is e1/(1-e1), as in the reference)."""
    e1, a = lp["e1"], lp["a"]
    s = 2 * a
    q = e1 / (1 - e1)
    w_pc = -1 / s
    d_pc = 3 * np.sqrt(q) / s
    return (w_pc - w).real ** 2 + (w_pc - w).imag ** 2 > c_p * (rho**2 + d_pc**2)


def hexadecapole_gate(w_points, rho, nlenses=2, perturb=0.0, seed=0, **params):
    """The per-point decision of mag() (lightcurve.py:168-227): returns
    (mu_hex, use_hexadecapole) with the reference's choices (no u1 in the multipole
    terms, planetary test for q < 0.01).  nlenses=3 always needs the contour."""
    lp, x_cm = lens.lens_params(nlenses, **params)
    wl = np.asarray(w_points, complex) + x_cm
    z, mask = lens.images_point_source(wl, nlenses, perturb=perturb, seed=seed, **lp)
    mu, dmu = mag_hexadecapole(z, mask, rho, 0.0, nlenses, **lp)
    if nlenses == 3:
        return mu, np.zeros(wl.shape, bool)
    if nlenses == 1:
        return mu, np.abs(wl) > 2 * rho
    t1 = caustics_proximity_test(wl, z, mask, rho, dmu, 2, **lp)
    if params["q"] < 0.01:
        t1 = t1 & planetary_caustic_test(wl, rho, **lp)
    return mu, t1
