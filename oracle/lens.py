"""NumPy FP64 restatement of the point-source part of the path. TEST INFRASTRUCTURE ONLY.

Follows (file:line in the reference, fbartolic/caustics):
  _poly_coeffs_binary ......... src/caustics/point_source.py:17-86   (expanded monomials, restated)
  _poly_coeffs_triple ......... src/caustics/point_source.py:89-1479 (same polynomial, built here
                                from the product form of the lens equation; pinned by the
                                notebook-output fixture and the forward-map invariant)
  lens_eq / lens_eq_det_jac ... src/caustics/point_source.py:1536-1580
  _images_point_source ........ src/caustics/point_source.py:1652-1709 (filter |lens_eq(z)-w|<1e-6)
  mag_point_source ............ src/caustics/point_source.py:1762-1830 (parameter map + x_cm shift)
  implicit root derivative .... src/caustics/ehrlich_aberth_primitive.py:254-324
                                (dz = -(sum_k z^k dp_k) / p'(z))
"""
import numpy as np

from . import ea


# ----------------------------------------------------------------------------
# parameter maps (point_source.py:1798-1817)
# ----------------------------------------------------------------------------
def lens_params(nlenses, s=None, q=None, q3=None, r3=None, psi=None):
    """Map user parameters to lens-frame (a, e1, e2, r3) and the x_cm shift."""
    if nlenses == 1:
        return dict(), 0.0
    a = 0.5 * s
    x_cm = a * (1 - q) / (1 + q)
    if nlenses == 2:
        e1 = 1 / (1 + q)
        return dict(a=a, e1=e1), x_cm
    if nlenses == 3:
        e1 = q / (1 + q + q3)          # reference mapping, point_source.py:1810-1811
        e2 = q * e1
        r3c = r3 * np.exp(1j * psi)     # +i psi, point_source.py:1812
        return dict(a=a, e1=e1, e2=e2, r3=r3c), x_cm
    raise ValueError("nlenses must be 1, 2 or 3")


def lens_positions_masses(nlenses, a=None, e1=None, e2=None, r3=None):
    if nlenses == 2:
        return np.array([a, -a], complex), np.array([e1, 1.0 - e1])
    if nlenses == 3:
        return np.array([a, -a, r3], complex), np.array([e1, e2, 1.0 - e1 - e2])
    raise ValueError


# ----------------------------------------------------------------------------
# coefficients
# ----------------------------------------------------------------------------
def poly_coeffs_binary(w, a, e1):
    """Restatement of _poly_coeffs_binary (point_source.py:17-86), highest-first."""
    w = np.asarray(w, complex)
    wbar = np.conj(w)
    p_0 = -(a**2) + wbar**2
    p_1 = a**2 * w - 2 * a * e1 + a - w * wbar**2 + wbar
    p_2 = 2 * a**4 - 2 * a**2 * wbar**2 + 4 * a * wbar * e1 - 2 * a * wbar - 2 * w * wbar
    p_3 = (-2 * a**4 * w + 4 * a**3 * e1 - 2 * a**3 + 2 * a**2 * w * wbar**2
           - 4 * a * w * wbar * e1 + 2 * a * w * wbar + 2 * a * e1 - a - w)
    p_4 = (-(a**6) + a**4 * wbar**2 - 4 * a**3 * wbar * e1 + 2 * a**3 * wbar
           + 2 * a**2 * w * wbar + 4 * a**2 * e1**2 - 4 * a**2 * e1 + 2 * a**2
           - 4 * a * w * e1 + 2 * a * w)
    p_5 = (a**6 * w - 2 * a**5 * e1 + a**5 - a**4 * w * wbar**2 - a**4 * wbar
           + 4 * a**3 * w * wbar * e1 - 2 * a**3 * w * wbar + 2 * a**3 * e1 - a**3
           - 4 * a**2 * w * e1**2 + 4 * a**2 * w * e1 - a**2 * w)
    return np.stack([p_0, p_1, p_2, p_3, p_4, p_5], axis=-1)


def _pmul(p, q):
    """Product of polynomials stored ascending along the last axis."""
    n = p.shape[-1] + q.shape[-1] - 1
    out = np.zeros(np.broadcast_shapes(p.shape[:-1], q.shape[:-1]) + (n,), complex)
    for i in range(p.shape[-1]):
        out[..., i:i + q.shape[-1]] += p[..., i:i + 1] * q
    return out


def _padd(p, q):
    n = max(p.shape[-1], q.shape[-1])
    shp = np.broadcast_shapes(p.shape[:-1], q.shape[:-1]) + (n,)
    out = np.zeros(shp, complex)
    out[..., :p.shape[-1]] += p
    out[..., :q.shape[-1]] += q
    return out


def poly_coeffs_product(w, r, e):
    """Lens polynomial of m point masses at r (complex) with masses e, highest-first.

    From w = z - sum_j e_j/(conj(z) - conj(r_j)): with D = prod_j (z - r_j),
    N = sum_j e_j prod_{k!=j} (z - r_k) and M_j = (conj(w) - conj(r_j)) D + N,
        P(z) = (z - w) prod_j M_j - D sum_j e_j prod_{k!=j} M_k,
    a polynomial of degree m^2 + 1 whose leading coefficient is prod_j (conj(w) - conj(r_j)).
    For m = 2 this is term-for-term the polynomial of point_source.py:35-86 and for
    m = 3 the one of point_source.py:107-1479.
    """
    w = np.asarray(w, complex)[..., None]
    wb = np.conj(w)
    m = len(r)
    one = np.ones(w.shape[:-1] + (1,), complex)
    lin = [np.concatenate([-r[j] * one, one], -1) for j in range(m)]  # (z - r_j)
    D = one
    for j in range(m):
        D = _pmul(D, lin[j])
    Nn = np.zeros(w.shape[:-1] + (1,), complex)
    for j in range(m):
        t = e[j] * one
        for k in range(m):
            if k != j:
                t = _pmul(t, lin[k])
        Nn = _padd(Nn, t)
    M = [_padd((wb - np.conj(r[j])) * D, Nn) for j in range(m)]
    prodM = one
    for j in range(m):
        prodM = _pmul(prodM, M[j])
    zw = np.concatenate([-w, one], -1)
    P = _pmul(zw, prodM)
    S = np.zeros_like(P[..., :1])
    for j in range(m):
        t = e[j] * one
        for k in range(m):
            if k != j:
                t = _pmul(t, M[k])
        S = _padd(S, t)
    P = _padd(P, -_pmul(D, S))
    return P[..., ::-1]


def poly_coeffs_triple(w, a, r3, e1, e2):
    r, e = lens_positions_masses(3, a=a, e1=e1, e2=e2, r3=r3)
    return poly_coeffs_product(w, r, e)


def poly_coeffs(w, nlenses, **lp):
    if nlenses == 2:
        return poly_coeffs_binary(w, lp["a"], lp["e1"])
    if nlenses == 3:
        return poly_coeffs_triple(w, lp["a"], lp["r3"], lp["e1"], lp["e2"])
    raise ValueError


# ----------------------------------------------------------------------------
# lens equation, Jacobian determinant, images (point_source.py:1536-1709)
# ----------------------------------------------------------------------------
def lens_eq(z, nlenses=2, **lp):
    zb = np.conj(z)
    if nlenses == 1:
        return z - 1 / zb
    if nlenses == 2:
        a, e1 = lp["a"], lp["e1"]
        return z - e1 / (zb - a) - (1.0 - e1) / (zb + a)
    a, r3, e1, e2 = lp["a"], lp["r3"], lp["e1"], lp["e2"]
    return z - e1 / (zb - a) - e2 / (zb + a) - (1.0 - e1 - e2) / (zb - np.conj(r3))


def lens_eq_det_jac(z, nlenses=2, reference_single_lens=False, **lp):
    """det J = 1 - |sum_j e_j/(conj z - conj r_j)^2|^2 (point_source.py:1558-1580).

    For one lens the reference returns 1 - 1/|z|^2 (point_source.py:1562-1563), which
    is not the Jacobian determinant 1 - 1/|z|^4 of w = z - 1/conj(z); the sign agrees,
    the magnification does not (the Paczynski closed form pins the correct one).
    ``reference_single_lens=True`` restates the reference expression."""
    zb = np.conj(z)
    if nlenses == 1:
        if reference_single_lens:
            return 1.0 - 1.0 / np.abs(zb**2)
        return 1.0 - 1.0 / np.abs(zb**2) ** 2
    if nlenses == 2:
        a, e1 = lp["a"], lp["e1"]
        return 1.0 - np.abs(e1 / (zb - a) ** 2 + (1.0 - e1) / (zb + a) ** 2) ** 2
    a, r3, e1, e2 = lp["a"], lp["r3"], lp["e1"], lp["e2"]
    return 1.0 - np.abs(e1 / (zb - a) ** 2 + e2 / (zb + a) ** 2
                        + (1.0 - e1 - e2) / (zb - np.conj(r3)) ** 2) ** 2


def images_point_source(w, nlenses=2, itmax=2500, compensated=False, init_mode=0, perturb=0.0,
                        seed=0, polish=False, **lp):
    """(n, *w.shape) images and real-image mask (point_source.py:1652-1709).

    perturb > 0 multiplies every coefficient by (1 + perturb*(u + i v)), u, v uniform in
    [-1, 1] (seeded): the spread of results under perturb ~ 1e-16 measures how many
    digits the FP64 polynomial route can determine at all."""
    w = np.asarray(w, complex)
    if nlenses == 1:
        wa = w.real**2 + w.imag**2
        z1 = 0.5 * w * (1.0 + np.sqrt(1 + 4 / wa))
        z2 = 0.5 * w * (1.0 - np.sqrt(1 + 4 / wa))
        z = np.stack([z1, z2])
        return z, np.ones(z.shape, bool)
    c = poly_coeffs(w, nlenses, **lp)
    if perturb:
        g = np.random.default_rng(seed)
        c = c * (1 + perturb * (g.uniform(-1, 1, c.shape) + 1j * g.uniform(-1, 1, c.shape)))
    z = ea.poly_roots(c, itmax=itmax, compensated=compensated, init_mode=init_mode,
                      polish=polish)
    z = np.moveaxis(z, -1, 0)
    mask = np.abs(lens_eq(z, nlenses, **lp) - w) < 1e-6
    return z, mask


def images_point_source_sequential(w, nlenses=2, itmax=2500, compensated=False, init_mode=0,
                                   **lp):
    """(n, N) images and mask along the 1-D sequence w (point_source.py:1711-1759): the
    first position from the default initial estimates, each later one with the previous
    position's roots as initial roots (custom_init, ehrlich_aberth.h:89-94)."""
    w = np.asarray(w, complex).reshape(-1)
    if nlenses == 1:
        return images_point_source(w, nlenses)
    c = poly_coeffs(w, nlenses, **lp)              # (N, n+1) highest-first
    n = c.shape[-1] - 1
    z = np.zeros((w.size, n), complex)
    prev = None
    for k in range(w.size):
        ck = c[k:k + 1, ::-1].copy()               # ascending
        zk = ea.roots_asc(ck, itmax=itmax, compensated=compensated, roots_init=prev,
                          init_mode=init_mode)
        z[k] = zk.reshape(-1)
        prev = z[k:k + 1]
    z = z.T
    mask = np.abs(lens_eq(z, nlenses, **lp) - w) < 1e-6
    return z, mask


def mag_point_source(w, nlenses=2, itmax=2500, compensated=False, init_mode=0, perturb=0.0,
                     seed=0, perturb_inputs=False, **params):
    """Point-source magnification (point_source.py:1762-1830).

    perturb_inputs: multiply w and every user parameter by (1 + u 2^-52), u uniform in
    [-1, 1] (seeded) before evaluating -- the 1-ulp input uncertainty every FP64
    implementation starts from."""
    if perturb_inputs:
        g = np.random.default_rng(seed)
        w = np.asarray(w, complex)
        w = w * (1 + 2.0**-52 * (g.uniform(-1, 1, w.shape) + 1j * g.uniform(-1, 1, w.shape)))
        params = {k: v * (1 + 2.0**-52 * g.uniform(-1, 1, np.shape(v)))
                  for k, v in params.items()}
    lp, x_cm = lens_params(nlenses, **params)
    w = np.asarray(w, complex) + x_cm
    z, mask = images_point_source(w, nlenses, itmax=itmax, compensated=compensated,
                                  init_mode=init_mode, perturb=perturb, seed=seed, **lp)
    det = lens_eq_det_jac(z, nlenses, **lp)
    return ((1.0 / np.abs(det)) * mask).sum(axis=0)


def mag_point_source_c(w, nlenses=2, itmax=2500, nthreads=1, **params):
    """mag_point_source in compiled C (lens_oracle.c: ora_mag_point_source) with OpenMP
    over positions: the same steps, for the host CPU baseline.  Parameters broadcast
    against w."""
    import ctypes

    w = np.ascontiguousarray(np.asarray(w, complex).reshape(-1))
    names = ("s", "q") if nlenses == 2 else ("s", "q", "q3", "r3", "psi")
    prm = np.ascontiguousarray(np.stack([np.broadcast_to(np.asarray(params[k], float), w.shape)
                                         for k in names], -1))
    out = np.empty(w.shape, float)
    L = ea.lib()
    f = L.ora_mag_point_source
    f.restype = ctypes.c_int64
    f.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                  ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    f(nlenses, w.size, w.ctypes.data, prm.ctypes.data, prm.shape[1], out.ctypes.data, itmax,
      nthreads)
    return out


# ----------------------------------------------------------------------------
# gradient oracle: the reference's route (implicit root derivative of the
# polynomial, ehrlich_aberth_primitive.py:305-319) chained through the
# coefficients by forward-mode dual numbers with (w, conj(w)) as independent
# inputs, then through the filter-masked sum of 1/|det J|.
# ----------------------------------------------------------------------------
class Dual:
    """Complex value with a vector of complex tangents (forward mode)."""
    __slots__ = ("v", "d")

    def __init__(self, v, d):
        self.v, self.d = v, d

    @staticmethod
    def lift(x, nd):
        if isinstance(x, Dual):
            return x
        x = np.asarray(x, complex)
        return Dual(x, np.zeros(x.shape + (nd,), complex))

    def _nd(self):
        return self.d.shape[-1]

    def __add__(self, o):
        o = Dual.lift(o, self._nd())
        return Dual(self.v + o.v, self.d + o.d)
    __radd__ = __add__

    def __sub__(self, o):
        o = Dual.lift(o, self._nd())
        return Dual(self.v - o.v, self.d - o.d)

    def __rsub__(self, o):
        return Dual.lift(o, self._nd()) - self

    def __neg__(self):
        return Dual(-self.v, -self.d)

    def __mul__(self, o):
        o = Dual.lift(o, self._nd())
        return Dual(self.v * o.v, self.d * o.v[..., None] + o.d * self.v[..., None])
    __rmul__ = __mul__

    def __truediv__(self, o):
        o = Dual.lift(o, self._nd())
        v = self.v / o.v
        return Dual(v, (self.d - o.d * v[..., None]) / o.v[..., None])

    def __rtruediv__(self, o):
        return Dual.lift(o, self._nd()) / self

    def __pow__(self, k):
        assert isinstance(k, int) and k >= 1
        out = self
        for _ in range(k - 1):
            out = out * self
        return out


def mag_point_source_and_grad(w, nlenses=2, itmax=2500, **params):
    """Magnification and its gradient with respect to (Re w, Im w) and the lens
    parameters (s, q) for nlenses=2 or (s, q, q3, r3, psi) for nlenses=3.

    Follows the reference's autodiff route: tangents of the coefficients are
    pushed through the implicit root derivative dz = -(sum_k z^k dp_k)/p'(z)
    (ehrlich_aberth_primitive.py:313-319), and then through lens_eq_det_jac and
    the masked sum (point_source.py:1828-1830); the mask is piecewise constant.
    Returns mag (shape w.shape) and grad (w.shape + (nparams,)).
    """
    w = np.asarray(w, complex).ravel()
    names = ["s", "q"] if nlenses == 2 else ["s", "q", "q3", "r3", "psi"]
    nd = 2 + len(names)
    P = {k: Dual(np.full(w.shape, params[k], complex), np.zeros(w.shape + (nd,), complex))
         for k in names}
    for i, k in enumerate(names):
        P[k].d[..., 2 + i] = 1.0
    wd = Dual(w.copy(), np.zeros(w.shape + (nd,), complex))
    wd.d[..., 0] = 1.0
    wd.d[..., 1] = 1j
    a = 0.5 * P["s"]
    x_cm = a * (1 - P["q"]) / (1 + P["q"])
    ws = wd + x_cm
    wbs = Dual(np.conj(ws.v), np.conj(ws.d))
    if nlenses == 2:
        e1 = 1 / (1 + P["q"])
        rs = [a, -a]
        es = [e1, 1 - e1]
    else:
        e1 = P["q"] / (1 + P["q"] + P["q3"])
        e2 = P["q"] * e1
        ang = P["psi"].v
        # r3 * exp(i psi): d/dpsi = i r3 exp(i psi)
        eip = np.exp(1j * ang)
        r3 = Dual(P["r3"].v * eip, P["r3"].d * eip[..., None]
                  + (1j * P["r3"].v * eip)[..., None] * P["psi"].d)
        rs = [a, -a, r3]
        es = [e1, e2, 1 - e1 - e2]
    # coefficients by the product form, tracked with duals (ascending)
    rbs = [Dual(np.conj(r.v), np.conj(r.d)) for r in rs]
    m = len(rs)

    def pmul(p, q):
        out = [Dual.lift(0.0 * w, nd) for _ in range(len(p) + len(q) - 1)]
        for i in range(len(p)):
            for j in range(len(q)):
                out[i + j] = out[i + j] + p[i] * q[j]
        return out

    def padd(p, q):
        n = max(len(p), len(q))
        out = []
        for i in range(n):
            t = Dual.lift(0.0 * w, nd)
            if i < len(p):
                t = t + p[i]
            if i < len(q):
                t = t + q[i]
            out.append(t)
        return out

    one = Dual.lift(np.ones_like(w), nd)
    lin = [[-rs[j], one] for j in range(m)]
    D = [one]
    for j in range(m):
        D = pmul(D, lin[j])
    Nn = [Dual.lift(0.0 * w, nd)]
    for j in range(m):
        t = [es[j] * one]
        for k in range(m):
            if k != j:
                t = pmul(t, lin[k])
        Nn = padd(Nn, t)
    M = [padd([(wbs - rbs[j]) * d for d in D], Nn) for j in range(m)]
    prodM = [one]
    for j in range(m):
        prodM = pmul(prodM, M[j])
    Pz = pmul([-ws, one], prodM)
    S = [Dual.lift(0.0 * w, nd)]
    for j in range(m):
        t = [es[j] * one]
        for k in range(m):
            if k != j:
                t = pmul(t, M[k])
        S = padd(S, t)
    Pz = padd(Pz, [-(x) for x in pmul(D, S)])
    deg = len(Pz) - 1
    c_asc = np.stack([p.v for p in Pz], -1)
    dc_asc = np.stack([p.d for p in Pz], -2)  # (N, deg+1, nd)
    z = ea.roots_asc(c_asc, itmax=itmax)      # (N, deg)
    # p'(z), sum_k z^k dp_k
    k = np.arange(deg + 1)
    zk = z[..., None] ** k                     # (N, deg, deg+1)
    dpz = (k[1:] * c_asc[:, None, 1:] * z[..., None] ** (k[1:] - 1)).sum(-1)
    dz = -np.einsum("nik,nkd->nid", zk, dc_asc) / dpz[..., None]  # (N, deg, nd)
    # lens-frame quantities with duals: only z (from roots) and params matter
    zd = Dual(z, dz)
    zbd = Dual(np.conj(z), np.conj(dz))
    # broadcast lens params over roots
    def bc(x):
        x = Dual.lift(x, nd)
        return Dual(np.broadcast_to(x.v[..., None], z.shape),
                    np.broadcast_to(x.d[..., None, :], z.shape + (nd,)))
    rsb = [bc(r) for r in rs]
    esb = [bc(e) for e in es]
    wsb = bc(ws)
    lens = zd
    for r_, e_ in zip(rsb, esb):
        lens = lens - e_ / (zbd - Dual(np.conj(r_.v), np.conj(r_.d)))
    mask = np.abs(lens.v - wsb.v) < 1e-6
    kap = None
    for r_, e_ in zip(rsb, esb):
        t = e_ / ((zbd - Dual(np.conj(r_.v), np.conj(r_.d))) ** 2)
        kap = t if kap is None else kap + t
    # det = 1 - |kap|^2 ; mag_i = 1/|det|
    det_v = 1.0 - np.abs(kap.v) ** 2
    ddet = -2.0 * np.real(np.conj(kap.v)[..., None] * kap.d)
    mag_i = 1.0 / np.abs(det_v)
    dmag_i = -np.sign(det_v)[..., None] / det_v[..., None] ** 2 * ddet
    mag = (mag_i * mask).sum(-1)
    grad = np.real((dmag_i * mask[..., None]).sum(-2))
    return mag, grad


# ----------------------------------------------------------------------------
# critical curves and caustics (point_source.py:1482-1534, 1582-1649)
# ----------------------------------------------------------------------------
def poly_coeffs_critical(phi, nlenses, **lp):
    """Highest-first coefficients of the critical-curve polynomial
    sum_j e_j / (z - r_j)^2 = exp(-i phi), times prod_j (z - r_j)^2 (notebook cells 19/22)."""
    x = np.exp(-1j * np.asarray(phi, float))
    if nlenses == 2:
        a, e1 = lp["a"], lp["e1"]
        p = [x, 0 * x, -2 * a**2 * x - 1.0, (-4 * a * e1 + 2 * a) * np.ones_like(x),
             a**4 * x - a**2]
        return np.stack(p, -1)
    a, r3, e1, e2 = lp["a"], lp["r3"], lp["e1"], lp["e2"]
    p0 = x
    p1 = -2 * x * r3
    p2 = -2 * a**2 * x - 1 + x * r3**2
    p3 = 4 * a**2 * x * r3 - 2 * a * e1 + 2 * a * e2 + 2 * e1 * r3 + 2 * e2 * r3
    p4 = (a**4 * x - 3 * a**2 * e1 - 3 * a**2 * e2 + 2 * a**2 - 2 * a**2 * x * r3**2
          + 4 * a * e1 * r3 - 4 * a * e2 * r3 - e1 * r3**2 - e2 * r3**2)
    p5 = (-2 * a**4 * x * r3 + 2 * a**2 * e1 * r3 + 2 * a**2 * e2 * r3 - 2 * a * e1 * r3**2
          + 2 * a * e2 * r3**2)
    p6 = (a**4 * e1 + a**4 * e2 - a**4 + a**4 * x * r3**2 - a**2 * e1 * r3**2
          - a**2 * e2 * r3**2)
    return np.stack([p0, p1, p2, p3, p4, p5, p6 * np.ones_like(x)], -1)


def critical_and_caustic_curves(npts=200, nlenses=2, **params):
    """(critical curves, caustics), each (n_roots, npts), in the centre-of-mass frame."""
    from .extended import match_points

    phi = np.linspace(-np.pi, np.pi, npts)
    if nlenses == 1:
        return np.exp(-1j * phi), np.zeros(npts, complex)
    lp, x_cm = lens_params(nlenses, **params)
    c = poly_coeffs_critical(phi, nlenses, **lp)
    z = ea.poly_roots(c)
    carry = z[0]
    out = []
    for k in range(npts):
        idx = match_points(carry, z[k])
        carry = z[k][idx]
        out.append(carry)
    zc = np.array(out).T
    za = lens_eq(zc, nlenses, **lp)
    return zc - x_cm, za - x_cm
