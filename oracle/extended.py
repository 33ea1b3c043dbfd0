"""NumPy FP64 restatement of the extended-source contour pipeline. TEST INFRASTRUCTURE ONLY.

Follows (fbartolic/caustics, src/caustics/):
  _images_of_source_limb ....... extended_source.py:56-153
  _permute_images / match_points extended_source.py:34-53, utils.py:15-40
  _split_segment ............... extended_source.py:156-216
  _process_segments ............ extended_source.py:220-248
  _get_segments ................ extended_source.py:251-309
  _concatenate_segments ........ extended_source.py:312-314
  _get_segment_length .......... extended_source.py:317-321
  _connection_condition ........ extended_source.py:324-442
  _merge_two_segments .......... extended_source.py:445-492
  _merge_open_segments ......... extended_source.py:495-667
  _contours_from_*_segments .... extended_source.py:670-727
  mag_extended_source .......... extended_source.py:730-901
  _integrate_unif / _integrate_ld / _brightness_profile / trapz_zero_avoiding
                                 integrate.py:18-121, utils.py:43-100
  mag_extended_source_c ........ the same steps in compiled C++ (extended_oracle.cpp),
                                 OpenMP over sources: the host CPU baseline
  mag_disk_quadrature .......... not a restatement: the same magnification as a
                                 source-plane integral of mag_point_source, the
                                 truth the contour route is judged against
Arrays keep the reference's dense, zero-padded layout.

One deliberate difference: the reference keeps the identity of each image along the
limb by warm-starting every root solve from the previous limb point (sequential
scan, point_source.py:1711-1759) or from the left neighbour (refinement, :79-98,
with +-1e-6 threefry noise).  Here every limb point is solved from the default
initial estimates and its roots are put in the order of the neighbour they continue
by the reference's own greedy nearest-point assignment (match_points): the sequence
of points is chained from the first, refined points are matched to their left
neighbour.  The image sets are identical; only the row labels of coincident or
crossing images can differ, which the final _permute_images pass removes.  The
+-1e-9 perturbation of bitwise-duplicate roots (:139-148) is not applied (no
duplicate can arise from distinct limb points in FP64 except at exact symmetry).
"""
import numpy as np

from . import lens

LIMB_NITER = 10


# ---------------------------------------------------------------------------
# utils.py
# ---------------------------------------------------------------------------
def match_points(a, b):
    """Greedy: for each a_i in order, the nearest not-yet-used b (utils.py:15-40)."""
    vals = np.argsort(np.abs(b[None, :] - a[:, None]), axis=1, kind="stable")
    used = []
    for i in range(len(a)):
        for j in vals[i]:
            if j not in used:
                used.append(int(j))
                break
    return np.array(used)


def first_nonzero(x):
    return int(np.argmax(x != 0))


def first_zero(x):
    return int(np.argmax(x == 0))


def last_nonzero(x):
    if not np.any(x):
        return 0
    return len(x) - 1 - int(np.argmax(x[::-1] != 0))


def trapz(y, x):
    return np.sum(0.5 * (y[1:] + y[:-1]) * (x[1:] - x[:-1]))


def trapz_zero_avoiding(y, x, tidx):
    I = trapz(y, x)
    if tidx == len(x) - 1:
        return I
    return I - 0.5 * ((y[tidx] + y[tidx + 1]) * (x[tidx + 1] - x[tidx]))


def dslice2(x, start):
    """lax.dynamic_slice(x, (start,), (2,)): start clamped to [0, len-2]."""
    s = min(max(start, 0), len(x) - 2)
    return x[s:s + 2]


# ---------------------------------------------------------------------------
# limb images
# ---------------------------------------------------------------------------
def _solve_points(w0, rho, theta, nlenses, lp, perturb=0.0, seed=0, polish=False):
    z, mask = lens.images_point_source(rho * np.exp(1j * theta) + w0, nlenses, perturb=perturb,
                                       seed=seed, polish=polish, **lp)
    par = np.sign(lens.lens_eq_det_jac(z, nlenses, **lp))
    return z, mask, par


def chain_permute(z, mask, par):
    """_permute_images (extended_source.py:34-53)."""
    z, mask, par = z.copy(), mask.copy(), par.copy()
    carry = z[:, 0].copy()
    for k in range(z.shape[1]):
        idx = match_points(carry, z[:, k])
        z[:, k], mask[:, k], par[:, k] = z[idx, k], mask[idx, k], par[idx, k]
        carry = z[:, k].copy()
    return z, mask, par


def limb_thetas_initial(npts):
    npts_init = int(0.5 * npts)
    th = np.linspace(-np.pi, np.pi, npts_init - 1, endpoint=False)
    return np.append(th, np.pi - 1e-8)


def images_of_source_limb(w0, rho, nlenses, npts, lp, niter=LIMB_NITER, theta=None,
                          perturb=0.0, seed=0, polish=False):
    """theta given: images at exactly these limb angles (sorted, no refinement) with the
    same identity continuation -- the discretisation of an earlier call, frozen, so
    that finite differences see only the smooth dependence on the inputs.
    perturb > 0: every limb polynomial's coefficients are multiplied by
    (1 + perturb (u + i v)), u, v uniform in [-1, 1] (lens.images_point_source), so the
    spread of results over seeds measures what the FP64 route determines at all.
    polish=True ends every limb solve with the canonical compensated polish (ea.py):
    the roots become the doubles nearest the exact roots of each FP64 polynomial, the
    canonicalisation the product applies; the reference's plain stopping test leaves
    ill-conditioned roots (clusters near low-mass lenses, near caustics) further away,
    far enough to move an image across the 1e-6 filter."""
    if theta is not None:
        z, mask, par = _solve_points(w0, rho, np.asarray(theta), nlenses, lp, perturb, seed,
                                     polish)
        z, mask, par = chain_permute(z, mask, par)
        return z, mask, par, np.asarray(theta)
    theta = limb_thetas_initial(npts)
    z, mask, par = _solve_points(w0, rho, theta, nlenses, lp, perturb, seed, polish)
    z, mask, par = chain_permute(z, mask, par)
    n = int(int(0.5 * npts) / niter)
    for it in range(niter):
        dz = np.abs(z[:, 1:] - z[:, :-1])
        dz = np.where(mask[:, 1:] | mask[:, :-1], dz, 0.0)
        dmax = dz.max(axis=0)
        idcs = np.argsort(dmax, kind="stable")[::-1][:n]
        th_new = 0.5 * (theta[idcs] + theta[idcs + 1])
        zn, mn, pn = _solve_points(w0, rho, th_new, nlenses, lp, perturb, seed + 1 + it,
                                   polish)
        for c, k in enumerate(idcs):
            perm = match_points(z[:, k], zn[:, c])
            zn[:, c], mn[:, c], pn[:, c] = zn[perm, c], mn[perm, c], pn[perm, c]
        theta = np.insert(theta, idcs + 1, th_new)
        z = np.insert(z, idcs + 1, zn, axis=1)
        mask = np.insert(mask, idcs + 1, mn, axis=1)
        par = np.insert(par, idcs + 1, pn, axis=1)
    z, mask, par = chain_permute(z, mask, par)
    return z, mask, par, theta


# ---------------------------------------------------------------------------
# segments
# ---------------------------------------------------------------------------
def split_segment(seg, n_parts=5):
    """extended_source.py:156-216; seg = (3, N) rows z, parity, mask."""
    z, zp, zm = seg
    N = len(z)
    zd = z[1:] - z[:-1]
    ddist = zd.real ** 2 + zd.imag ** 2 > 0.1 ** 2
    dpar = zp[1:].real - zp[:-1].real
    dmask = zm[1:].real - zm[:-1].real
    cp = ddist | (dpar != 0) | (dmask != 0)
    cp = np.concatenate([[zm[0] != 0], cp, [zm[-1] != 0]])
    dm = np.concatenate([[0.0 if zm[0] == 0 else 1.0], dmask, [0.0 if zm[-1] == 0 else -1.0]])
    starts = np.nonzero(cp & (dm >= 0))[0]
    ends = np.nonzero(cp & (dm <= 0))[0]
    size = 2 * n_parts
    starts = np.concatenate([starts, np.zeros(size, int)])[:size]
    ends = np.concatenate([ends, np.zeros(size, int)])[:size]
    n = np.arange(N)
    out = np.zeros((size,) + seg.shape, complex)
    for k in range(size):
        l, r = starts[k], ends[k]
        if l == 0 and r == 0:
            continue
        out[k] = seg * ((n >= l) & (n < r))
    return out


def process_segments(segments, nr_of_segments):
    segs = np.concatenate([split_segment(s) for s in segments])
    small = np.sum(segs[:, 0] != 0, axis=1) < 2
    segs = segs * (~small)[:, None, None]
    nonempty = np.any(np.abs(segs[:, 0, :]) > 0.0, axis=1)
    order = np.argsort(nonempty, kind="stable")[::-1]
    segs = segs[order][:nr_of_segments]
    out = np.zeros_like(segs)
    for i, s in enumerate(segs):
        h = first_nonzero(np.abs(s[0]))
        out[i] = np.roll(s, -h, axis=-1)
    return out


def get_segments(z, mask, par, nlenses):
    n_images = nlenses ** 2 + 1
    nr = 3 * n_images
    zz = z * mask
    pp = par * mask
    segs = np.moveaxis(np.stack([zz, pp, mask.astype(complex)]), 0, 1)
    closed = (np.abs(zz[:, 0] - zz[:, -1]) < 1e-5) & np.all(mask, axis=1)
    seg_c = segs * closed[:, None, None]
    seg_o = segs * (~closed)[:, None, None]
    all_closed = bool(np.all(closed))
    if all_closed:
        seg_o = np.concatenate([seg_o, np.zeros((nr - seg_o.shape[0],) + seg_o.shape[1:])])
    else:
        seg_o = process_segments(seg_o, nr)
    return seg_c[:, :2], seg_o[:, :2], all_closed


def concatenate_segments(a, b):
    L = first_zero(np.abs(a[0]))
    return a + np.roll(b, L, axis=-1)


def segment_length(seg, tidx):
    d = np.diff(seg[0])
    if tidx < len(d):   # jnp .at[].set drops an out-of-range index
        d[tidx] = 0.0
    return np.abs(d).sum()


def _head(seg, t):
    x = seg[0]
    cond = (np.abs(x[1] - x[0]) > 1e-5) | (t <= 1)
    line = x[:2] if cond else x[1:3]
    return line[::-1]


def _tail(seg, t):
    x = seg[0]
    cond = (np.abs(x[t] - x[t - 1]) > 1e-5) | (t <= 1)
    return dslice2(x, t - 1) if cond else dslice2(x, t - 2)


def connection_condition(s1, s2, t1, t2, ctype, min_dist=1e-5, max_dist=1e-1, max_ang=60.0):
    same = s1[1, 0].real * s2[1, 0].real > 0.0
    cpar = [same, same, not same, not same][ctype]
    if ctype == 0:
        l1, l2 = _tail(s1, t1), _head(s2, t2)
    elif ctype == 1:
        l1, l2 = _head(s1, t1), _tail(s2, t2)
    elif ctype == 2:
        l1, l2 = _head(s1, t1), _head(s2, t2)
    else:
        l1, l2 = _tail(s1, t1), _tail(s2, t2)
    dist = np.abs(l1[1] - l2[1])
    c1 = dist < max_dist
    with np.errstate(invalid="ignore", divide="ignore"):
        v1 = (l1[1] - l1[0]) / np.abs(l1[1] - l1[0])
        v2 = (l2[1] - l2[0]) / np.abs(l2[1] - l2[0])
        alpha = np.arccos(v1.real * v2.real + v1.imag * v2.imag)
    c2 = (180.0 - np.rad2deg(alpha)) < max_ang
    c3 = np.abs(l1[1] - l2[1]) < np.abs(l1[0] - l2[0])
    return bool(cpar and ((c1 and c2 and c3) or dist < min_dist))


def merge_two(s1, s2, t1, t2, ctype):
    if ctype in (2, 3):
        s2 = s2[:, ::-1]
        s2 = np.roll(s2, -(s2.shape[-1] - t2 - 1), axis=-1)
        s2 = s2.copy()
        s2[1] = -1 * s2[1]
    if ctype == 0 or ctype == 3:
        m = concatenate_segments(s1, s2)
    else:
        m = concatenate_segments(s2, s1)
    return m, t1 + t2 + 1


def merge_open_segments(segments, max_nr_of_contours=3, max_nr_of_segments_in_contour=20):
    tails = np.array([last_nonzero(s[0].real) for s in segments])
    N = segments.shape[-1]
    segments = np.concatenate([segments, np.zeros(segments.shape[:2] + (3 * N,), complex)], -1)
    merged = []
    for _ in range(max_nr_of_contours):
        lengths = np.array([segment_length(s, t) for s, t in zip(segments, tails)])
        key = np.where(lengths != 0, lengths, np.nan)
        order = np.argsort(key, kind="stable")
        segments, tails = segments[order], tails[order]
        act, tact = segments[0].copy(), int(tails[0])
        rest, trest = segments[1:].copy(), tails[1:].copy()
        for _ in range(max_nr_of_segments_in_contour):
            if not np.any(rest[:, 0, 0]):
                continue
            d_th = np.abs(act[0, tact] - rest[:, 0, 0])
            d_ht = np.array([np.abs(act[0, 0] - s[0, t]) for s, t in zip(rest, trest)])
            d_hh = np.abs(act[0, 0] - rest[:, 0, 0])
            d_tt = np.array([np.abs(act[0, tact] - s[0, t]) for s, t in zip(rest, trest)])
            dist = np.stack([d_th, d_ht, d_hh, d_tt])
            flat = np.argsort(dist.reshape(-1), kind="stable")[:4]
            cands = [np.unravel_index(f, dist.shape) for f in flat]
            ok = [connection_condition(act, rest[i], tact, int(trest[i]), int(ct))
                  for ct, i in cands]
            if any(ok):
                b = ok.index(True)
                ct, i = cands[b]
                act, tact = merge_two(act, rest[i], tact, int(trest[i]), int(ct))
                rest[i] = 0
        merged.append(act)
        segments, tails = rest, trest
        max_nr_of_segments_in_contour -= 2
    return np.stack(merged)


def contours_from_open_segments(segments):
    m = merge_open_segments(segments)
    cp = m[:, 1, 0].real
    c = m[:, 0].copy()
    for k in range(c.shape[0]):
        t = last_nonzero(c[k].real)
        if t + 1 < c.shape[1]:
            c[k, t + 1] = c[k, 0]
    return c, cp


def contours_from_closed_segments(segments):
    cp = segments[:, 1, 0].real
    c = segments[:, 0]
    return np.concatenate([c, c[:, :1]], axis=1), cp


# ---------------------------------------------------------------------------
# integrals
# ---------------------------------------------------------------------------
def integrate_unif(z, tidx):
    return (trapz_zero_avoiding(0.5 * z.real, z.imag, tidx)
            + trapz_zero_avoiding(-0.5 * z.imag, z.real, tidx))


def brightness(z, rho, w0, u1, nlenses, lp):
    w = lens.lens_eq(z, nlenses, **lp)
    r = np.abs(w - w0) / rho
    with np.errstate(invalid="ignore", divide="ignore"):
        Br = np.where(r <= 1.0, 1 + np.sqrt(np.where(1 - r**2 > 0, 1 - r**2, 0.0)),
                      1 - np.sqrt(np.where(1 - 1.0 / r**2 > 0, 1 - 1.0 / r**2, 0.0)))
    return 3.0 / (3.0 - u1) * (u1 * Br + 1.0 - 2.0 * u1)


def _gl(f, a, b, n):
    x, wts = np.polynomial.legendre.leggauss(n)
    pts = 0.5 * (b - a) * x[:, None] + 0.5 * (b + a)
    return np.sum(0.5 * (b - a) * f(pts) * wts[:, None], axis=0)


def _split_point(a, b, rho):
    d = np.abs(b - a)
    sp = np.where(b > a, b - 2 * rho, b + 2 * rho)
    return np.where(0.5 * d <= 2 * rho, a + 0.5 * d, sp)


def integrate_ld(z, tidx, w0, rho, u1, nlenses, lp, npts=100):
    z0 = z.sum() / (tidx + 1)
    n1 = int(npts / 2)
    n2 = npts - n1
    x, y = z.real, z.imag
    a = z0.imag * np.ones_like(x)
    sp = _split_point(a, y, rho)
    fP = lambda yy: brightness(x + 1j * yy, rho, w0, u1, nlenses, lp)
    P = -0.5 * (_gl(fP, a, sp, n1) + _gl(fP, sp, y, n2))
    a = z0.real * np.ones_like(y)
    sp = _split_point(a, x, rho)
    fQ = lambda xx: brightness(xx + 1j * y, rho, w0, u1, nlenses, lp)
    Q = 0.5 * (_gl(fQ, a, sp, n1) + _gl(fQ, sp, x, n2))
    return trapz_zero_avoiding(P, x, tidx) + trapz_zero_avoiding(Q, y, tidx)


# ---------------------------------------------------------------------------
# magnification
# ---------------------------------------------------------------------------
def mag_from_limb(z, mask, par, w0, rho, nlenses, lp, limb_darkening=False, u1=0.0,
                  npts_ld=100):
    """Segments, stitching and integrals (extended_source.py:156-901 after
    _images_of_source_limb) for given limb images: z, mask, par (n, N) in limb order,
    rows continued along the limb; w0 in the lens frame."""
    if limb_darkening:
        integ = lambda c, t: integrate_ld(c, t, w0, rho, u1, nlenses, lp, npts_ld)
    else:
        integ = integrate_unif
    if nlenses == 1:
        c, cp = contours_from_closed_segments(np.moveaxis(np.stack([z, par]), 0, 1))
        I = np.array([integ(c[k], z.shape[1] - 1) for k in range(2)])
        return np.abs(np.sum(I * cp)) / (np.pi * rho**2)
    seg_c, seg_o, all_closed = get_segments(z, mask, par, nlenses)
    c1, p1 = contours_from_closed_segments(seg_c)
    I1 = np.array([integ(c1[k], c1.shape[1] - 1) for k in range(c1.shape[0])]) * p1
    I2 = np.zeros(3)
    if not all_closed:
        c2, p2 = contours_from_open_segments(seg_o)
        I2 = np.array([integ(c2[k], last_nonzero(c2[k].real)) for k in range(3)]) * p2
    return np.abs(I1.sum() + I2.sum()) / (np.pi * rho**2)


def mag_extended_source(w0, rho, nlenses=2, npts_limb=150, limb_darkening=False, u1=0.0,
                        npts_ld=100, return_parts=False, theta=None, perturb=0.0, seed=0,
                        polish=False, **params):
    lp, x_cm = lens.lens_params(nlenses, **params)
    w0 = complex(w0) + x_cm
    z, mask, par, theta = images_of_source_limb(w0, rho, nlenses, npts_limb, lp, theta=theta,
                                                perturb=perturb, seed=seed, polish=polish)
    if not return_parts:
        return mag_from_limb(z, mask, par, w0, rho, nlenses, lp, limb_darkening, u1, npts_ld)
    if limb_darkening:
        integ = lambda c, t: integrate_ld(c, t, w0, rho, u1, nlenses, lp, npts_ld)
    else:
        integ = integrate_unif
    if nlenses == 1:
        c, cp = contours_from_closed_segments(np.moveaxis(np.stack([z, par]), 0, 1))
        I = np.array([integ(c[k], z.shape[1] - 1) for k in range(2)])
        return np.abs(np.sum(I * cp)) / (np.pi * rho**2)
    seg_c, seg_o, all_closed = get_segments(z, mask, par, nlenses)
    c1, p1 = contours_from_closed_segments(seg_c)
    I1 = np.array([integ(c1[k], c1.shape[1] - 1) for k in range(c1.shape[0])]) * p1
    I2 = np.zeros(3)
    if not all_closed:
        c2, p2 = contours_from_open_segments(seg_o)
        I2 = np.array([integ(c2[k], last_nonzero(c2[k].real)) for k in range(3)]) * p2
    mag = np.abs(I1.sum() + I2.sum()) / (np.pi * rho**2)
    if return_parts:
        return mag, dict(z=z, mask=mask, par=par, theta=theta, all_closed=all_closed)
    return mag


def mag_extended_source_c(w0, rho, nlenses=2, npts_limb=150, limb_darkening=False, u1=0.0,
                          npts_ld=100, nthreads=1, polish=False, **params):
    """mag_extended_source in compiled C++ (extended_oracle.cpp: ora_mag_extended_source)
    with OpenMP over sources: the same steps, for the host CPU baseline.  w0, rho,
    npts_limb, u1 and the lens parameters broadcast against each other; polish as in
    images_of_source_limb.  Returns the magnifications and the number of sources with an
    unconverged limb solve."""
    import ctypes

    from . import ea

    names = ("s", "q") if nlenses == 2 else ("s", "q", "q3", "r3", "psi")
    arrs = np.broadcast_arrays(np.asarray(w0, complex), np.asarray(rho, float),
                               np.asarray(npts_limb), np.asarray(u1, float),
                               *[np.asarray(params[k], float) for k in names])
    shape = arrs[0].shape
    w = np.ascontiguousarray(arrs[0].reshape(-1))
    r = np.ascontiguousarray(arrs[1].reshape(-1))
    npts = np.ascontiguousarray(arrs[2].reshape(-1), dtype=np.int32)
    u = np.ascontiguousarray(arrs[3].reshape(-1))
    prm = np.ascontiguousarray(np.stack([a.reshape(-1) for a in arrs[4:]], -1))
    n1 = int(npts_ld / 2)
    x1, w1 = np.polynomial.legendre.leggauss(n1)
    x2, w2 = np.polynomial.legendre.leggauss(npts_ld - n1)
    out = np.empty(w.shape, float)
    f = ea.lib().ora_mag_extended_source
    f.restype = ctypes.c_int64
    vp, i32 = ctypes.c_void_p, ctypes.c_int
    f.argtypes = [i32, ctypes.c_int64, vp, vp, vp, ctypes.c_int64, vp, i32, vp, vp, vp, i32, vp,
                  vp, i32, i32, vp, i32]
    bad = f(nlenses, w.size, w.ctypes.data, r.ctypes.data, prm.ctypes.data, prm.shape[1],
            npts.ctypes.data, int(bool(limb_darkening)), u.ctypes.data, x1.ctypes.data,
            w1.ctypes.data, n1, x2.ctypes.data, w2.ctypes.data, npts_ld - n1, int(bool(polish)),
            out.ctypes.data, nthreads)
    if bad < 0:
        raise ValueError("nlenses must be 2 or 3")
    return out.reshape(shape), int(bad)


def mag_disk_quadrature(w0, rho, nlenses=2, u1=0.0, nr=100, **params):
    """Independent route to the magnification that mag_extended_source approximates: the
    brightness-weighted mean of the point-source magnification (lens.mag_point_source,
    point_source.py:1762-1830) over the source disk, by tensor Gauss-Legendre quadrature
    in polar coordinates.  The radius is r = rho sin(t), t in [0, pi/2], which makes the
    linear limb-darkening profile (integrate.py:18-44, normalised to unit flux)
        I(r) = 3/(3 - u1) (1 - u1 (1 - cos t))
    and the area element r dr = rho^2 sin t cos t dt smooth; nr nodes in t and 2 nr
    equally spaced angles.  Converges fast where the disk holds no caustic; near a
    caustic, compare two resolutions."""
    nt = 2 * nr
    x, wx = np.polynomial.legendre.leggauss(nr)
    t = 0.25 * np.pi * (x + 1.0)
    wt = 0.25 * np.pi * wx
    r = rho * np.sin(t)
    area = rho * rho * np.sin(t) * np.cos(t) * wt
    prof = 3.0 / (3.0 - u1) * (1.0 - u1 * (1.0 - np.cos(t)))
    ang = 2.0 * np.pi * (np.arange(nt) + 0.5) / nt
    pts = complex(w0) + r[:, None] * np.exp(1j * ang)[None, :]
    A = lens.mag_point_source(pts.reshape(-1), nlenses, **params).reshape(pts.shape)
    return float((A * (area * prof)[:, None]).sum() * (2.0 * np.pi / nt) / (np.pi * rho * rho))
