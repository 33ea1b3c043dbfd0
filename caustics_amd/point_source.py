"""Point-source lensing on the GPU (reference: src/caustics/point_source.py).

``mag_point_source(w, nlenses=2, roots_itmax=2500, roots_compensated=False, **params)``
(:1762-1830) runs in one fused HIP kernel per evaluation: parameter map and x_cm
shift, FP64 coefficients, all-roots solve, lens-equation filter (1e-6), sum of
1/|det J|.  Its gradient (torch.autograd) comes from the fused VJP kernel, which
re-solves the roots and pushes implicit-derivative tangents of w and of every lens
parameter through det J.

``lens_eq`` and ``lens_eq_det_jac`` (:1536-1580) take lens-frame parameters
(a, e1[, e2, r3]) exactly as in the reference and are elementwise tensor expressions.

``images_point_source`` and ``images_point_source_sequential`` (:1652-1759, the
reference's private ``_images_point_source`` and ``_images_point_source_sequential``)
return the images and the real-image mask for lens-frame source positions and
lens-frame parameters; the sequential form warm-starts each position from the previous
one's roots, so that row j follows one image along the sequence.  They return no
gradients (``mag_point_source`` and ``ehrlich_aberth`` do).
"""
import numpy as np
import torch

from . import _lib

__all__ = [
    "mag_point_source",
    "images_point_source",
    "images_point_source_sequential",
    "lens_eq",
    "lens_eq_det_jac",
    "critical_and_caustic_curves",
    "PARAM_NAMES",
]

PARAM_NAMES = {1: (), 2: ("s", "q"), 3: ("s", "q", "q3", "r3", "psi")}


def _mode_id(roots_compensated, mode):
    if mode is None:
        mode = "fp64_comp" if roots_compensated else "fp64"
    return _lib.MODES[mode]


def _tensor(x, like=None):
    if isinstance(x, torch.Tensor):
        return x
    return torch.as_tensor(np.asarray(x), device="cuda")


def _prepare(w, nlenses, params):
    w = _tensor(w)
    if not w.is_complex():
        w = w.to(torch.float64)
    w = w.to(device="cuda", dtype=torch.complex128)
    names = PARAM_NAMES[nlenses]
    missing = [n for n in names if n not in params]
    if missing:
        raise ValueError(f"nlenses={nlenses} needs parameters {names}; missing {missing}")
    ps = [_tensor(params[n]).to(device="cuda", dtype=torch.float64) for n in names]
    return w, ps


class _MagPointSource(torch.autograd.Function):
    @staticmethod
    def forward(ctx, nlenses, itmax, mode, w, *ps):
        shape = torch.broadcast_shapes(w.shape, *[p.shape for p in ps]) if ps else w.shape
        wf = w.expand(shape).contiguous().reshape(-1)
        N = wf.numel()
        P = len(ps)
        if P and all(p.numel() == 1 for p in ps):
            prm = torch.stack([p.reshape(()) for p in ps]).contiguous()
            pstride = 0
        elif P:
            prm = torch.stack([p.expand(shape).reshape(-1) for p in ps], -1).contiguous()
            pstride = P
        else:
            prm = None
            pstride = 0
        mag = torch.empty(N, dtype=torch.float64, device=w.device)
        status = torch.empty(N, dtype=torch.uint8, device=w.device)
        rc = _lib.lib().caus_mag_point_source(_lib.stream_ptr(), nlenses, _lib.ptr(wf), N,
                                              _lib.ptr(prm), pstride, _lib.ptr(mag),
                                              _lib.ptr(status), itmax, mode)
        _lib.check(rc, "caus_mag_point_source")
        ctx.save_for_backward(wf, *( [prm] if prm is not None else []))
        ctx.meta = (nlenses, itmax, mode, shape, pstride, w.shape, [p.shape for p in ps])
        return mag.reshape(shape)

    @staticmethod
    def backward(ctx, g):
        nlenses, itmax, mode, shape, pstride, wshape, pshapes = ctx.meta
        saved = ctx.saved_tensors
        wf = saved[0]
        prm = saved[1] if len(saved) > 1 else None
        N = wf.numel()
        P = len(pshapes)
        gf = g.expand(shape).contiguous().reshape(-1).to(torch.float64)
        mag = torch.empty(N, dtype=torch.float64, device=wf.device)
        g_w = torch.empty(N, dtype=torch.complex128, device=wf.device)
        g_p = torch.empty((N, max(P, 1)), dtype=torch.float64, device=wf.device)
        rc = _lib.lib().caus_mag_point_source_vjp(
            _lib.stream_ptr(), nlenses, _lib.ptr(wf), N, _lib.ptr(prm), pstride, _lib.ptr(gf),
            _lib.ptr(mag), _lib.ptr(g_w), _lib.ptr(g_p), None, itmax, mode)
        _lib.check(rc, "caus_mag_point_source_vjp")
        gw = g_w.reshape(shape).sum_to_size(wshape) if ctx.needs_input_grad[3] else None
        gps = []
        for k, ps_shape in enumerate(pshapes):
            if ctx.needs_input_grad[4 + k]:
                gps.append(g_p[:, k].reshape(shape).sum_to_size(ps_shape))
            else:
                gps.append(None)
        return (None, None, None, gw, *gps)


def mag_point_source(w, nlenses=2, roots_itmax=2500, roots_compensated=False, mode=None,
                     **params):
    """Point-source magnification at source positions ``w`` (complex), in the frame of
    the reference: origin at the centre of mass of lenses 1-2, which lie on the real
    axis.  Parameters: ``s, q`` (nlenses=2) or ``s, q, q3, r3, psi`` (nlenses=3); each
    may be a float or a tensor broadcastable against ``w``.  Differentiable in ``w``
    and every parameter.  ``mode``: "fp64" (default), "fp64_comp" (= roots_compensated)
    or "cf32"."""
    if nlenses not in (1, 2, 3):
        raise ValueError("`nlenses` has to be set to be <= 3.")
    w, ps = _prepare(w, nlenses, params)
    return _MagPointSource.apply(nlenses, int(roots_itmax), _mode_id(roots_compensated, mode),
                                 w, *ps)


_DEGREE = {1: 2, 2: 5, 3: 10}


def _as64(v):
    """v on the device in 64-bit precision (float64, or complex128 when complex);
    torch.as_tensor alone makes python floats and lists float32."""
    if torch.is_tensor(v):
        return v.to(device="cuda", dtype=torch.complex128 if v.is_complex() else torch.float64)
    a = np.asarray(v)
    return torch.as_tensor(a.astype(np.complex128 if np.iscomplexobj(a) else np.float64),
                           device="cuda")


def _lens_rows(nlenses, shape, params):
    """Lens-frame parameter rows (a, e1) or (a, e1, e2, Re r3, Im r3) broadcast to
    `shape` (the source positions' batch shape, sequences flattened)."""
    if nlenses == 1:
        return None
    names = ("a", "e1") if nlenses == 2 else ("a", "e1", "e2", "r3")
    cols = []
    for n in names:
        v = params[n]
        if n == "r3":
            # python floats and complexes straight to 64-bit (as_tensor would make float32)
            v = torch.as_tensor(v, dtype=torch.complex128, device="cuda")
            cols += [v.real, v.imag]
        else:
            v = (v.to(device="cuda", dtype=torch.float64) if torch.is_tensor(v)
                 else torch.as_tensor(v, dtype=torch.float64, device="cuda"))
            cols.append(v)
    cols = [c.expand(shape).reshape(-1) for c in cols]
    return torch.stack(cols, -1).contiguous()


def _images(w, nlenses, itmax, mode, z_init, sequential, params):
    L = _lib.lib()
    n = _DEGREE[nlenses]
    B, N = w.shape
    lens = _lens_rows(nlenses, (B,) if sequential else (B, N), params)
    if lens is not None and not sequential:
        lens = lens.reshape(B * N, -1)
    z = torch.empty((B, N, n), dtype=torch.complex128, device="cuda")
    mask = torch.empty((B, N, n), dtype=torch.uint8, device="cuda")
    status = torch.empty((B, N), dtype=torch.uint8, device="cuda")
    zi = None
    if z_init is not None:
        zi = (z_init.to(device="cuda", dtype=torch.complex128) if torch.is_tensor(z_init)
              else torch.as_tensor(np.asarray(z_init, np.complex128), device="cuda"))
        zi = zi.reshape(n, B, N).permute(1, 2, 0).contiguous()
    # independent positions: one lens row per position, addressed as B*N sequences of 1
    Bc, Nc = (B, N) if sequential else (B * N, 1)
    rc = L.caus_images_point_source(_lib.stream_ptr(), nlenses, _lib.ptr(w), Bc, Nc,
                                    _lib.ptr(lens), 0 if lens is None else lens.shape[1],
                                    _lib.ptr(zi), _lib.ptr(z), _lib.ptr(mask), _lib.ptr(status),
                                    int(itmax), mode, 1 if sequential else 0)
    _lib.check(rc, "caus_images_point_source")
    return z, mask.bool(), status


def images_point_source(w, nlenses=2, roots_itmax=2500, roots_compensated=False,
                        custom_init=False, z_init=None, mode=None, **params):
    """Images (n, *w.shape) complex and the real-image mask (|lens_eq(z) - w| < 1e-6) of
    lens-frame source positions w; parameters a, e1 (nlenses=2) or a, e1, e2, r3
    (nlenses=3), floats or tensors broadcastable against w.  custom_init with z_init
    (n, *w.shape) starts the solver from those roots (point_source.py:1652-1709)."""
    if nlenses not in (1, 2, 3):
        raise ValueError("`nlenses` has to be set to be <= 3.")
    w = (w.to(device="cuda", dtype=torch.complex128) if torch.is_tensor(w)
         else torch.as_tensor(np.asarray(w, np.complex128), device="cuda"))
    shape = w.shape
    wf = w.reshape(1, -1).contiguous()
    prm = {k: (_as64(v).expand(shape).reshape(1, -1) if torch.is_tensor(v) or np.ndim(v) else v)
           for k, v in params.items()}
    z, mask, _ = _images(wf, nlenses, roots_itmax, _mode_id(roots_compensated, mode),
                         z_init if custom_init else None, False, prm)
    n = _DEGREE[nlenses]
    return (z.reshape(*shape, n).movedim(-1, 0), mask.reshape(*shape, n).movedim(-1, 0))


def images_point_source_sequential(w, nlenses=2, roots_itmax=2500, roots_compensated=False,
                                   mode=None, **params):
    """Images along a sequence of lens-frame source positions w (N,) solved in order, each
    position started from the previous one's roots (point_source.py:1711-1759): returns
    z (n, N) and the mask (n, N).  A 2-D w (B, N) runs B sequences at once, with
    parameters of shape (B,) or scalars, and returns (B, n, N)."""
    if nlenses not in (1, 2, 3):
        raise ValueError("`nlenses` has to be set to be <= 3.")
    w = (w.to(device="cuda", dtype=torch.complex128) if torch.is_tensor(w)
         else torch.as_tensor(np.asarray(w, np.complex128), device="cuda"))
    one = w.dim() == 1
    w2 = (w.reshape(1, -1) if one else w).contiguous()
    B = w2.shape[0]
    prm = {k: (_as64(v).reshape(-1) if torch.is_tensor(v) or np.ndim(v) else v)
           for k, v in params.items()}
    z, mask, _ = _images(w2, nlenses, roots_itmax, _mode_id(roots_compensated, mode), None,
                         True, {k: (v if not torch.is_tensor(v) else v.expand(B)) for k, v in
                                prm.items()})
    z, mask = z.permute(0, 2, 1), mask.permute(0, 2, 1)   # (B, n, N)
    return (z[0], mask[0]) if one else (z, mask)


def _lens_terms(z, nlenses, params):
    zb = torch.conj(z)
    if nlenses == 1:
        return [(1.0, 0.0)], zb
    a, e1 = params["a"], params["e1"]
    if nlenses == 2:
        return [(e1, a), (1.0 - e1, -a)], zb
    e2, r3 = params["e2"], params["r3"]
    r3 = r3 if isinstance(r3, torch.Tensor) else torch.as_tensor(complex(r3), device=z.device,
                                                                 dtype=torch.complex128)
    return [(e1, a), (e2, -a), (1.0 - e1 - e2, r3)], zb


def lens_eq(z, nlenses=2, **params):
    """w = z - sum_j e_j / (conj z - conj r_j) with lens-frame parameters
    (point_source.py:1536-1556)."""
    z = _tensor(z).to(torch.complex128)
    terms, zb = _lens_terms(z, nlenses, params)
    out = z
    for e, r in terms:
        r_c = torch.conj(r) if isinstance(r, torch.Tensor) else r
        out = out - e / (zb - r_c)
    return out


def lens_eq_det_jac(z, nlenses=2, **params):
    """det J = 1 - |sum_j e_j / (conj z - conj r_j)^2|^2 (point_source.py:1558-1580)."""
    z = _tensor(z).to(torch.complex128)
    terms, zb = _lens_terms(z, nlenses, params)
    K = 0
    for e, r in terms:
        r_c = torch.conj(r) if isinstance(r, torch.Tensor) else r
        K = K + e / (zb - r_c) ** 2
    return 1.0 - torch.abs(K) ** 2


def _critical_coeffs(phi, nlenses, a, e1, e2=None, r3=None):
    """Highest-first coefficients of prod_j (z - r_j)^2 exp(-i phi) - sum_j e_j
    prod_{k!=j} (z - r_k)^2, the critical-curve polynomial (point_source.py:1482-1534)."""
    x = torch.exp(-1j * phi) * torch.ones_like(a)   # (B, npts)
    one = torch.ones_like(x)
    if nlenses == 2:
        return torch.stack([x, 0 * x, -2 * a**2 * x - 1.0, (-4 * a * e1 + 2 * a) * one,
                            a**4 * x - a**2], -1)
    p = [x, -2 * x * r3, -2 * a**2 * x - 1 + x * r3**2,
         4 * a**2 * x * r3 - 2 * a * e1 + 2 * a * e2 + 2 * e1 * r3 + 2 * e2 * r3,
         (a**4 * x - 3 * a**2 * e1 - 3 * a**2 * e2 + 2 * a**2 - 2 * a**2 * x * r3**2
          + 4 * a * e1 * r3 - 4 * a * e2 * r3 - e1 * r3**2 - e2 * r3**2),
         (-2 * a**4 * x * r3 + 2 * a**2 * e1 * r3 + 2 * a**2 * e2 * r3 - 2 * a * e1 * r3**2
          + 2 * a * e2 * r3**2),
         (a**4 * e1 + a**4 * e2 - a**4 + a**4 * x * r3**2 - a**2 * e1 * r3**2
          - a**2 * e2 * r3**2) * one]
    return torch.stack(p, -1)


def critical_and_caustic_curves(npts=200, nlenses=2, **params):
    """Critical curves and caustics, each (n_curves, npts) complex, in the frame of the
    reference API (point_source.py:1582-1649): roots of the critical-curve polynomial
    on a grid of phi (GPU solver), continued along phi with match_points (HIP kernel),
    mapped through the lens equation and shifted by x_cm.  Parameters may also be
    arrays of length B, giving (B, n_curves, npts)."""
    from .ehrlich_aberth_primitive import poly_roots

    phi = torch.linspace(-np.pi, np.pi, npts, dtype=torch.float64, device="cuda")
    if nlenses == 1:
        return torch.exp(-1j * phi), torch.zeros(npts, dtype=torch.complex128, device="cuda")
    names = PARAM_NAMES[nlenses]
    vals = [torch.as_tensor(np.asarray(params[k]) if not isinstance(params[k], torch.Tensor)
                            else params[k], device="cuda").to(torch.float64) for k in names]
    batch = vals[0].dim() > 0
    vals = [v.reshape(-1, 1) for v in vals]
    s, q = vals[0], vals[1]
    a = 0.5 * s
    x_cm = a * (1 - q) / (1 + q)
    if nlenses == 2:
        e1 = 1 / (1 + q)
        coeffs = _critical_coeffs(phi[None, :], 2, a, e1)
        lf = dict(a=a, e1=e1)
    else:
        q3, r3m, psi = vals[2], vals[3], vals[4]
        e1 = q / (1 + q + q3)
        e2 = q * e1
        r3 = r3m * torch.exp(1j * psi)
        coeffs = _critical_coeffs(phi[None, :], 3, a, e1, e2, r3)
        lf = dict(a=a, e1=e1, e2=e2, r3=r3)
    z = poly_roots(coeffs).contiguous()            # (B, npts, n)
    B, _, n = z.shape
    zc = torch.empty_like(z)
    rc = _lib.lib().caus_match_chain(_lib.stream_ptr(), _lib.ptr(z), _lib.ptr(zc), npts, n, B)
    _lib.check(rc, "caus_match_chain")
    zc = zc.transpose(1, 2)                        # (B, n, npts)
    lfb = {k: (v[:, :, None] if isinstance(v, torch.Tensor) else v) for k, v in lf.items()}
    za = lens_eq(zc, nlenses=nlenses, **lfb)
    zc = zc - x_cm[:, :, None]
    za = za - x_cm[:, :, None]
    if not batch:
        return zc[0], za[0]
    return zc, za
