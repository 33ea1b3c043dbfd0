"""Light curves on the GPU.

``mag_point_source_lightcurve``: point-source magnification of L parameter sets at T
epochs along rectilinear trajectories w(t) = e^{-i alpha}(u0 + i (t - t0)/tE)
(src/caustics/trajectory.py:152-155 with no parallax and psi -> alpha), one fused
HIP kernel per (light curve, epoch).  Its gradient with respect to every parameter
(t0, u0, tE, alpha and the lens parameters) comes from the fused forward+VJP kernel,
which re-solves the roots, forms per-epoch implicit-derivative tangents and reduces
them per light curve in a fixed order (deterministic).

``lightcurve_fwd_vjp`` exposes that fused kernel directly: one root solve gives the
magnifications and the vector-Jacobian product for a cotangent ``g_mag``.
"""
import numpy as np
import torch

from . import _lib
from .point_source import PARAM_NAMES, _mode_id

__all__ = ["mag_point_source_lightcurve", "lightcurve_fwd_vjp", "hexadecapole_gate",
           "hexadecapole_vjp", "mag", "mag_fwd_vjp", "TRAJ_NAMES"]

TRAJ_NAMES = ("t0", "u0", "tE", "alpha")


def _call(t, lcp, g_mag, nlenses, itmax, mode, want_status=False):
    L_ = _lib.lib()
    T = t.numel()
    L = lcp.shape[0]
    mag = torch.empty((L, T), dtype=torch.float64, device=t.device)
    status = torch.empty((L, T), dtype=torch.uint8, device=t.device) if want_status else None
    g_params = ws = None
    wsb = 0
    if g_mag is not None:
        g_params = torch.empty_like(lcp)
        wsb = L_.caus_lightcurve_workspace_bytes(nlenses, T, L)
        ws = torch.empty(max(wsb // 8, 1), dtype=torch.float64, device=t.device)
    rc = L_.caus_lightcurve_point_source(_lib.stream_ptr(), nlenses, _lib.ptr(t), T,
                                         _lib.ptr(lcp), L, _lib.ptr(g_mag), _lib.ptr(mag),
                                         _lib.ptr(g_params), _lib.ptr(ws), wsb, _lib.ptr(status),
                                         itmax, mode)
    _lib.check(rc, "caus_lightcurve_point_source")
    return mag, g_params, status


def lightcurve_fwd_vjp(t, lc_params, g_mag, nlenses=2, roots_itmax=2500, mode="fp64",
                       return_status=False):
    """Fused forward + VJP.  t (T,), lc_params (L, 4+P) rows (t0, u0, tE, alpha, lens
    params...), g_mag (L, T) -> (mag (L, T), g_params (L, 4+P)[, status (L, T)])."""
    t = torch.as_tensor(t, device="cuda", dtype=torch.float64).contiguous()
    lcp = torch.as_tensor(lc_params, device="cuda", dtype=torch.float64).contiguous()
    g = None if g_mag is None else torch.as_tensor(g_mag, device="cuda",
                                                   dtype=torch.float64).contiguous()
    mag, gp, st = _call(t, lcp, g, nlenses, int(roots_itmax), _lib.MODES[mode], return_status)
    return (mag, gp, st) if return_status else (mag, gp)


class _LightCurve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, nlenses, itmax, mode, t, lcp):
        mag, _, _ = _call(t, lcp, None, nlenses, itmax, mode)
        ctx.save_for_backward(t, lcp)
        ctx.meta = (nlenses, itmax, mode)
        return mag

    @staticmethod
    def backward(ctx, g):
        t, lcp = ctx.saved_tensors
        nlenses, itmax, mode = ctx.meta
        _, gp, _ = _call(t, lcp, g.contiguous().to(torch.float64), nlenses, itmax, mode)
        return None, None, None, None, gp


def mag_point_source_lightcurve(t, t0, u0, tE, alpha, nlenses=2, roots_itmax=2500,
                                roots_compensated=False, mode=None, **params):
    """Point-source light curves: returns (L, T) magnifications, L = broadcast length
    of the parameters (each a float or a (L,) tensor).  Differentiable in every
    parameter (not in t)."""
    names = TRAJ_NAMES + PARAM_NAMES[nlenses]
    vals = dict(t0=t0, u0=u0, tE=tE, alpha=alpha, **params)
    cols = [torch.as_tensor(np.asarray(vals[n]) if not isinstance(vals[n], torch.Tensor)
                            else vals[n], device="cuda").to(torch.float64) for n in names]
    L = max([c.numel() for c in cols])
    lcp = torch.stack([c.reshape(-1).expand(L) for c in cols], -1).contiguous()
    t = torch.as_tensor(t, device="cuda", dtype=torch.float64).contiguous().reshape(-1)
    return _LightCurve.apply(nlenses, int(roots_itmax), _mode_id(roots_compensated, mode), t,
                             lcp)


def hexadecapole_gate(w_points, rho, nlenses=2, u1=0.0, roots_itmax=2500, **params):
    """Per point: hexadecapole magnification, the size of its quadrupole + hexadecapole
    terms, and whether mag() accepts it (src/caustics/lightcurve.py:200-227;
    multipole.py:208-244).  Returns (mu_hex, delta_mu, use_hex) tensors."""
    from .extended_source import _col

    w = w_points if isinstance(w_points, torch.Tensor) else torch.as_tensor(np.asarray(w_points))
    shape = tuple(w.shape)
    w = w.to(device="cuda", dtype=torch.complex128).reshape(-1).contiguous()
    P = w.numel()
    rho_t = _col(rho, P)
    names = PARAM_NAMES[nlenses]
    prm = torch.stack([_col(params[k], P) for k in names], -1).contiguous() if names else None
    mu = torch.empty(P, dtype=torch.float64, device="cuda")
    dmu = torch.empty(P, dtype=torch.float64, device="cuda")
    use = torch.empty(P, dtype=torch.uint8, device="cuda")
    rc = _lib.lib().caus_hexadecapole_gate(_lib.stream_ptr(), nlenses, _lib.ptr(w), P,
                                           _lib.ptr(prm), len(names), _lib.ptr(rho_t), 1,
                                           float(u1), _lib.ptr(mu), _lib.ptr(dmu),
                                           _lib.ptr(use), None, int(roots_itmax))
    _lib.check(rc, "caus_hexadecapole_gate")
    return mu.reshape(shape), dmu.reshape(shape), use.reshape(shape).bool()


def hexadecapole_vjp(w_points, rho, g_mu, nlenses=2, u1=0.0, roots_itmax=2500, **params):
    """Hexadecapole magnification and g_mu * its derivative (HIP kernel): returns
    (mu_hex (P,), g_out (P, 4 + n_params)) with g_out columns Re w, Im w, rho, u1 and the
    lens parameters (reverse mode of src/caustics/multipole.py:208-244)."""
    from .extended_source import _col

    w = w_points if isinstance(w_points, torch.Tensor) else torch.as_tensor(np.asarray(w_points))
    w = w.to(device="cuda", dtype=torch.complex128).reshape(-1).contiguous()
    P = w.numel()
    names = PARAM_NAMES[nlenses]
    prm = torch.stack([_col(params[k], P) for k in names], -1).contiguous() if names else None
    return _hex_vjp(nlenses, w, _col(rho, P), float(u1), prm, _col(g_mu, P), roots_itmax)


def _hex_vjp(nlenses, w, rho_t, u1, prm, g, itmax):
    P = w.numel()
    NP = 0 if prm is None else prm.shape[1]
    mu = torch.empty(P, dtype=torch.float64, device="cuda")
    go = torch.empty((P, 4 + NP), dtype=torch.float64, device="cuda")
    if P == 0:
        return mu, go
    rc = _lib.lib().caus_hexadecapole_vjp(_lib.stream_ptr(), nlenses, _lib.ptr(w), P,
                                          _lib.ptr(prm), NP, _lib.ptr(rho_t), 1, float(u1),
                                          _lib.ptr(g), _lib.ptr(mu), _lib.ptr(go), None,
                                          int(itmax))
    _lib.check(rc, "caus_hexadecapole_vjp")
    return mu, go


def _gate(nlenses, w, rho_t, prm, itmax):
    P = w.numel()
    mu = torch.empty(P, dtype=torch.float64, device="cuda")
    use = torch.empty(P, dtype=torch.uint8, device="cuda")
    NP = 0 if prm is None else prm.shape[1]
    rc = _lib.lib().caus_hexadecapole_gate(_lib.stream_ptr(), nlenses, _lib.ptr(w), P,
                                           _lib.ptr(prm), NP, _lib.ptr(rho_t), 1, 0.0,
                                           _lib.ptr(mu), None, _lib.ptr(use), None, int(itmax))
    _lib.check(rc, "caus_hexadecapole_gate")
    return mu, use.bool()


def _mag_fwd(cfg, w, rho_t, u1_t, prm, g=None, work=None):
    """Gate, then the contour pipeline on the points the gate rejects; with g also the
    VJP of both branches.  The gated points are skipped inside the contour launches
    (npts_limb 0, set on the device from the gate's output) rather than gathered on the
    host, so nothing waits for the host.  Returns (mag, use, g_out or None)."""
    from .extended_source import _run

    nlenses, npts_host, npts_max, ld, npts_ld, itmax, mode_id = cfg
    mu, use = _gate(nlenses, w, rho_t, prm, itmax)
    m, _, gr = _run(nlenses, w, rho_t, prm, npts_host, npts_max, ld, u1_t if ld else None,
                    npts_ld, itmax, mode_id, g=g, work=work, skip=use)
    out = torch.where(use, mu, m)
    go = None
    if g is not None:
        _, ga = _hex_vjp(nlenses, w, rho_t, 0.0, prm, g, itmax)
        ga[:, 3] = 0.0   # the gate's multipoles use no limb darkening
        go = torch.where(use[:, None], ga, gr)
    return out, use, go


class _Mag(torch.autograd.Function):
    """Per-point outputs: the forward keeps the block-diagonal Jacobian (a unit-cotangent
    VJP of both branches) when a gradient is needed, and the backward scales it."""

    @staticmethod
    def forward(ctx, cfg, w, rho_t, u1_t, *ps):
        prm = torch.stack(list(ps), -1).contiguous() if ps else None
        need = any(ctx.needs_input_grad[1:])
        ones = torch.ones(w.numel(), dtype=torch.float64, device="cuda") if need else None
        out, _, jac = _mag_fwd(cfg, w, rho_t, u1_t, prm, ones)
        ctx.nps = len(ps)
        if need:
            ctx.save_for_backward(jac)
        return out

    @staticmethod
    def backward(ctx, g):
        (jac,) = ctx.saved_tensors
        go = g.to(torch.float64)[:, None] * jac
        grads = [torch.complex(go[:, 0], go[:, 1]), go[:, 2], go[:, 3]]
        grads += [go[:, 4 + k] for k in range(ctx.nps)]
        return (None, *grads)


def _mag_args(w_points, rho, nlenses, npts_limb, limb_darkening, u1, npts_ld, roots_itmax,
              roots_compensated, mode, params):
    from .extended_source import _col, _host_npts

    if nlenses not in (1, 2, 3):
        raise ValueError("`nlenses` has to be set to be <= 3.")
    w = w_points if isinstance(w_points, torch.Tensor) else torch.as_tensor(np.asarray(w_points))
    shape = tuple(w.shape)
    w = w.to(device="cuda", dtype=torch.complex128).reshape(-1).contiguous()
    P = w.numel()
    names = PARAM_NAMES[nlenses]
    cfg = (nlenses, _host_npts(npts_limb, P), int(np.max(npts_limb)), bool(limb_darkening),
           int(npts_ld), int(roots_itmax), _mode_id(roots_compensated, mode))
    return cfg, shape, w, _col(rho, P), _col(u1, P), [_col(params[k], P) for k in names]


def mag(w_points, rho, nlenses=2, npts_limb=200, limb_darkening=False, u1=0.0, npts_ld=100,
        roots_itmax=2500, roots_compensated=False, mode=None, **params):
    """Extended-source magnification at source positions w_points (reference:
    src/caustics/lightcurve.py:88-254): the hexadecapole value where the gate accepts
    it, the full contour integral elsewhere.  As in the reference the multipole terms
    of the gate use no limb darkening; three-lens systems always use the contour.
    Differentiable in w_points, rho, u1 and the lens parameters (hexadecapole VJP on
    the accepted points, contour VJP on the others)."""
    cfg, shape, w, rho_t, u1_t, ps = _mag_args(w_points, rho, nlenses, npts_limb,
                                               limb_darkening, u1, npts_ld, roots_itmax,
                                               roots_compensated, mode, params)
    return _Mag.apply(cfg, w, rho_t, u1_t, *ps).reshape(shape)


def mag_fwd_vjp(w_points, rho, g_mag, nlenses=2, npts_limb=200, limb_darkening=False, u1=0.0,
                npts_ld=100, roots_itmax=2500, roots_compensated=False, mode=None, **params):
    """Fused forward + VJP of mag(): (mag (P,), g_out (P, 4+n_params), use_hex (P,)), g_out
    columns Re w, Im w, rho, u1 and the lens parameters."""
    cfg, shape, w, rho_t, u1_t, ps = _mag_args(w_points, rho, nlenses, npts_limb,
                                               limb_darkening, u1, npts_ld, roots_itmax,
                                               roots_compensated, mode, params)
    from .extended_source import _col

    prm = torch.stack(ps, -1).contiguous() if ps else None
    m, use, go = _mag_fwd(cfg, w, rho_t, u1_t, prm, _col(g_mag, w.numel()))
    return m, go, use
