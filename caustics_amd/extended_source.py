"""Extended-source magnification on the GPU (reference: src/caustics/extended_source.py).

``mag_extended_source(w0, rho, nlenses=2, npts_limb=150, limb_darkening=False, u1=0.0,
npts_ld=100, roots_itmax=2500, roots_compensated=False, **params)`` (:730-901) computes,
per source, the images of the source limb with the reference's 10 rounds of gap
refinement, chains the image rows along the limb, splits and stitches the image
segments into closed contours and integrates them (uniform: Green's theorem with the
trapezoid rule; linear limb darkening: Dominik's P/Q line integrals with Gauss-Legendre
quadrature), all inside one HIP workgroup per source.  Gradients with respect to the
source centre, rho, u1 and the lens parameters come from the fused forward+VJP kernel.

Unlike the reference (one scalar source per call) ``w0`` may be an array of sources;
``rho``, ``u1``, ``npts_limb`` and the lens parameters broadcast against it, so a
ragged batch (different ``npts_limb`` per source) is one launch.
"""
import numpy as np
import torch

from . import _lib
from .point_source import PARAM_NAMES

__all__ = ["mag_extended_source", "extended_workspace_bytes", "extended_fwd_vjp"]


def extended_workspace_bytes(nlenses, nsources, npts_max):
    return int(_lib.load_library().caus_extended_workspace_bytes(nlenses, nsources, npts_max))


def _col(x, n, dtype=torch.float64):
    t = x if isinstance(x, torch.Tensor) else torch.as_tensor(np.asarray(x))
    t = t.to(device="cuda", dtype=dtype).reshape(-1)
    if t.numel() == 1:
        t = t.expand(n)
    if t.numel() != n:
        raise ValueError("argument does not broadcast against w0")
    return t.contiguous()


def _run(nlenses, w, rho_t, prm, npts_t, npts_max, ld, u1_t, npts_ld, itmax, mode_id, g=None,
         max_chunk=None, work=None, streams=2):
    """One or more launch sequences of the contour pipeline, one per chunk of sources.
    A ragged batch (per-source npts_limb) is processed in descending npts_limb order, so
    that every launch holds sources of similar cost (no workgroup of a long contour
    trailing a launch of short ones) and each chunk's workspace is sized by its own
    largest contour; results return in the caller's order.  With several chunks they
    alternate over `streams` streams, each with its own workspace, so that one chunk's
    low-occupancy launches (refinement rounds, continuation, stitching) overlap another
    chunk's solves and integrals; the caller's stream waits for all of them.  Per-source
    results do not depend on the batch around them.  `work`, a (5,) int64 CUDA tensor,
    accumulates the kernels' work counters (FP64, FP32 and compensated root-iterations,
    limb-point solves, contour points)."""
    L = _lib.lib()
    S = w.numel()
    P = 0 if prm is None else prm.shape[1]
    if extended_workspace_bytes(nlenses, 1, npts_max) < 0:
        raise ValueError(f"npts_limb={npts_max} out of range (4..4096)")
    order = None
    if npts_t is not None and S > 1:
        eff = np.clip(npts_t.detach().cpu().numpy().astype(np.int64), 4, npts_max)
        order_np = np.argsort(-eff, kind="stable")
        eff = eff[order_np]
        if not np.array_equal(order_np, np.arange(S)):
            order = torch.as_tensor(order_np, device="cuda")
            pick = lambda x: None if x is None else x[order].contiguous()
            w, rho_t, prm, npts_t, u1_t, g = (pick(w), pick(rho_t), pick(prm), pick(npts_t),
                                              pick(u1_t), pick(g))
    else:
        eff = None
    mag = torch.empty(S, dtype=torch.float64, device="cuda")
    status = torch.empty(S, dtype=torch.uint8, device="cuda")
    g_out = torch.empty((S, 4 + P), dtype=torch.float64, device="cuda") if g is not None else None
    # chunks (lo, hi, largest npts_limb of the chunk)
    chunks = []
    lo = 0
    while lo < S:
        nm = npts_max if eff is None else int(eff[lo])
        per = extended_workspace_bytes(nlenses, 1, nm)
        cnt = max_chunk or max(1, (2 << 30) // per)
        hi = min(S, lo + cnt)
        chunks.append((lo, hi, nm))
        lo = hi
    nstr = max(1, min(int(streams), len(chunks)))
    main = torch.cuda.current_stream()
    lanes = [main] if nstr == 1 else _lib.side_streams(nstr)
    need = [0] * nstr
    for ci, (lo, hi, nm) in enumerate(chunks):
        need[ci % nstr] = max(need[ci % nstr], extended_workspace_bytes(nlenses, hi - lo, nm))
    ws = [torch.empty((b + 7) // 8, dtype=torch.float64, device="cuda") for b in need]
    for st in lanes:
        if st is not main:
            st.wait_stream(main)   # inputs and outputs were produced on the caller's stream
    try:
        for ci, (lo, hi, nm) in enumerate(chunks):
            sl = slice(lo, hi)
            k = ci % nstr
            with torch.cuda.stream(lanes[k]):
                rc = L.caus_mag_extended_source_work(
                    _lib.stream_ptr(), nlenses, _lib.ptr(w[sl]), hi - lo, _lib.ptr(rho_t[sl]),
                    None if prm is None else _lib.ptr(prm[sl]), P,
                    None if npts_t is None else _lib.ptr(npts_t[sl]), nm, 1 if ld else 0,
                    None if u1_t is None else _lib.ptr(u1_t[sl]), 1, int(npts_ld),
                    None if g is None else _lib.ptr(g[sl]), _lib.ptr(mag[sl]),
                    None if g_out is None else _lib.ptr(g_out[sl]), _lib.ptr(status[sl]),
                    _lib.ptr(ws[k]), ws[k].numel() * 8, int(itmax), mode_id,
                    None if work is None else _lib.ptr(work))
            _lib.check(rc, "caus_mag_extended_source")
    finally:
        # also after a failed launch: what was enqueued still orders every later use
        for st in lanes:
            if st is not main:
                main.wait_stream(st)
                for t in [*ws, mag, status, g_out, w, rho_t, prm, npts_t, u1_t, g, work]:
                    if t is not None:
                        t.record_stream(st)   # no reuse of their memory before the work ends
    if order is not None:
        back = lambda x: None if x is None else torch.empty_like(x).index_copy_(0, order, x)
        mag, status, g_out = back(mag), back(status), back(g_out)
    return mag, status, g_out


class _Extended(torch.autograd.Function):
    @staticmethod
    def forward(ctx, cfg, w, rho_t, u1_t, *ps):
        nlenses, npts_t, npts_max, ld, npts_ld, itmax, mode_id, shapes = cfg
        prm = torch.stack(list(ps), -1).contiguous() if ps else None
        mag, status, _ = _run(nlenses, w, rho_t, prm, npts_t, npts_max, ld,
                              u1_t if ld else None, npts_ld, itmax, mode_id)
        ctx.cfg = cfg
        ctx.save_for_backward(w, rho_t, u1_t, *ps)
        ctx.mark_non_differentiable(status)
        return mag, status

    @staticmethod
    def backward(ctx, g, _gs):
        nlenses, npts_t, npts_max, ld, npts_ld, itmax, mode_id, shapes = ctx.cfg
        w, rho_t, u1_t, *ps = ctx.saved_tensors
        prm = torch.stack(list(ps), -1).contiguous() if ps else None
        _, _, go = _run(nlenses, w, rho_t, prm, npts_t, npts_max, ld, u1_t if ld else None,
                        npts_ld, itmax, mode_id, g=g.contiguous().to(torch.float64))
        gw = torch.complex(go[:, 0], go[:, 1])
        grads = [gw, go[:, 2], go[:, 3] if ld else torch.zeros_like(go[:, 3])]
        grads += [go[:, 4 + k] for k in range(len(ps))]
        return (None, *grads)


def mag_extended_source(w0, rho, nlenses=2, npts_limb=150, limb_darkening=False, u1=0.0,
                        npts_ld=100, roots_itmax=2500, roots_compensated=False, mode=None,
                        return_status=False, **params):
    """Magnification of uniform or linearly limb-darkened circular sources of radius
    ``rho`` centred at ``w0`` (same frame and parameters as ``mag_point_source``).
    Differentiable in w0, rho, u1 and the lens parameters."""
    if nlenses not in (1, 2, 3):
        raise ValueError("`nlenses` has to be set to be <= 3.")
    if roots_compensated:
        raise NotImplementedError("the contour kernels run the FP64 or CF32 solver")
    mode_id = _lib.MODES[mode or "fp64"]
    w = w0 if isinstance(w0, torch.Tensor) else torch.as_tensor(np.asarray(w0))
    shape = tuple(w.shape)
    w = w.to(device="cuda", dtype=torch.complex128).reshape(-1).contiguous()
    S = w.numel()
    names = PARAM_NAMES[nlenses]
    rho_t = _col(rho, S)
    u1_t = _col(u1, S)
    ps = [_col(params[k], S) for k in names]
    npts_arr = np.asarray(npts_limb)
    npts_max = int(npts_arr.max())
    npts_t = None if npts_arr.ndim == 0 else _col(npts_arr, S, torch.int32)
    cfg = (nlenses, npts_t, npts_max, bool(limb_darkening), int(npts_ld), int(roots_itmax),
           mode_id, None)
    mag, status = _Extended.apply(cfg, w, rho_t, u1_t, *ps)
    mag = mag.reshape(shape)
    return (mag, status.reshape(shape)) if return_status else mag


def extended_fwd_vjp(w0, rho, g_mag, nlenses=2, npts_limb=150, limb_darkening=False, u1=0.0,
                     npts_ld=100, roots_itmax=2500, mode="fp64", **params):
    """Fused forward + VJP: (mag (S,), g_params (S, 4+P), status (S,)); g_params columns
    are Re w0, Im w0, rho, u1 and the lens parameters."""
    w = torch.as_tensor(np.asarray(w0)) if not isinstance(w0, torch.Tensor) else w0
    w = w.to(device="cuda", dtype=torch.complex128).reshape(-1).contiguous()
    S = w.numel()
    names = PARAM_NAMES[nlenses]
    prm = torch.stack([_col(params[k], S) for k in names], -1).contiguous() if names else None
    npts_arr = np.asarray(npts_limb)
    npts_t = None if npts_arr.ndim == 0 else _col(npts_arr, S, torch.int32)
    mag, status, g_out = _run(nlenses, w, _col(rho, S), prm, npts_t, int(npts_arr.max()),
                              limb_darkening, _col(u1, S) if limb_darkening else None, npts_ld,
                              roots_itmax, _lib.MODES[mode], g=_col(g_mag, S))
    return mag, g_out, status
