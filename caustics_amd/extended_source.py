"""Extended-source magnification on the GPU (reference: src/caustics/extended_source.py).

``mag_extended_source(w0, rho, nlenses=2, npts_limb=150, limb_darkening=False, u1=0.0,
npts_ld=100, roots_itmax=2500, roots_compensated=False, **params)`` (:730-901) computes,
per source, the images of the source limb with the reference's 10 rounds of gap
refinement, chains the image rows along the limb, splits and stitches the image
segments into closed contours and integrates them (uniform: Green's theorem with the
trapezoid rule; linear limb darkening: Dominik's P/Q line integrals with Gauss-Legendre
quadrature).  The work runs as one fixed sequence of HIP launches over the whole batch:
flat launches for the limb solves (one limb point per lane), one wave or workgroup per
source for continuation, gap selection, stitching and integration.  Gradients with
respect to the source centre, rho, u1 and the lens parameters come from the fused
forward+VJP pass.  ``roots_compensated=True`` runs the limb solves with the
compensated stage (ehrlich_aberth.h:153-292).

Unlike the reference (one scalar source per call) ``w0`` may be an array of sources;
``rho``, ``u1``, ``npts_limb`` and the lens parameters broadcast against it, so a
ragged batch (different ``npts_limb`` per source) is one call.  A call synchronises
nothing with the host, so it can be captured in a HIP graph.
"""
import numpy as np
import torch

from . import _lib
from .point_source import PARAM_NAMES, _mode_id

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


def _host_npts(npts_limb, S):
    """npts_limb as a host int64 array of length S (or None for one value): the batch
    is ordered and cut into chunks on the host, from values the caller holds there."""
    if isinstance(npts_limb, torch.Tensor):
        raise TypeError("npts_limb must be a host value (int, list or numpy array)")
    a = np.asarray(npts_limb, dtype=np.int64)
    if a.ndim == 0:
        return None
    a = np.broadcast_to(a.reshape(-1), (S,)) if a.size == 1 else a.reshape(-1)
    if a.size != S:
        raise ValueError("npts_limb does not broadcast against w0")
    return a


_LAYOUTS = {}


def _ragged_layout(npts_host, S, npts_max, need_npts):
    """Host-side order of a ragged batch (descending npts_limb, stable), the sorted
    per-source sizes, and their device copies (order, npts as int32).  The device copies
    are kept per layout (a few, least recently used first out), so a repeated call, and
    in particular a call captured in a HIP graph after a warm-up call, copies nothing
    from the host."""
    if npts_host is None and not need_npts:
        return None, None, None
    key = (S, npts_max, None if npts_host is None else npts_host.tobytes())
    hit = _LAYOUTS.pop(key, None)
    if hit is None:
        if npts_host is not None:
            eff = np.clip(npts_host, 4, npts_max)
            order_np = np.argsort(-eff, kind="stable")
            eff = eff[order_np]
            sorted_n = np.asarray(npts_host)[order_np]
            order = (None if np.array_equal(order_np, np.arange(S))
                     else torch.as_tensor(order_np, device="cuda"))
        else:
            eff, order, sorted_n = None, None, np.full(S, npts_max, np.int64)
        npts_t = torch.as_tensor(np.ascontiguousarray(sorted_n), dtype=torch.int32, device="cuda")
        hit = (order, eff, npts_t)
    _LAYOUTS[key] = hit
    while len(_LAYOUTS) > 8:
        _LAYOUTS.pop(next(iter(_LAYOUTS)))
    return hit


_flat_set = [None]


def _run(nlenses, w, rho_t, prm, npts_host, npts_max, ld, u1_t, npts_ld, itmax, mode_id, g=None,
         max_chunk=None, work=None, streams=None, skip=None, chunk_bytes=None):
    """One or more launch sequences of the contour pipeline, one per chunk of sources.
    A ragged batch (npts_host, per-source npts_limb held on the host) is processed in
    descending npts_limb order, so that every launch holds sources of similar cost (no
    workgroup of a long contour trailing a launch of short ones) and each chunk's
    workspace is sized by its own largest contour; results return in the caller's order.
    With several chunks they alternate over `streams` streams, each with its own
    workspace, so that one chunk's low-occupancy launches (refinement rounds,
    continuation, stitching) overlap another chunk's solves and integrals; the caller's
    stream waits for all of them.  `skip`, a (S,) bool CUDA tensor, marks sources to
    leave out (npts_limb 0 for the kernels: no work, magnification and gradient 0); it
    is applied on the device, so nothing waits for the host.  Per-source results do not
    depend on the batch around them.  `work`, a (5,) int64 CUDA tensor, accumulates the
    kernels' work counters (FP64, FP32 and compensated root-iterations, limb-point
    solves, contour points).  The schedule (chunk workspace budget `chunk_bytes`,
    `streams`, the flat kernels' workgroups per CU) comes from the tuned table of
    caustics_amd.autotune unless the caller fixes it; it changes no result."""
    L = _lib.lib()
    S = w.numel()
    if max_chunk is None and (chunk_bytes is None or streams is None):
        from . import autotune
        sch = autotune.lookup(nlenses, ld, g is not None, npts_max)
        chunk_bytes = chunk_bytes or sch["chunk_bytes"]
        streams = streams or sch["streams"]
        if _flat_set[0] != sch["flat_per_cu"]:
            _lib.check(L.caus_ext_set_flat_blocks_per_cu(int(sch["flat_per_cu"])), "schedule")
            _flat_set[0] = sch["flat_per_cu"]
    chunk_bytes = chunk_bytes or (2 << 30)
    streams = streams or 2
    P = 0 if prm is None else prm.shape[1]
    if extended_workspace_bytes(nlenses, 1, npts_max) < 0:
        raise ValueError(f"npts_limb={npts_max} out of range (4..4096)")
    order, eff, npts_t = _ragged_layout(npts_host, S, npts_max, skip is not None)
    if order is not None:
        pick = lambda x: None if x is None else x[order].contiguous()
        w, rho_t, prm, u1_t, g, skip = (pick(w), pick(rho_t), pick(prm), pick(u1_t), pick(g),
                                        pick(skip))
    if skip is not None:
        npts_t = torch.where(skip, torch.zeros_like(npts_t), npts_t).contiguous()
    mag = torch.empty(S, dtype=torch.float64, device="cuda")
    status = torch.empty(S, dtype=torch.uint8, device="cuda")
    g_out = torch.empty((S, 4 + P), dtype=torch.float64, device="cuda") if g is not None else None
    # chunks (lo, hi, largest npts_limb of the chunk)
    chunks = []
    lo = 0
    while lo < S:
        nm = npts_max if eff is None else int(eff[lo])
        per = extended_workspace_bytes(nlenses, 1, nm)
        cnt = max_chunk or max(1, int(chunk_bytes) // per)
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
    """Each magnification depends only on its own source's inputs, so the Jacobian is
    block diagonal: when a gradient is needed the forward runs the fused forward+VJP
    pass with unit cotangents and keeps dmag_i/dtheta_i (S, 4+P); the backward scales
    it by the incoming cotangents instead of running the contour pipeline again."""

    @staticmethod
    def forward(ctx, cfg, w, rho_t, u1_t, *ps):
        nlenses, npts_host, npts_max, ld, npts_ld, itmax, mode_id, shapes = cfg
        prm = torch.stack(list(ps), -1).contiguous() if ps else None
        need = any(ctx.needs_input_grad[1:])
        ones = torch.ones(w.numel(), dtype=torch.float64, device="cuda") if need else None
        mag, status, jac = _run(nlenses, w, rho_t, prm, npts_host, npts_max, ld,
                                u1_t if ld else None, npts_ld, itmax, mode_id, g=ones)
        ctx.nps = len(ps)
        ctx.ld = ld
        if need:
            ctx.save_for_backward(jac)
        ctx.mark_non_differentiable(status)
        return mag, status

    @staticmethod
    def backward(ctx, g, _gs):
        (jac,) = ctx.saved_tensors
        go = g.to(torch.float64)[:, None] * jac
        gw = torch.complex(go[:, 0], go[:, 1])
        grads = [gw, go[:, 2], go[:, 3] if ctx.ld else torch.zeros_like(go[:, 3])]
        grads += [go[:, 4 + k] for k in range(ctx.nps)]
        return (None, *grads)


def mag_extended_source(w0, rho, nlenses=2, npts_limb=150, limb_darkening=False, u1=0.0,
                        npts_ld=100, roots_itmax=2500, roots_compensated=False, mode=None,
                        return_status=False, **params):
    """Magnification of uniform or linearly limb-darkened circular sources of radius
    ``rho`` centred at ``w0`` (same frame and parameters as ``mag_point_source``).
    Differentiable in w0, rho, u1 and the lens parameters."""
    if nlenses not in (1, 2, 3):
        raise ValueError("`nlenses` has to be set to be <= 3.")
    mode_id = _mode_id(roots_compensated, mode)
    w = w0 if isinstance(w0, torch.Tensor) else torch.as_tensor(np.asarray(w0))
    shape = tuple(w.shape)
    w = w.to(device="cuda", dtype=torch.complex128).reshape(-1).contiguous()
    S = w.numel()
    names = PARAM_NAMES[nlenses]
    rho_t = _col(rho, S)
    u1_t = _col(u1, S)
    ps = [_col(params[k], S) for k in names]
    npts_host = _host_npts(npts_limb, S)
    npts_max = int(np.max(npts_limb))
    cfg = (nlenses, npts_host, npts_max, bool(limb_darkening), int(npts_ld), int(roots_itmax),
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
    mag, status, g_out = _run(nlenses, w, _col(rho, S), prm, _host_npts(npts_limb, S),
                              int(np.max(npts_limb)), limb_darkening,
                              _col(u1, S) if limb_darkening else None, npts_ld, roots_itmax,
                              _lib.MODES[mode], g=_col(g_mag, S))
    return mag, g_out, status
