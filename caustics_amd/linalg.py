"""Flux-marginalised log-likelihood of light curves on the GPU.

``marginalized_log_likelihood(A_list, fobs_list, C_inv_list, dense_covariance=False,
Lam_sd=1e04)`` keeps the reference's signature and return value
(src/caustics/linalg.py:10-70): a list of magnification vectors, one per independent
set of observations (observatory), the observed fluxes and the diagonal inverse
covariances; it returns ``(beta_list, ll)`` with the least-squares (F_s, F_b) of each
set and the summed log-likelihood.  Magnification vectors may also carry a leading
batch dimension B (one row per parameter set, e.g. the chains of an HMC run): then every
beta is (B, 2) and ll is (B,).

One HIP kernel (``caus_marginalized_log_likelihood``) computes ll, beta and, in the
backward pass, d ll / d A; the gradient then flows into whatever produced A (e.g. the
fused light-curve VJP of ``mag_point_source_lightcurve``).  Gradients flow to A only:
the observations are data.  The reference's dense-covariance branch does not run
(linalg.py:80 calls ``jnp.linals``, :84 uses an undefined ``M``), so
``dense_covariance=True`` raises.  ``Lam_sd`` belongs to that branch and is unused, as
in the reference.
"""
import numpy as np
import torch

from . import _lib

__all__ = ["marginalized_log_likelihood"]


def _call(A, fobs, c_inv, off, nseg, g_ll=None, want_grad=False):
    B, n = A.shape
    ll = torch.empty(B, dtype=torch.float64, device=A.device)
    beta = torch.empty((B, nseg, 2), dtype=torch.float64, device=A.device)
    gA = torch.empty((B, n), dtype=torch.float64, device=A.device) if want_grad else None
    rc = _lib.lib().caus_marginalized_log_likelihood(
        _lib.stream_ptr(), _lib.ptr(A), B, n, n, _lib.ptr(fobs), _lib.ptr(c_inv), _lib.ptr(off),
        nseg, _lib.ptr(ll), _lib.ptr(beta), _lib.ptr(g_ll), _lib.ptr(gA), n)
    _lib.check(rc, "caus_marginalized_log_likelihood")
    return ll, beta, gA


class _MargLL(torch.autograd.Function):
    @staticmethod
    def forward(ctx, A, fobs, c_inv, off, nseg):
        ll, beta, _ = _call(A, fobs, c_inv, off, nseg)
        ctx.save_for_backward(A, fobs, c_inv, off)
        ctx.nseg = nseg
        ctx.mark_non_differentiable(beta)
        return ll, beta

    @staticmethod
    def backward(ctx, g_ll, _g_beta):
        A, fobs, c_inv, off = ctx.saved_tensors
        _, _, gA = _call(A, fobs, c_inv, off, ctx.nseg,
                         g_ll=g_ll.contiguous().to(torch.float64), want_grad=True)
        return gA, None, None, None, None


def _f64(x, device):
    if torch.is_tensor(x):
        return x.to(device=device, dtype=torch.float64).contiguous()
    return torch.as_tensor(np.asarray(x, np.float64), device=device).contiguous()


def marginalized_log_likelihood(A_list, fobs_list, C_inv_list, dense_covariance=False,
                                Lam_sd=1e04):
    if dense_covariance:
        raise NotImplementedError(
            "dense_covariance=True: the reference's dense branch does not run "
            "(src/caustics/linalg.py:80, :84); only the diagonal branch is provided")
    if not (len(A_list) == len(fobs_list) == len(C_inv_list)) or len(A_list) == 0:
        raise ValueError("A_list, fobs_list and C_inv_list must be non-empty and of equal length")
    dev = torch.device("cuda")
    As = [a if isinstance(a, torch.Tensor) else torch.as_tensor(np.asarray(a, np.float64))
          for a in A_list]
    batched = As[0].dim() == 2
    if any((a.dim() == 2) != batched for a in As):
        raise ValueError("all magnification arrays must be 1-D, or all (B, n_k)")
    As = [(a if batched else a.reshape(1, -1)).to(dev, torch.float64) for a in As]
    B = As[0].shape[0]
    if any(a.shape[0] != B for a in As):
        raise ValueError("batched magnification arrays must share the leading dimension")
    fs = [_f64(f, dev).reshape(-1) for f in fobs_list]
    cs = [_f64(c, dev).reshape(-1) for c in C_inv_list]
    lens = [a.shape[1] for a in As]
    if [f.numel() for f in fs] != lens or [c.numel() for c in cs] != lens:
        raise ValueError("each fobs / C_inv must have the length of its magnification array")
    A = torch.cat(As, 1).contiguous()
    fobs = torch.cat(fs)
    c_inv = torch.cat(cs)
    off = torch.tensor([0] + list(torch.tensor(lens).cumsum(0).tolist()), dtype=torch.int64,
                       device=dev)
    ll, beta = _MargLL.apply(A, fobs, c_inv, off, len(As))
    if batched:
        return [beta[:, k] for k in range(len(As))], ll
    return [beta[0, k] for k in range(len(As))], ll[0]
