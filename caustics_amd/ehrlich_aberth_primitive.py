"""Batched Ehrlich-Aberth root solver on the GPU, with its implicit-function VJP.

Mirrors the reference operator (src/caustics/ehrlich_aberth_primitive.py):
  ``poly_roots(coeffs, itmax=2000, compensated=False, custom_init=False, roots_init=None)``
      (:34-94) highest-first coefficients of any batch shape -> roots (..., n)
  ``ehrlich_aberth(coeffs, roots_init, itmax, compensated, custom_init)``
      (:98-106) ascending coefficients (N, n+1) -> flat roots (N*n,)
Differentiation: the reference's JVP (:254-324) is dz_j = -(sum_k z_j^k dp_k)/p'(z_j);
here the transpose of that linear map runs in the hand-written kernel behind
``caus_roots_vjp``, inside a torch.autograd.Function.  Gradients follow torch's
conjugate-Wirtinger convention (grad = sum_j g_j * conj(dz_j/dp_k)); JAX's
convention differs by a conjugation for complex inputs and agrees for real ones.

Extra keyword arguments beyond the reference API:
  mode  "fp64" (default; "fp64_comp" when compensated=True) or "cf32"
  init  "reference" (default: the reference's initial estimates and branch
        structure, same root order) or "complex"
"""
import numpy as np
import torch

from . import _lib

__all__ = ["poly_roots", "ehrlich_aberth"]


def _as_cuda_c128(x):
    if isinstance(x, torch.Tensor):
        t = x
    else:
        t = torch.as_tensor(np.asarray(x))
    if not t.is_complex():
        t = t.to(torch.float64)
    t = t.to(device="cuda", dtype=torch.complex128)
    return t.contiguous()


def _resolve_mode(compensated, mode):
    if mode is None:
        mode = "fp64_comp" if compensated else "fp64"
    if mode not in _lib.MODES:
        raise ValueError(f"mode must be one of {sorted(_lib.MODES)}")
    return _lib.MODES[mode]


def _solve(coeffs_asc, roots_init, itmax, mode, init, with_stats=False):
    L = _lib.lib()
    N, n1 = coeffs_asc.shape
    deg = n1 - 1
    roots = torch.empty((N, deg), dtype=torch.complex128, device=coeffs_asc.device)
    status = torch.empty((N,), dtype=torch.uint8, device=coeffs_asc.device)
    stats = (torch.empty((N, 4), dtype=torch.int32, device=coeffs_asc.device)
             if with_stats else None)
    rc = L.caus_roots(_lib.stream_ptr(), _lib.ptr(coeffs_asc), _lib.ptr(roots_init),
                      _lib.ptr(roots), _lib.ptr(stats), _lib.ptr(status), N, deg, int(itmax),
                      mode, init)
    _lib.check(rc, "caus_roots")
    return roots, status, stats


class _RootsFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, coeffs_asc, roots_init, itmax, mode, init):
        roots, status, _ = _solve(coeffs_asc, roots_init, itmax, mode, init)
        ctx.save_for_backward(coeffs_asc, roots)
        ctx.mark_non_differentiable(status)
        return roots, status

    @staticmethod
    def backward(ctx, g_roots, _g_status):
        coeffs_asc, roots = ctx.saved_tensors
        g = g_roots.contiguous().to(torch.complex128)
        gc = torch.empty_like(coeffs_asc)
        N, n1 = coeffs_asc.shape
        rc = _lib.lib().caus_roots_vjp(_lib.stream_ptr(), _lib.ptr(coeffs_asc), _lib.ptr(roots),
                                       _lib.ptr(g), _lib.ptr(gc), N, n1 - 1)
        _lib.check(rc, "caus_roots_vjp")
        return gc, None, None, None, None


def ehrlich_aberth(coeffs, roots_init=None, itmax=2000, compensated=False, custom_init=False,
                   mode=None, init="reference", return_status=False):
    """The operator itself: ascending coefficients (N, n+1) -> flat roots (N*n,).

    ``roots_init`` (N, n) is used only when ``custom_init`` is true (as in the
    reference, ehrlich_aberth_primitive.py:73-94)."""
    c = _as_cuda_c128(coeffs)
    if c.dim() != 2:
        raise ValueError("coeffs must have shape (N, n+1)")
    N, n1 = c.shape
    if not 2 <= n1 <= 13:
        raise NotImplementedError(f"polynomial degree {n1 - 1} not supported (1..12)")
    ri = None
    if custom_init:
        if roots_init is None:
            raise ValueError("custom_init=True needs roots_init")
        ri = _as_cuda_c128(roots_init).reshape(N, n1 - 1).contiguous()
    if init not in _lib.INITS:
        raise ValueError(f"init must be one of {sorted(_lib.INITS)}")
    roots, status = _RootsFn.apply(c, ri, itmax, _resolve_mode(compensated, mode),
                                   _lib.INITS[init])
    out = roots.reshape(-1)
    return (out, status) if return_status else out


def poly_roots(coeffs, itmax=2000, compensated=False, custom_init=False, roots_init=None,
               mode=None, init="reference", return_status=False):
    """Roots of polynomials with highest-first coefficients along the last axis.

    Same shapes as the reference ``poly_roots`` (ehrlich_aberth_primitive.py:34-94):
    input (..., n+1) -> output (..., n), complex128 on the GPU."""
    c = _as_cuda_c128(coeffs)
    n1 = c.shape[-1]
    out_shape = tuple(c.shape[:-1]) + (n1 - 1,)
    flat = torch.flip(c.reshape(-1, n1), dims=(-1,)).contiguous()
    ri = None
    if custom_init:
        ri = _as_cuda_c128(roots_init).reshape(flat.shape[0], n1 - 1)
    r = ehrlich_aberth(flat, ri, itmax=itmax, compensated=compensated, custom_init=custom_init,
                       mode=mode, init=init, return_status=return_status)
    if return_status:
        r, st = r
        return r.reshape(out_shape), st.reshape(out_shape[:-1])
    return r.reshape(out_shape)


def solve_with_stats(coeffs_asc, itmax=2000, mode="fp64", init="complex", roots_init=None):
    """Roots, status and per-polynomial work counters (no autograd)."""
    c = _as_cuda_c128(coeffs_asc)
    ri = None if roots_init is None else _as_cuda_c128(roots_init)
    return _solve(c, ri, itmax, _lib.MODES[mode], _lib.INITS[init], with_stats=True)
