"""ctypes wrapper of the C oracle solver (ea_oracle.c). Test infrastructure only.

Follows the reference batch op ``cpu_ehrlich_aberth`` (lib/ehrlich_aberth/cpu_ops.cc:15-81):
coefficients ``(N, n+1)`` complex128 in ascending order, roots ``(N, n)``.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libea_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: run `make -C {_HERE}`")
        L = ctypes.CDLL(path)
        L.ora_roots_batch.restype = ctypes.c_int64
        L.ora_roots_batch.argtypes = [
            ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
            ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
        ]
        L.ora_num_threads.restype = ctypes.c_int
        L.ora_polish_batch.restype = None
        L.ora_polish_batch.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_int]
        _LIB = L
    return _LIB


def roots_asc(coeffs_asc, itmax=2000, compensated=False, roots_init=None, init_mode=0,
              nthreads=0, return_stats=False, polish=False):
    """Roots of polynomials given ascending coefficients (N, n+1).  polish=True ends
    with the canonical compensated polish (ora_polish: no reference counterpart)."""
    c = np.ascontiguousarray(coeffs_asc, dtype=np.complex128)
    N, n1 = c.shape
    deg = n1 - 1
    out = np.zeros((N, deg), np.complex128)
    stats = np.zeros((N, 4), np.int64)
    status = np.zeros(N, np.uint8)
    ri = None
    if roots_init is not None:
        ri = np.ascontiguousarray(roots_init, dtype=np.complex128).reshape(N, deg)
    lib().ora_roots_batch(
        N, deg, itmax, int(bool(compensated)), int(ri is not None), init_mode,
        c.ctypes.data, None if ri is None else ri.ctypes.data, out.ctypes.data,
        stats.ctypes.data, status.ctypes.data, nthreads,
    )
    if polish:
        lib().ora_polish_batch(N, deg, c.ctypes.data, out.ctypes.data, nthreads)
    if return_stats:
        return out, stats, status
    return out


def poly_roots(coeffs, itmax=2000, compensated=False, roots_init=None, init_mode=0, nthreads=0,
               polish=False):
    """Same contract as the reference ``poly_roots`` (ehrlich_aberth_primitive.py:34-94):
    highest-first coefficients of any batch shape -> roots (..., n)."""
    coeffs = np.asarray(coeffs)
    n1 = coeffs.shape[-1]
    flat = coeffs.reshape(-1, n1)[:, ::-1]
    ri = None if roots_init is None else np.asarray(roots_init).reshape(-1, n1 - 1)
    r = roots_asc(flat, itmax=itmax, compensated=compensated, roots_init=ri,
                  init_mode=init_mode, nthreads=nthreads, polish=polish)
    return r.reshape(coeffs.shape[:-1] + (n1 - 1,))
