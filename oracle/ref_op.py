"""Loader for the reference's own CPU root-solver op, compiled from
lib/ehrlich_aberth/cpu_ops.cc into ``oracle/_ref/`` by ``make -C oracle ref``.
Test infrastructure only: used to pin the C oracle and as the timed CPU baseline.

The op is the XLA custom-call target ``cpu_ehrlich_aberth(void* out, const void** in)``
(cpu_ops.cc:15-81) exported in a PyCapsule named "xla._CUSTOM_CALL_TARGET"
(pybind11_kernel_helpers.h:21-24); operands in[0..6] = size, deg, itmax (int64
literals read as int), compensated, custom_init (bool), coeffs, roots_init.
"""
import ctypes
import glob
import importlib.util
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_FN = None


def available():
    return bool(glob.glob(os.path.join(_HERE, "_ref", "ehrlich_aberth_cpu_op*.so")))


def _fn():
    global _FN
    if _FN is None:
        paths = glob.glob(os.path.join(_HERE, "_ref", "ehrlich_aberth_cpu_op*.so"))
        if not paths:
            raise RuntimeError("reference CPU op not built (make -C oracle ref)")
        spec = importlib.util.spec_from_file_location("ehrlich_aberth_cpu_op", paths[0])
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        cap = mod.registrations()["cpu_ehrlich_aberth"]
        get = ctypes.pythonapi.PyCapsule_GetPointer
        get.restype = ctypes.c_void_p
        get.argtypes = [ctypes.py_object, ctypes.c_char_p]
        ptr = get(cap, b"xla._CUSTOM_CALL_TARGET")
        _FN = (mod, ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p))(ptr))
    return _FN[1]


def roots_asc(coeffs_asc, itmax=2000, compensated=False, roots_init=None):
    """Run the reference op on ascending coefficients (N, n+1) -> roots (N, n)."""
    c = np.ascontiguousarray(coeffs_asc, dtype=np.complex128)
    N, n1 = c.shape
    deg = n1 - 1
    out = np.zeros(N * deg, np.complex128)
    custom = roots_init is not None
    ri = (np.ascontiguousarray(roots_init, dtype=np.complex128).reshape(N, deg) if custom
          else np.zeros((N, deg), np.complex128))
    scal = [ctypes.c_int64(N), ctypes.c_int64(deg), ctypes.c_int64(itmax),
            ctypes.c_bool(bool(compensated)), ctypes.c_bool(custom)]
    ptrs = (ctypes.c_void_p * 7)(*[ctypes.addressof(v) for v in scal], c.ctypes.data,
                                 ri.ctypes.data)
    _fn()(out.ctypes.data, ptrs)
    return out.reshape(N, deg)
