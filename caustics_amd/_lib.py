"""ctypes binding of the gfx950 library ``libcaustics_hip.so`` (C ABI in
``include/caustics_hip.h``).

The library is the only compute path: if it is missing, or no HIP device is
visible, every operation raises instead of computing anything on the CPU.
"""
import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcaustics_hip.so")

CAUS_OK = 0
CAUS_ERR_ARG = -1
CAUS_ERR_LAUNCH = -2

LC_NOALLOC_MAX_BYTES = 8 << 20   # CAUS_LC_NOALLOC_MAX_BYTES

MODE_FP64 = 0
MODE_FP64_COMP = 1
MODE_CF32 = 2
MODES = {"fp64": MODE_FP64, "fp64_comp": MODE_FP64_COMP, "cf32": MODE_CF32}

INIT_REFERENCE = 0
INIT_COMPLEX = 1
INITS = {"reference": INIT_REFERENCE, "complex": INIT_COMPLEX}

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64

# name -> (restype, argtypes); mirrors include/caustics_hip.h
SIGNATURES = {
    "caus_abi_version": (ctypes.c_int, []),
    "caus_roots": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _i32, _i32, _i32, _i32]),
    "caus_gpu_ehrlich_aberth": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.c_char_p,
                                               ctypes.c_size_t]),
    "caus_gpu_ehrlich_aberth_void": (None, [_vp, ctypes.POINTER(_vp), ctypes.c_char_p,
                                            ctypes.c_size_t]),
    "caus_gpu_ehrlich_aberth_status": (ctypes.c_int, []),
    "caus_gpu_ehrlich_aberth_vjp_void": (None, [_vp, ctypes.POINTER(_vp), ctypes.c_char_p,
                                                ctypes.c_size_t]),
    "caus_ext_set_flat_blocks_per_cu": (ctypes.c_int, [_i32]),
    "caus_roots_vjp": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _i64, _i32]),
    "caus_mag_point_source": (ctypes.c_int, [_vp, _i32, _vp, _i64, _vp, _i64, _vp, _vp, _i32,
                                             _i32]),
    "caus_mag_point_source_vjp": (ctypes.c_int, [_vp, _i32, _vp, _i64, _vp, _i64, _vp, _vp, _vp,
                                                 _vp, _vp, _i32, _i32]),
    "caus_lightcurve_workspace_bytes": (ctypes.c_int64, [_i32, _i64, _i64]),
    "caus_lightcurve_point_source": (ctypes.c_int, [_vp, _i32, _vp, _i64, _vp, _i64, _vp, _vp,
                                                    _vp, _vp, _i64, _vp, _i32, _i32]),
    "caus_lightcurve_point_source_noalloc": (ctypes.c_int, [_vp, _i32, _vp, _i64, _vp, _i64, _vp,
                                                            _vp, _vp, _vp, _i32, _i32]),
    "caus_extended_workspace_bytes": (ctypes.c_int64, [_i32, _i64, _i32]),
    "caus_mag_extended_source": (ctypes.c_int, [_vp, _i32, _vp, _i64, _vp, _vp, _i64, _vp, _i32,
                                                _i32, _vp, _i64, _i32, _vp, _vp, _vp, _i64, _i32,
                                                _i32]),
    "caus_mag_extended_source_vjp": (ctypes.c_int, [_vp, _i32, _vp, _i64, _vp, _vp, _i64, _vp,
                                                    _i32, _i32, _vp, _i64, _i32, _vp, _vp, _vp,
                                                    _vp, _vp, _i64, _i32, _i32]),
    "caus_mag_extended_source_work": (ctypes.c_int, [_vp, _i32, _vp, _i64, _vp, _vp, _i64, _vp,
                                                     _i32, _i32, _vp, _i64, _i32, _vp, _vp, _vp,
                                                     _vp, _vp, _i64, _i32, _i32, _vp]),
    "caus_match_chain": (ctypes.c_int, [_vp, _vp, _vp, _i64, _i32, _i64]),
    "caus_images_point_source": (ctypes.c_int, [_vp, _i32, _vp, _i64, _i64, _vp, _i64, _vp, _vp,
                                                _vp, _vp, _i32, _i32, _i32]),
    "caus_marginalized_log_likelihood": (ctypes.c_int, [_vp, _vp, _i64, _i64, _i64, _vp, _vp,
                                                        _vp, _i32, _vp, _vp, _vp, _vp, _i64]),
    "caus_trajectory_annual_parallax": (ctypes.c_int, [_vp, _vp, _i64, _vp, _vp, _i64, _vp,
                                                       _i64, _i32, _vp, _vp, _vp]),
    "caus_hexadecapole_gate": (ctypes.c_int, [_vp, _i32, _vp, _i64, _vp, _i64, _vp, _i64,
                                              ctypes.c_double, _vp, _vp, _vp, _vp, _i32]),
    "caus_hexadecapole_vjp": (ctypes.c_int, [_vp, _i32, _vp, _i64, _vp, _i64, _vp, _i64,
                                             ctypes.c_double, _vp, _vp, _vp, _vp, _i32]),
}

_LIB = None


def load_library():
    """Load the shared library and bind every declared symbol (no device needed)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"caustics_amd: {LIB_PATH} is missing; build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'` (no CPU fallback exists)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def lib():
    """The library, after checking that a HIP device is present."""
    L = load_library()
    if not torch.cuda.is_available():
        raise RuntimeError("caustics_amd: no HIP device visible; the hot path runs only on the GPU")
    return L


_SIDE_STREAMS = {}


def side_streams(n=2):
    """The library's pool of extra default-priority streams on the current device,
    created once and shared by every caller: a process has few hardware queues (4 by
    default), and two streams that share one run their work one after the other."""
    dev = torch.cuda.current_device()
    got = _SIDE_STREAMS.setdefault(dev, [])
    while len(got) < n:
        got.append(torch.cuda.Stream())
    return got[:n]


def stream_ptr(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def check(rc, what):
    if rc != CAUS_OK:
        kind = {CAUS_ERR_ARG: "invalid argument", CAUS_ERR_LAUNCH: "launch failed"}.get(rc, str(rc))
        raise RuntimeError(f"caustics_amd: {what}: {kind} (code {rc})")
