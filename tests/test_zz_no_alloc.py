"""caus_lightcurve_point_source_noalloc (no allocation: its partial sums live in the
library's static device scratch) against caus_lightcurve_point_source with a caller
workspace, through the C ABI: two calls in a row, the second with a smaller batch, must
give bit-identical magnifications and gradients; a batch beyond the scratch is refused
before anything is enqueued."""
import ctypes

import numpy as np
import pytest
import torch

from caustics_amd import _lib
from caustics_amd import workloads as W

pytestmark = pytest.mark.gpu

_vp, _i32, _i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64


def _bind():
    L = _lib.lib()
    f = L.caus_lightcurve_point_source_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [_vp, _i32, _vp, _i64, _vp, _i64, _vp, _vp, _vp, _vp, _i32, _i32]
    return L, f


def _batch(L_, T, seed):
    p = W.lightcurve_params(L_, seed)
    lcp = np.stack([p[k] for k in ("t0", "u0", "tE", "alpha", "s", "q")], -1)
    t = np.linspace(-60.0, 60.0, T)
    g = np.random.default_rng(seed).normal(size=(L_, T))
    dev = lambda x: torch.as_tensor(x, dtype=torch.float64, device="cuda").contiguous()
    return dev(t), dev(lcp), dev(g)


def _with_workspace(L, t, lcp, g, mode):
    Lc, T = lcp.shape[0], t.numel()
    mag = torch.empty((Lc, T), dtype=torch.float64, device="cuda")
    gp = torch.empty_like(lcp)
    st = torch.empty((Lc, T), dtype=torch.uint8, device="cuda")
    wsb = L.caus_lightcurve_workspace_bytes(2, T, Lc)
    ws = torch.empty(wsb // 8, dtype=torch.float64, device="cuda")
    rc = L.caus_lightcurve_point_source(_lib.stream_ptr(), 2, _lib.ptr(t), T, _lib.ptr(lcp), Lc,
                                        _lib.ptr(g), _lib.ptr(mag), _lib.ptr(gp), _lib.ptr(ws),
                                        wsb, _lib.ptr(st), 2500, mode)
    assert rc == _lib.CAUS_OK
    return mag, gp, st


def _no_alloc(f, t, lcp, g, mode):
    Lc, T = lcp.shape[0], t.numel()
    mag = torch.empty((Lc, T), dtype=torch.float64, device="cuda")
    gp = torch.empty_like(lcp)
    st = torch.empty((Lc, T), dtype=torch.uint8, device="cuda")
    rc = f(_lib.stream_ptr(), 2, _lib.ptr(t), T, _lib.ptr(lcp), Lc, _lib.ptr(g), _lib.ptr(mag),
           _lib.ptr(gp), _lib.ptr(st), 2500, mode)
    assert rc == _lib.CAUS_OK
    return mag, gp, st


@pytest.mark.parametrize("mode", [_lib.MODE_FP64, _lib.MODE_CF32])
def test_no_alloc_matches_workspace_entry(gpu, mode):
    L, f = _bind()
    # the second call is smaller: it reads only the slots its own kernels wrote, not
    # what the first call left in the shared scratch
    for Lc, T, seed in ((256, 2000, 3), (37, 700, 4)):
        t, lcp, g = _batch(Lc, T, seed)
        a = _no_alloc(f, t, lcp, g, mode)
        b = _with_workspace(L, t, lcp, g, mode)
        torch.cuda.synchronize()
        for x, y in zip(a, b):
            assert torch.equal(x, y)
        assert torch.isfinite(a[1]).all()


def test_no_alloc_refuses_batches_beyond_its_scratch(gpu):
    L, f = _bind()
    T, Lc = 200000, 4096
    need = L.caus_lightcurve_workspace_bytes(2, T, Lc)
    assert need > _lib.LC_NOALLOC_MAX_BYTES
    # nothing is read or launched: null buffers are never touched
    rc = f(_lib.stream_ptr(), 2, None, T, None, Lc, ctypes.c_void_p(8), None, None, None, 2500,
           _lib.MODE_FP64)
    assert rc == _lib.CAUS_ERR_ARG
