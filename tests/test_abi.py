"""The C-ABI library loads and exports every entry point include/*.h declares
(no GPU needed: nothing is called)."""
import ctypes
import glob
import os
import re

from caustics_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(caus_\w+)\s*\(", text))
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "caus_roots" in names and "caus_roots_vjp" in names


def test_library_exports_all_declared_symbols():
    L = _lib.load_library()
    missing = [n for n in sorted(declared_symbols()) if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert declared_symbols() <= set(_lib.SIGNATURES)


def test_version_call_needs_no_device():
    assert _lib.load_library().caus_abi_version() == 1


def test_argument_errors_enqueue_nothing():
    # invalid arguments are rejected before any HIP call, so this runs without a GPU
    L = _lib.load_library()
    assert L.caus_roots(None, None, None, None, None, None, 10, 5, 100, 0, 0) == _lib.CAUS_ERR_ARG
    assert L.caus_roots(None, None, None, None, None, None, 0, 5, 100, 0, 0) == _lib.CAUS_OK
    assert L.caus_roots(None, None, None, None, None, None, -1, 5, 100, 0, 0) == _lib.CAUS_ERR_ARG
    # trajectories: a grid of fewer than two points, an unknown parametrization and a
    # cotangent without its gradient output are rejected before any HIP call
    tr = L.caus_trajectory_annual_parallax
    assert tr(None, None, 10, None, None, 1, None, 4, 0, None, None, None) == _lib.CAUS_ERR_ARG
    assert tr(None, None, 10, None, None, 5, None, 4, 2, None, None, None) == _lib.CAUS_ERR_ARG
    assert tr(None, None, 10, None, None, 5, None, 4, 0, None, 8, None) == _lib.CAUS_ERR_ARG
    assert tr(None, None, 10, None, None, 5, None, 0, 0, None, None, None) == _lib.CAUS_OK
