"""softlight_amd -- MI355X (gfx950) implementation of SoftLight's triangle draw path.

The product is the HIP library `softlight_amd/_build/libsoftlight_amd.so`,
which exports the C ABI of include/sl_gpu.h.  This package only loads it
(`load_library`) and provides the ctypes mirror of SL_Context used by the
tests and the benchmark (`api.SLContext`) plus seeded workload builders
(`scenes`).  There is no CPU fallback: a missing or unloadable library raises.
"""
from __future__ import annotations

import ctypes
import os

from . import api  # noqa: F401

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "_build", "libsoftlight_amd.so")

_lib = None


def load_library() -> "api.Library":
    """Load the HIP library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"softlight_amd: HIP library not built ({LIB_PATH}); run __graft_entry__.build()")
        _lib = api.Library(ctypes.CDLL(LIB_PATH), "sl_")
        missing = _lib.missing_symbols()
        if missing:
            raise RuntimeError(f"softlight_amd: library lacks symbols {missing}")
    return _lib


def new_context() -> "api.SLContext":
    return api.SLContext(load_library())
