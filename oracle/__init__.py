"""CPU oracle (test infrastructure): loader for oracle/_build/libslo.so.

Only the tests, the smoke check and the benchmark's CPU-baseline leg use this.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libslo.so")

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def load():
    """Return an api.Library bound to the oracle (prefix slo_)."""
    global _lib
    if _lib is None:
        from softlight_amd import api
        if not os.path.exists(LIB_PATH):
            build()
        _lib = api.Library(ctypes.CDLL(LIB_PATH), "slo_")
    return _lib


def new_context():
    from softlight_amd import api
    return api.SLContext(load())
