"""The C-ABI boundary: the HIP library loads without a GPU and exports every
entry point include/sl_gpu.h declares; the oracle exports the same set under
the slo_ prefix; struct layouts used across the boundary have the sizes the
header implies.  No compute calls are made here."""
import ctypes
import os
import re

import pytest

from softlight_amd import api as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sl_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(sl_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_the_draw_api():
    names = declared_functions()
    for required in ("sl_draw", "sl_draw_multiple", "sl_draw_instanced", "sl_create_shader",
                     "sl_clear_framebuffer", "sl_texture_read", "sl_fbo_valid"):
        assert required in names
    assert len(names) >= 39


def test_hip_library_exports_every_declared_symbol():
    import softlight_amd
    lib = ctypes.CDLL(softlight_amd.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_oracle_exports_the_same_api():
    import oracle
    oracle.load()
    lib = ctypes.CDLL(oracle.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, "slo_" + n[3:])]
    assert not missing, missing


def test_struct_sizes():
    assert ctypes.sizeof(A.sl_mesh) == 32                    # SL_Mesh.hpp: "32 bytes"
    assert ctypes.sizeof(A.sl_std_uniforms) == 16 * 4 * 2 + 16 * 4 + 16
    assert ctypes.sizeof(A.sl_std_uniforms) <= 1024          # fits an SL_UniformBuffer
    assert ctypes.sizeof(A.sl_frame_stats) == 6 * 8 + 6 * 8 + 2 * 8 + 8 + 8 + 8


def test_product_loader_refuses_missing_library(monkeypatch, tmp_path):
    import softlight_amd
    monkeypatch.setattr(softlight_amd, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(softlight_amd, "_lib", None)
    with pytest.raises(RuntimeError):
        softlight_amd.load_library()
