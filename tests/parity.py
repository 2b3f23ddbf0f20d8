"""Helpers shared by the parity tests: run one call sequence on the HIP
library and on the CPU oracle and compare the results to the parity bar
(coverage, depth-test outcomes and per-pixel primitive order bit-exact,
RGBA8 within 1 LSB, FP32 depth bit-exact)."""
from __future__ import annotations

import numpy as np

from softlight_amd import api


def render_both(gpu_lib, oracle_lib, build, **kw):
    outs = []
    for lib in (gpu_lib, oracle_lib):
        ctx = api.SLContext(lib)
        scene = build(ctx, **kw)
        ctx.reset_stats()
        scene.render()
        out = scene.read()
        out["stats"] = ctx.stats()
        out["scene"] = scene
        outs.append(out)
    return outs


def diff_report(g, o):
    rep = {}
    bits = np.uint16 if g["depth"].dtype.itemsize == 2 else np.uint32
    gd = g["depth"].view(bits)
    od = o["depth"].view(bits)
    rep["depth_mismatch"] = int((gd != od).sum())
    if "last" in g:
        rep["last_mismatch"] = int((g["last"] != o["last"]).sum())
        rep["count_mismatch"] = int((g["count"] != o["count"]).sum())
    if g["color"].dtype == np.uint8:
        dc = np.abs(g["color"].astype(np.int16) - o["color"].astype(np.int16))
        rep["color_max_lsb"] = int(dc.max()) if dc.size else 0
        rep["color_mismatch"] = int((dc > 0).sum())
    else:
        rep["color_max_abs"] = float(np.nanmax(np.abs(g["color"] - o["color"]))) if g["color"].size else 0.0
    for k in ("triangles_submitted", "triangles_binned", "fragments_tested", "fragments_shaded"):
        rep[k] = (g["stats"][k], o["stats"][k])
    return rep


def assert_parity(g, o, color_tol=1, float_tol=None):
    rep = diff_report(g, o)
    assert rep["depth_mismatch"] == 0, rep
    if "last_mismatch" in rep:
        assert rep["last_mismatch"] == 0, rep
        assert rep["count_mismatch"] == 0, rep
    if "color_max_lsb" in rep:
        assert rep["color_max_lsb"] <= color_tol, rep
    elif float_tol is not None:
        assert rep["color_max_abs"] <= float_tol, rep
    for k in ("triangles_submitted", "triangles_binned", "fragments_tested", "fragments_shaded"):
        assert rep[k][0] == rep[k][1], (k, rep)
    return rep
