"""Build the gfx950 HIP library softlight_amd/_build/libsoftlight_amd.so.

Run as `python -m softlight_amd.build` or through __graft_entry__.build().
Every HIP source is compiled for gfx950 only, with floating-point
contraction off (the draw path's arithmetic order is part of its
specification) and packed-FP32 instruction selection disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libsoftlight_amd.so")
SOURCES = ["sl_kernels.hip", "sl_backward.hip", "sl_smooth.hip", "sl_runtime.hip",
           # k_raster instances, split so that they compile in parallel
           "sl_raster_plain.hip", "sl_raster_plain_dbg.hip", "sl_raster_modes.hip", "sl_raster_modes_dbg.hip",
           "sl_raster_mrt.hip", "sl_raster_vis.hip", "sl_raster_reject.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed on {cmd[-1]}")
    return r


def build(verbose: bool = False) -> str:
    os.makedirs(OUT, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(OUT, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC, *CFLAGS, "-c", "-o", obj, os.path.join(CSRC, src)])
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        for r in ex.map(_run, cmds):
            if verbose and r.stderr:
                sys.stderr.write(r.stderr)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    # the C++ facade example: a host program written against
    # include/softlight_amd/SL_Context.hpp, linked to the library
    ex = os.path.join(OUT, "facade_draw")
    _run(["g++", "-std=c++17", "-O1", "-I", os.path.join(HERE, "..", "include"), "-o", ex + ".tmp",
          os.path.join(HERE, "examples", "facade_draw.cpp"), "-L", OUT, "-lsoftlight_amd",
          f"-Wl,-rpath,$ORIGIN"])
    os.replace(ex + ".tmp", ex)
    # the scene-file example (include/softlight_amd/SL_SceneFileLoader.hpp):
    # the full program, and a host-only build of its graph dump
    inc = ["-std=c++17", "-O2", "-ffp-contract=off", "-I", os.path.join(HERE, "..", "include")]
    src = os.path.join(HERE, "examples", "scene_graph_draw.cpp")
    for name, extra in (("scene_graph_draw", ["-L", OUT, "-lsoftlight_amd", "-Wl,-rpath,$ORIGIN"]),
                        ("scene_graph_dump", ["-DSL_SCENE_DUMP_ONLY"])):
        exe = os.path.join(OUT, name)
        _run(["g++", *inc, "-o", exe + ".tmp", src, *extra])
        os.replace(exe + ".tmp", exe)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
