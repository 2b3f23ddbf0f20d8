"""Raster tile-configuration autotuner.

Times each of the library's raster tile configurations (every index that
sl_get_raster_config accepts: tile size and column segments per row) on a
caller-supplied frame and
keeps the fastest.  Results are cached in a JSON file keyed by the device
name and a workload signature, so later runs reuse them without timing.
The choice never changes results, only speed (the parity tests run every
configuration).
"""
from __future__ import annotations

import ctypes
import json
import os

DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_configs.json")


def num_configs(ctx) -> int:
    """Number of raster configurations the library offers (indices 0..n-1)."""
    n = 0
    while n < 64 and ctx.raster_config(n)[0] == n:
        n += 1
    return n


def device_name() -> str:
    """The HIP device name, or the library's only target when HIP gives none."""
    fallback = "gfx950"
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        try:
            hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
        except OSError:
            return fallback
    buf = ctypes.create_string_buffer(256)
    if hip.hipDeviceGetName(buf, ctypes.c_int(256), ctypes.c_int(0)) != 0:
        return fallback
    return buf.value.decode(errors="replace").strip() or fallback


def load_cache(path: str = DEFAULT_CACHE) -> dict:
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def lookup(signature: str, path: str = DEFAULT_CACHE):
    entry = load_cache(path).get(device_name(), {}).get(signature)
    return None if entry is None else int(entry["config"])


def cached_config(signature: str, path: str = DEFAULT_CACHE):
    """The cached configuration of `signature` for this device, else the one
    cached for any device (the cache may have been written on a box whose HIP
    reported another name), else None."""
    own = lookup(signature, path)
    if own is not None:
        return own
    for entries in load_cache(path).values():
        if signature in entries:
            return int(entries[signature]["config"])
    return None


def tune(ctx, render_frame, signature: str, frames: int = 3, path: str | None = DEFAULT_CACHE) -> dict:
    """Time render_frame() under each raster configuration; return the
    per-configuration raster milliseconds and the best index (also cached)."""
    timings = {}
    for cfg in range(num_configs(ctx)):
        ctx.set_raster_config(cfg)
        render_frame()                     # warm-up (buffer growth, code load)
        ctx.reset_stats()
        for _ in range(frames):
            render_frame()
        st = ctx.stats()
        timings[cfg] = {"raster_ms": st["ms_raster"] / max(st["frames"], 1),
                        "frame_ms": st["ms_total"] / max(st["frames"], 1),
                        "tile": ctx.raster_config(cfg)[1]}
    best = min(timings, key=lambda k: timings[k]["frame_ms"])
    ctx.set_raster_config(best)
    result = {"config": best, "tile": timings[best]["tile"], "timings": timings}
    if path:
        cache = load_cache(path)
        cache.setdefault(device_name(), {})[signature] = {"config": best, "tile": timings[best]["tile"],
                                                          "frame_ms": timings[best]["frame_ms"]}
        with open(path, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    return result
