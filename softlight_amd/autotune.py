"""Raster autotuner.

Times a caller-supplied frame under the library's raster knobs and keeps the
fastest: the tile configuration (every index sl_get_raster_config accepts:
tiles from 8x8 to 64x64 and column segments per row), then for that tile the
occupancy cap (resident raster workgroups per CU, sl_set_raster_occupancy),
then the bin capacity of a raster workgroup (records per part of a heavy
tile's bin, sl_set_raster_split: the library's default policy, never split,
or 64 / 256 / 1024 records), then the raster variant with per-wave depth
rejection on or off (sl_set_raster_depth_reject).  A knob setting the frame cannot use (a tile
configuration with more tiles than the binning's 16384) is recorded as such
and skipped.  Results are cached in a JSON file keyed by the target (gfx950) and a
workload signature, so later runs reuse them without timing.  The choice
never changes results, only speed (the parity tests run every setting).
"""
from __future__ import annotations

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
    """The cache's device key: the library's only target (HIP's device name
    is the generic "AMD Radeon Graphics" on some boxes and unavailable on
    others, which would split one device's entries over two keys)."""
    return "gfx950"


def load_cache(path: str = DEFAULT_CACHE) -> dict:
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


OCCUPANCY = (0, 1, 2, 4, 8)            # raster workgroups per CU (0: no cap)
SPLIT_DEFAULT = 0xFFFFFFFF               # SL_RASTER_SPLIT_DEFAULT
SPLITS = (SPLIT_DEFAULT, 0, 64, 256, 1024)


def lookup(signature: str, path: str = DEFAULT_CACHE):
    entry = load_cache(path).get(device_name(), {}).get(signature)
    return None if entry is None else int(entry["config"])


def cached_entry(signature: str, path: str = DEFAULT_CACHE):
    """The cached knobs of `signature` for this device, else those cached for
    any device (the cache may have been written on a box whose HIP reported
    another name), else None: {"config", "wg_per_cu", "split", ...}."""
    cache = load_cache(path)
    entry = cache.get(device_name(), {}).get(signature)
    if entry is None:
        for entries in cache.values():
            if signature in entries:
                entry = entries[signature]
                break
    if entry is None:
        return None
    out = dict(entry)
    out["config"] = int(out["config"])
    out.setdefault("wg_per_cu", 0)
    out.setdefault("split", SPLIT_DEFAULT)
    out.setdefault("depth_reject", False)
    return out


def cached_config(signature: str, path: str = DEFAULT_CACHE):
    """The cached tile configuration of `signature` (see cached_entry)."""
    entry = cached_entry(signature, path)
    return None if entry is None else entry["config"]


def apply(ctx, entry) -> None:
    """Set a cached entry's knobs on a context."""
    ctx.set_raster_config(int(entry["config"]))
    ctx.set_raster_occupancy(int(entry.get("wg_per_cu", 0)))
    ctx.set_raster_split(int(entry.get("split", SPLIT_DEFAULT)))
    ctx.set_raster_depth_reject(bool(entry.get("depth_reject", False)))


def _time(ctx, render_frame, frames):
    render_frame()                         # warm-up (buffer growth, code load)
    ctx.reset_stats()
    for _ in range(frames):
        render_frame()
    st = ctx.stats()
    n = max(st["frames"], 1)
    return {"raster_ms": st["ms_raster"] / n, "frame_ms": st["ms_total"] / n}


def tune(ctx, render_frame, signature: str, frames: int = 3, path: str | None = DEFAULT_CACHE) -> dict:
    """Time render_frame() under each tile configuration, then each occupancy
    cap, each bin capacity and depth rejection off / on for the best tile
    (with rejection kept, every tile again with it on); set and return the
    best knobs with every timing (also cached)."""
    from softlight_amd.api import SLError

    configs, occupancy, splits, rejects = {}, {}, {}, {}
    ctx.set_raster_occupancy(0)
    ctx.set_raster_split(SPLIT_DEFAULT)
    ctx.set_raster_depth_reject(False)
    for cfg in range(num_configs(ctx)):
        ctx.set_raster_config(cfg)
        try:
            configs[cfg] = {**_time(ctx, render_frame, frames), "tile": ctx.raster_config(cfg)[1]}
        except SLError as e:
            configs[cfg] = {"unsupported": str(e), "tile": ctx.raster_config(cfg)[1]}
    ok = [k for k in configs if "frame_ms" in configs[k]]
    best = min(ok, key=lambda k: configs[k]["frame_ms"])
    ctx.set_raster_config(best)
    for wg in OCCUPANCY:
        ctx.set_raster_occupancy(wg)
        try:
            occupancy[wg] = _time(ctx, render_frame, frames)
        except SLError as e:
            occupancy[wg] = {"unsupported": str(e)}
    wg_best = min((k for k in occupancy if "frame_ms" in occupancy[k]), key=lambda k: occupancy[k]["frame_ms"])
    ctx.set_raster_occupancy(wg_best)
    for sp in SPLITS:
        ctx.set_raster_split(sp)
        try:
            splits[sp] = _time(ctx, render_frame, frames)
        except SLError as e:
            splits[sp] = {"unsupported": str(e)}
    sp_best = min((k for k in splits if "frame_ms" in splits[k]), key=lambda k: splits[k]["frame_ms"])
    ctx.set_raster_split(sp_best)
    rejects[False] = splits[sp_best]
    ctx.set_raster_depth_reject(True)
    rejects[True] = _time(ctx, render_frame, frames)
    # kept only for a clear gain: batches the variant does not cover (the
    # visibility attachment) time the same either way
    rj_best = rejects[True]["frame_ms"] < 0.98 * rejects[False]["frame_ms"]
    ctx.set_raster_depth_reject(rj_best)
    frame_ms = rejects[rj_best]["frame_ms"]
    configs_rj = {}
    if rj_best:
        # rejection changes which tile is fastest: the tiles again with it on
        for cfg in range(num_configs(ctx)):
            ctx.set_raster_config(cfg)
            try:
                configs_rj[cfg] = _time(ctx, render_frame, frames)
            except SLError as e:
                configs_rj[cfg] = {"unsupported": str(e)}
        ok = [k for k in configs_rj if "frame_ms" in configs_rj[k]]
        cfg_rj = min(ok, key=lambda k: configs_rj[k]["frame_ms"])
        if configs_rj[cfg_rj]["frame_ms"] < frame_ms:
            best, frame_ms = cfg_rj, configs_rj[cfg_rj]["frame_ms"]
        ctx.set_raster_config(best)
    entry = {"config": best, "tile": configs[best]["tile"], "wg_per_cu": wg_best, "split": sp_best,
             "depth_reject": rj_best, "frame_ms": frame_ms}
    result = {**entry, "timings": {"config": configs, "wg_per_cu": occupancy,
                                   "split": {("default" if k == SPLIT_DEFAULT else k): v for k, v in splits.items()},
                                   "depth_reject": {("on" if k else "off"): v for k, v in rejects.items()},
                                   "config_with_depth_reject": configs_rj}}
    if path:
        cache = load_cache(path)
        cache.setdefault(device_name(), {})[signature] = entry
        with open(path, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    return result


def describe(entry) -> dict:
    """The knobs of an entry for a report: tile, occupancy cap, bin capacity,
    depth rejection."""
    if entry is None:
        return None
    sp = int(entry.get("split", SPLIT_DEFAULT))
    return {"config": int(entry["config"]), "tile": list(entry.get("tile", [])),
            "workgroups_per_cu": int(entry.get("wg_per_cu", 0)) or "uncapped",
            "bin_capacity_records_per_workgroup": ("default (split point/line batches at 256)" if sp == SPLIT_DEFAULT
                                                   else ("unsplit" if sp == 0 else sp)),
            "depth_reject": bool(entry.get("depth_reject", False))}
