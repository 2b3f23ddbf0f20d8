"""Benchmark of the MI355X draw path on the textured-sphere workload at 1080p.

A step is one frame: clear + draw_instanced of a 256x256-segment UV sphere
(130,560 indexed u32 triangles) as 64 instances at random positions in a
40-unit cube, 60 deg perspective from the origin, bilinear REPEAT sampling of
a 1024x1024 RGBA8 value-noise texture, Lambert lighting, 1920x1080 RGBA8 +
R32F targets, depth LESS, back-face culling; inputs resident in HBM before
timing starts.  Prints one JSON line:

  value      submitted triangles per second of whole frames (Mtri/s, wall
             clock around `steps` back-to-back replays of the frame captured as
             a HIP graph, synchronised on both sides); eager frames (each with
             a host round trip after setup) give the per-stage device times
  mtri_per_s_stage    submitted triangles / (vertex+setup+binning device time)
  shaded_gfrag_per_s  shaded fragments / fragment-stage device time
  roofline   the setup kernel (the dominant kernel) against the 8 TB/s HBM
             roofline, with algorithmic bytes = 28.2 B per submitted
             triangle (12 B of u32 indices + 32 B vertex x 66,049/130,560)
  cpu_baseline        the CPU oracle (multithreaded, all host cores) on a
             bounded number of frames of the same workload

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
WORKLOAD = dict(seed=1, n_instances=64, stacks=256, sectors=256, W=1920, H=1080, tex_size=1024, radius=2.0,
                extent=20.0)
RANDOM_WORKLOAD = dict(seed=0, ntri=4096, W=640, H=480)
BATCHED_WORKLOAD = dict(seed=3, n_meshes=1000)
VIEWS_WORKLOAD = dict(seed=4, level=5, views=32, size=512, views_per_fbo=8, tex_size=512)
SMOOTH_WORKLOAD = dict(seed=6, meshes=8, min_tris=5000, max_tris=60000, views=64, size=512, views_per_fbo=8,
                       tex_size=512)
SMOOTH_SIGMAS = (1e-2, 1e-3, 1e-4)
RASTER_SPLIT = None          # --raster-split: records per part of a split heavy tile (None: the library's default)
CALIBRATION_BYTES = 256 << 20


def calibration_copy(repeats=3):
    """A plain device-to-device copy of CALIBRATION_BYTES (the copy kernel the
    HBM byte counters are scaled against)."""
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    n = ctypes.c_size_t(CALIBRATION_BYTES)
    assert hip.hipMalloc(ctypes.byref(a), n) == 0 and hip.hipMalloc(ctypes.byref(b), n) == 0
    hip.hipMemset(a, 1, n)
    for _ in range(repeats):
        assert hip.hipMemcpy(b, a, n, 3) == 0          # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


def algorithmic_bytes_per_triangle(info) -> float:
    tpi = info["triangles_per_instance"]
    vpi = info["vertices_per_instance"]
    return 12.0 + 32.0 * vpi / tpi


def frag_bytes(tested, shaded, texel_bytes=16.0) -> float:
    # depth read per tested sample; depth write + RGBA8 write (+ 4 RGBA8 texels) per shaded sample
    return 4.0 * tested + (4.0 + 4.0 + texel_bytes) * shaded


def run_gpu(steps, warmup, workload="spheres", raster_config=None, tune=False, raster_split=None):
    import softlight_amd
    from softlight_amd import api, autotune, scenes

    ctx = api.SLContext(softlight_amd.load_library())
    if raster_split is None:
        raster_split = RASTER_SPLIT
    if raster_split is not None:
        ctx.set_raster_split(raster_split)
    if workload == "spheres":
        scene = scenes.build_sphere_scene(ctx, **WORKLOAD)
    elif workload in RENDER_MODES:
        scene = scenes.build_sphere_scene(ctx, **WORKLOAD, **RENDER_MODES[workload])
    elif workload == "spheres_rgb8_half":
        scene = scenes.build_sphere_scene(ctx, **WORKLOAD, color_type=api.SL_COLOR_RGB_8U, depth_type=api.SL_COLOR_R_HALF)
    elif workload == "batched":
        scene = scenes.build_batched_scene(ctx, **BATCHED_WORKLOAD)
    else:
        scene = scenes.build_random_triangles(ctx, **RANDOM_WORKLOAD)
    signature = f"{workload}:{scene.W}x{scene.H}:{scene.n_triangles}"

    def frame():
        scene.render()
        ctx.finish()

    if tune:
        res = autotune.tune(ctx, frame, signature, path=os.environ.get("SL_TUNE_CACHE", autotune.DEFAULT_CACHE))
        print(json.dumps({"autotune": signature, **res}), file=sys.stderr, flush=True)
    elif raster_config is not None:
        ctx.set_raster_config(raster_config)
    else:
        cache = os.environ.get("SL_TUNE_CACHE", autotune.DEFAULT_CACHE)
        entry = autotune.cached_entry(signature, cache)
        if entry is not None:
            autotune.apply(ctx, entry)
            if raster_split is not None:
                ctx.set_raster_split(raster_split)
        elif workload.startswith("spheres_"):
            # the sphere frame's variants (other targets, other render modes)
            # bin the same geometry: start from the sphere frame's tile
            cached = autotune.cached_config(f"spheres:{scene.W}x{scene.H}:{scene.n_triangles}", cache)
            if cached is not None:
                ctx.set_raster_config(cached)
        scene.info["tuned"] = autotune.describe(entry)
    scene.info["raster_config"] = ctx.raster_config()
    for _ in range(warmup):
        frame()
    # eager frames: per-stage device times
    ctx.reset_stats()
    ctx.finish()
    t0 = time.perf_counter()
    for _ in range(steps):
        frame()
    t1 = time.perf_counter()
    st = ctx.stats()
    eager = (t1 - t0)
    # captured frames: the timed measurement
    scene.render()
    ctx.frame_graph_capture()
    ctx.finish()
    ctx.reset_stats()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.frame_graph_launch()
    ctx.finish()
    t1 = time.perf_counter()
    gst = ctx.stats()
    ctx.frame_graph_release()
    return scene, st, eager, gst, (t1 - t0)


def run_views(steps, warmup, tune=False):
    """Differentiable views: 32 cameras at 512x512 of a level-5 icosphere
    (20,480 triangles) with per-vertex colours and a 512x512 RGBA32F texture.
    Forward = clears + 32 draws (4 framebuffers of 8 viewports); backward =
    loss gradient + gradient zeroing + the fixed-visibility adjoint."""
    import softlight_amd
    from softlight_amd import api, autotune, scenes

    ctx = api.SLContext(softlight_amd.load_library())
    sc = scenes.DiffScene(ctx, **VIEWS_WORKLOAD)
    sc.make_targets()

    def fwd():
        sc.forward()
        ctx.finish()

    signature = f"views:{sc.views}x{sc.size}:{sc.n_triangles}"
    cache = os.environ.get("SL_TUNE_CACHE", autotune.DEFAULT_CACHE)
    if tune:
        res = autotune.tune(ctx, fwd, signature, path=cache)
        print(json.dumps({"autotune": signature, **res}), file=sys.stderr, flush=True)
    else:
        entry = autotune.cached_entry(signature, cache)
        if entry is not None:
            autotune.apply(ctx, entry)

    def fwd_bwd():
        sc.forward()
        sc.loss_grad_images()
        sc.backward()
        ctx.finish()

    out = {}
    for name, fn in (("fwd", fwd), ("fwd_bwd", fwd_bwd)):
        for _ in range(warmup):
            fn()
        ctx.reset_stats()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        t1 = time.perf_counter()
        st = ctx.stats()
        out[name] = dict(ms=(t1 - t0) * 1e3 / steps, views_per_s=sc.views * steps / (t1 - t0),
                         device_ms_forward=st["ms_total"] / steps, device_ms_backward=st["ms_backward"] / steps)
    out["raster_config"] = ctx.raster_config()
    return sc, out


def device_used_bytes():
    """Device memory in use on device 0 (hipMemGetInfo total - free), from the
    HIP runtime the library itself runs on."""
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    rc = hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total))
    if rc != 0:
        raise RuntimeError(f"hipMemGetInfo failed ({rc})")
    return total.value - free.value


def run_smooth(steps, warmup, tune=False):
    """Smoothed span coverage: 64 views at 512x512 of 8 meshes (5k..60k
    triangles, log-uniform), per-vertex RGB + 512x512 RGBA32F texture.
    Forward = clears + 512 draws (8 framebuffers of 8 viewports) + the
    smoothing pass; backward = loss gradient on the smoothed images, smoothing
    adjoint (image gradient + vertex screen-position gradients) and the
    fixed-visibility adjoint to colours and texels.  Per sigma; plus the
    sigma = 0 check against the hard image."""
    import numpy as np
    import softlight_amd
    from softlight_amd import api, scenes

    from softlight_amd import autotune

    base = device_used_bytes()
    ctx = api.SLContext(softlight_amd.load_library())
    sc = scenes.SmoothScene(ctx, **SMOOTH_WORKLOAD)
    # the raster knobs: timed on the forward at the first sigma (--tune), or
    # the cached entry of this workload
    signature = f"smooth:{sc.views}x{sc.size}:{sc.n_triangles}"
    cache = os.environ.get("SL_TUNE_CACHE", autotune.DEFAULT_CACHE)
    if tune:
        sc.make_targets(SMOOTH_SIGMAS[0])

        def fwd():
            sc.forward(SMOOTH_SIGMAS[0])
            ctx.finish()

        res = autotune.tune(ctx, fwd, signature, path=cache)
        print(json.dumps({"autotune": signature, **res}), file=sys.stderr, flush=True)
        entry = autotune.cached_entry(signature, cache)
    else:
        entry = autotune.cached_entry(signature, cache)
        if entry is not None:
            autotune.apply(ctx, entry)
    sc.forward(0.0)
    ctx.finish()
    t0 = sc.targets[0]
    hard_equal = bool(np.array_equal(ctx.texture_read(t0["smooth"]).view(np.uint32),
                                     ctx.texture_read(t0["color"]).view(np.uint32)))
    res = {}
    peak = device_used_bytes() - base
    for sigma in SMOOTH_SIGMAS:
        sc.make_targets(sigma)
        row = {}
        for name, fn in (("fwd", lambda: sc.forward(sigma)), ("fwd_bwd", lambda: (sc.forward(sigma), sc.backward(sigma)))):
            for _ in range(warmup):
                fn()
            ctx.finish()
            t = time.perf_counter()
            for _ in range(steps):
                fn()
            ctx.finish()
            dt = (time.perf_counter() - t) / steps
            row[name] = dict(ms=dt * 1e3, views_per_s=sc.views / dt)
            peak = max(peak, device_used_bytes() - base)
        ctx.finish()
        pg = ctx.vbo_read(t0["pos_grad"], len(t0["views"]) * len(sc.meshes) * sc.n_vertices_max * 8, dtype=np.float32)
        res[f"{sigma:g}"] = {"fwd_views_per_s": round(row["fwd"]["views_per_s"], 1),
                             "fwd_bwd_views_per_s": round(row["fwd_bwd"]["views_per_s"], 1),
                             "fwd_bwd_over_fwd_time": round(row["fwd_bwd"]["ms"] / row["fwd"]["ms"], 3),
                             "ms_fwd": round(row["fwd"]["ms"], 3), "ms_fwd_bwd": round(row["fwd_bwd"]["ms"], 3),
                             "nonzero_position_grads": int(np.count_nonzero(pg))}
    ctx_tile = ctx.raster_config()[1]
    ctx.close()
    shared = run_smooth_shared(steps, warmup, entry)
    return {"workload": f"{sc.views} views x {sc.size}^2 of {len(sc.meshes)} UV spheres "
                        f"({min(sc.mesh_tris)}..{max(sc.mesh_tris)} triangles, {sc.n_triangles} in all), "
                        "per-vertex RGB + 512^2 RGBA32F texture, L2 loss on the smoothed images against targets "
                        "rendered with vertices displaced by N(0, 0.02)",
            "sigma_0_equals_hard_image": hard_equal,
            "raster_tile": list(ctx_tile),
            "autotuned": autotune.describe(entry),
            "per_sigma": res,
            "peak_device_bytes": int(peak),
            "shared_visibility": shared}


def run_smooth_shared(steps, warmup, entry):
    """The smoothed-coverage leg at the first sigma with one visibility
    attachment shared by the 8 framebuffers: visibility recomputed (each
    backward pass draws its framebuffer again) instead of kept per
    framebuffer; views/s and the device-memory delta beside the kept case."""
    import softlight_amd
    from softlight_amd import api, autotune, scenes

    sigma = SMOOTH_SIGMAS[0]
    base = device_used_bytes()
    ctx = api.SLContext(softlight_amd.load_library())
    if entry is not None:
        autotune.apply(ctx, entry)
    sc = scenes.SmoothScene(ctx, **SMOOTH_WORKLOAD, shared_vis=True)
    sc.make_targets(sigma)
    peak = device_used_bytes() - base
    row = {}
    for name, fn in (("fwd", lambda: sc.forward(sigma)), ("fwd_bwd", lambda: (sc.forward(sigma), sc.backward(sigma)))):
        for _ in range(warmup):
            fn()
        ctx.finish()
        ctx.reset_stats()
        t = time.perf_counter()
        for _ in range(steps):
            fn()
        ctx.finish()
        dt = (time.perf_counter() - t) / steps
        row[name] = dict(ms=dt * 1e3, views_per_s=sc.views / dt, rebuilds=ctx.stats()["visibility_rebuilds"] / steps)
        peak = max(peak, device_used_bytes() - base)
    ctx.close()
    return {"sigma": sigma, "fwd_views_per_s": round(row["fwd"]["views_per_s"], 1),
            "fwd_bwd_views_per_s": round(row["fwd_bwd"]["views_per_s"], 1),
            "fwd_bwd_over_fwd_time": round(row["fwd_bwd"]["ms"] / row["fwd"]["ms"], 3),
            "visibility_rebuilds_per_step": row["fwd_bwd"]["rebuilds"],
            "visibility_bytes": 16 * sc.targets[0]["W"] * sc.targets[0]["H"],
            "peak_device_bytes": int(peak)}


# the sphere frame's index buffer read in the other render modes
RENDER_MODES = {"spheres_wire": dict(wire=True), "spheres_lines": dict(prim="lines"),
                "spheres_points": dict(prim="points")}


def run_demo_targets(steps, warmup):
    """The 1080p sphere frame into the reference demos' targets: RGB8 colour
    and R16F depth (sl_large_scene_test.cpp:778-782)."""
    scene, st, eager, gst, graph_s = run_gpu(steps, warmup, workload="spheres_rgb8_half")
    f = max(1, st["frames"])
    return {"targets": "RGB8 colour + R16F depth", "ms_per_frame": round(graph_s * 1e3 / steps, 4),
            "Mtri_per_s": round(scene.n_triangles * steps / graph_s / 1e6, 1),
            "raster_ms": round(st["ms_raster"] / f, 4),
            "fragments_shaded_per_frame": int(st["fragments_shaded"] // f)}


def run_scene_file(frames=12, n_objects=120, level=3, cpu=True):
    """Host ingestion: a scene file of `n_objects` textured icospheres
    (20*4^level triangles each) loaded by the C++ scene-file loader into a
    scene graph, then per frame: a root edit, the graph update, the draw-list
    build, one deferred draw per item and one synchronisation
    (softlight_amd/examples/scene_graph_draw.cpp)."""
    import subprocess
    import tempfile

    import softlight_amd
    from softlight_amd import scenes
    exe = os.path.join(os.path.dirname(softlight_amd.LIB_PATH), "scene_graph_draw")
    if not os.path.exists(exe):
        return {"skipped": "scene_graph_draw not built"}
    with tempfile.TemporaryDirectory() as d:
        obj, counts = scenes.write_icosphere_scene(d, n_objects=n_objects, level=level, extent=12.0)
        cmds = os.path.join(d, "nodes.txt")
        with open(cmds, "w") as f:
            f.write("")
        r = subprocess.run([exe, "render", obj, cmds, "1920", "1080", os.path.join(d, "c.bin"), os.path.join(d, "d.bin"),
                            str(frames)], capture_output=True, text=True, timeout=300)
        if r.returncode != 0:
            raise RuntimeError(f"scene_graph_draw failed ({r.returncode}): {r.stderr.strip()}")
        out = json.loads(r.stdout.strip().splitlines()[-1])
        out["file_MB"] = round(os.path.getsize(obj) / 2 ** 20, 1)
        out["Mtri_per_s"] = round(out["triangles"] / (out["frame_ms"] * 1e-3) / 1e6, 1)
        if cpu:
            # the same frame drawn by the oracle's thread pool from the dumped draw list
            import oracle
            from softlight_amd import scene_files
            dump = scene_files.run_dump(scene_files.dump_executable(), obj, "", d)
            threads = host_cpu()["threads"]
            times = sorted(scene_files.draw_dump(oracle.load(), dump, 1920, 1080, threads=threads)["ms"] for _ in range(5))
            ms = times[2]
            out["cpu_baseline"] = {"ms_per_frame": round(ms, 2), "Mtri_per_s": round(out["triangles"] / (ms * 1e-3) / 1e6, 1),
                                   "cores": threads, "kind": "port", "sample": "the same frame, median of 5"}
    return out


def run_render_modes(steps, warmup):
    """The 1080p sphere frame with its index buffer drawn as wireframe
    triangles, lines (index pairs) and points (every index): per-frame time
    of captured-graph replays and primitives per second."""
    out = {}
    for name, kw in RENDER_MODES.items():
        scene, st, eager, gst, graph_s = run_gpu(steps, warmup, workload=name)
        info = scene.info
        per = {"spheres_wire": 3, "spheres_lines": 2, "spheres_points": 1}[name]
        prims = info["instances"] * (3 * info["triangles_per_instance"] // per)
        f = max(1, st["frames"])
        out[name[len("spheres_"):]] = {
            "primitives_per_frame": prims,
            "ms_per_frame": round(graph_s * 1e3 / steps, 4),
            "Mprim_per_s": round(prims * steps / graph_s / 1e6, 1),
            "stage_ms": {"setup_kernel": round(st["ms_setup_kernel"] / f, 4), "binning": round(st["ms_bin"] / f, 4),
                         "raster": round(st["ms_raster"] / f, 4)},
            "fragments_tested_per_frame": int(st["fragments_tested"] // f),
            "fragments_shaded_per_frame": int(st["fragments_shaded"] // f),
        }
    return out


def run_blit(reps=50):
    """Present: blit of a 1920x1080 RGBA8 colour buffer to an RGBA8 back-buffer
    of the same size (SL_Context::blit), reps times back to back."""
    import softlight_amd
    from softlight_amd import api, scenes

    ctx = api.SLContext(softlight_amd.load_library())
    W, H = 1920, 1080
    src = ctx.create_texture(api.SL_COLOR_RGBA_8U, W, H, 1, scenes.uniform_u8(11, W * H * 4).reshape(H, W, 4))
    dst = ctx.create_texture(api.SL_COLOR_RGBA_8U, W, H)
    ctx.blit(dst, src)
    ctx.finish()
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.blit(dst, src)
    ctx.finish()
    dt = (time.perf_counter() - t0) / reps
    gbs = W * H * 8 / dt / 1e9
    return {"workload": "1920x1080 RGBA8 -> RGBA8 present blit", "us_per_blit": round(dt * 1e6, 2),
            "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_blit": W * H * 8},
            "timing": "wall clock over back-to-back blits (launch overhead included)"}


def host_cpu():
    """The host's online cores (nproc), the cores this process may run on, the
    CPU model, and the thread count the CPU legs use: the host-core share of
    one GPU (16 on these boxes), or fewer when fewer are online."""
    online = os.cpu_count() or 1
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = online
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": online, "usable_cores": usable, "model": model, "threads": min(usable, 16)}


def _median_frame_ms(ctx, render, frames):
    render()                             # warm-up frame
    ctx.finish()
    times = []
    for _ in range(frames):
        t0 = time.perf_counter()
        render()
        ctx.finish()
        times.append(time.perf_counter() - t0)
    times.sort()
    return times[len(times) // 2] * 1e3


def run_cpu_baseline(frames=5):
    """The oracle (the project's restatement of the reference draw path) on
    the host, multithreaded: the sphere frame (the headline's CPU figure)."""
    import oracle
    from softlight_amd import api, scenes

    cpu = host_cpu()
    ctx = api.SLContext(oracle.load())
    ctx.num_threads(cpu["threads"])
    scene = scenes.build_sphere_scene(ctx, **WORKLOAD)
    med = _median_frame_ms(ctx, scene.render, frames)
    ctx.close()
    return {
        "value": round(scene.n_triangles / (med * 1e-3) / 1e6, 3),
        "unit": "Mtri/s",
        "cores": cpu["threads"],
        "kind": "port",
        "sample": f"median of {frames} frames (+1 warm-up) of the same 1080p textured-sphere workload: {med:.1f} ms; "
                  f"{cpu['threads']} threads (the host-core share of one GPU) of {cpu['nproc']} online "
                  f"({cpu['model']})",
        "host": cpu,
    }


def run_cpu_baselines(frames=5):
    """The oracle on the other workloads: 4096 random triangles at 640x480 and
    the 1000-mesh batched frame (median of `frames` frames, multithreaded),
    and the differentiable workloads' forward and forward+backward (the scalar
    adjoint) at one thread and multithreaded: all 32 views of the
    fixed-visibility workload, and one framebuffer (8 of the 64 views) of the
    smoothed-coverage workload at sigma = 1e-3."""
    import oracle
    from softlight_amd import api, scenes

    cpu = host_cpu()
    out = {"host": cpu}
    for name, build, kw in (("random_640x480", scenes.build_random_triangles, RANDOM_WORKLOAD),
                            ("batched_1080p", scenes.build_batched_scene, BATCHED_WORKLOAD)):
        ctx = api.SLContext(oracle.load())
        ctx.num_threads(cpu["threads"])
        scene = build(ctx, **kw)
        med = _median_frame_ms(ctx, scene.render, frames)
        ctx.close()
        out[name] = {"ms_per_frame": round(med, 2), "Mtri_per_s": round(scene.n_triangles / (med * 1e-3) / 1e6, 3),
                     "threads": cpu["threads"], "sample": f"median of {frames} frames"}

    def diff_leg(make, fwd, fwd_bwd, views, threads):
        ctx = api.SLContext(oracle.load())
        ctx.num_threads(threads)
        sc = make(ctx)
        res = {}
        for name, fn in (("fwd", fwd), ("fwd_bwd", fwd_bwd)):
            ms = _median_frame_ms(ctx, lambda: fn(sc), 3)
            res[name] = ms
        ctx.close()
        return {"threads": threads, "ms_fwd": round(res["fwd"], 2), "ms_fwd_bwd": round(res["fwd_bwd"], 2),
                "fwd_views_per_s": round(views / (res["fwd"] * 1e-3), 2),
                "fwd_bwd_views_per_s": round(views / (res["fwd_bwd"] * 1e-3), 2),
                "fwd_bwd_over_fwd_time": round(res["fwd_bwd"] / res["fwd"], 3)}

    def make_views(ctx):
        sc = scenes.DiffScene(ctx, **VIEWS_WORKLOAD)
        sc.make_targets()
        return sc

    smooth_kw = dict(SMOOTH_WORKLOAD, views=8)

    def make_smooth(ctx):
        sc = scenes.SmoothScene(ctx, **smooth_kw)
        sc.make_targets(1e-3)
        return sc

    out["views_32x512"] = {
        "workload": "the fixed-visibility workload: 32 views x 512^2, forward and forward + scalar adjoint, median of 3",
        "per_threads": [diff_leg(make_views, lambda s: s.forward(),
                                 lambda s: (s.forward(), s.loss_grad_images(), s.backward()), VIEWS_WORKLOAD["views"], t)
                        for t in (1, cpu["threads"])]}
    out["smooth_8_of_64_views"] = {
        "workload": "one framebuffer (8 of the 64 views) of the smoothed-coverage workload at sigma = 1e-3, forward and "
                    "forward + scalar adjoint, median of 3",
        "per_threads": [diff_leg(make_smooth, lambda s: s.forward(1e-3), lambda s: (s.forward(1e-3), s.backward(1e-3)),
                                 smooth_kw["views"], t)
                        for t in (1, cpu["threads"])]}
    return out


def run_workload_leg(workload, steps, warmup):
    """One of the other frames (4096 random triangles at 640x480, the
    1000-mesh batched frame at 1080p) on its tuned raster knobs: graph-replay
    frame time, triangles/s, shaded fragments/s and the stage split."""
    scene, st, eager_s, gst, elapsed = run_gpu(steps, warmup, workload)
    tris = scene.n_triangles
    frames = max(st["frames"], 1)
    ms_raster = st["ms_raster"] / frames
    if gst["raster_launches"] > 0:
        ms_raster = gst["ms_raster"] / gst["raster_launches"]
    shaded = st["fragments_shaded"] / frames
    tested = st["fragments_tested"] / frames
    texel_bytes = 0.0 if workload == "random" else 16.0
    fbytes = frag_bytes(tested, shaded, texel_bytes)
    frag_gbs = fbytes / (ms_raster * 1e-3) / 1e9
    return {
        "triangles_per_frame": tris,
        "ms_per_frame": round(elapsed * 1e3 / steps, 4),
        "Mtri_per_s": round(tris * steps / elapsed / 1e6, 3),
        "shaded_gfrag_per_s": round(shaded / (ms_raster * 1e-3) / 1e9, 4),
        "fragments_shaded_per_frame": int(shaded),
        "stage_ms": {"vertex_setup": round(st["ms_vertex"] / frames, 4), "binning": round(st["ms_bin"] / frames, 4),
                     "raster": round(ms_raster, 4)},
        "fragment_roofline": {"kernel": "k_raster", "bound": "hbm", "achieved": round(frag_gbs, 2),
                              "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(frag_gbs / HBM_PEAK_GBS, 5),
                              "bytes_per_frame": fbytes},
        "raster_tile": list(scene.info["raster_config"][1]),
        "autotuned": scene.info.get("tuned"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--headline-only", action="store_true",
                    help="spheres: only the headline frame (no differentiable, smoothing, render-mode, target or blit legs)")
    ap.add_argument("--workload", choices=["spheres", "random", "batched", "views", "smooth", "modes", "scene_file"],
                    default="spheres",
                    help="spheres: the 1080p textured-sphere frame (default); random: 4096 random triangles at "
                         "640x480; batched: 1000 meshes with mixed opaque/blended materials at 1080p; views: "
                         "32 differentiable 512x512 views, forward and forward+backward; smooth: the smoothed-"
                         "coverage leg (64 views of 8 meshes, each sigma); modes: the sphere frame's "
                         "index buffer drawn as wireframe triangles, lines and points (its leg only); scene_file: "
                         "a 120-object OBJ file loaded into the scene graph and drawn per frame (its leg only)")
    ap.add_argument("--calibrate", action="store_true", help="also run a 256 MiB device copy (counter calibration)")
    ap.add_argument("--raster-config", type=int, default=None, help="raster tile configuration 0..11 (default: tuned)")
    ap.add_argument("--tune", action="store_true",
                    help="time the raster knobs first (tile, workgroups per CU, bin capacity) and cache the best")
    ap.add_argument("--raster-split", type=int, default=None,
                    help="records per part of a split heavy raster tile (0: no split; default: the library's)")
    args = ap.parse_args()
    global RASTER_SPLIT
    RASTER_SPLIT = args.raster_split
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported (a single-GPU box)",
              file=sys.stderr)
        sys.exit(2)

    if args.calibrate:
        calibration_copy()
    if args.workload == "views":
        main_views(args)
        return
    if args.workload == "smooth":
        print(json.dumps({"smoothed_coverage": run_smooth(max(2, args.steps // 8), 1, args.tune)}), flush=True)
        return
    if args.workload == "modes":
        print(json.dumps({"render_modes": run_render_modes(args.steps, args.warmup),
                          "demo_targets": run_demo_targets(args.steps, args.warmup)}), flush=True)
        return
    if args.workload == "scene_file":
        print(json.dumps({"scene_file": run_scene_file(max(2, args.steps), cpu=not args.no_cpu)}), flush=True)
        return
    scene, st, eager_s, gst, elapsed = run_gpu(args.steps, args.warmup, args.workload, args.raster_config, args.tune,
                                               args.raster_split)
    tris = scene.n_triangles
    frames = max(st["frames"], 1)
    ms_step = elapsed * 1e3 / args.steps
    value = tris * args.steps / elapsed / 1e6

    ms_vertex = st["ms_vertex"] / frames
    ms_bin = st["ms_bin"] / frames
    ms_raster = st["ms_raster"] / frames
    ms_setup = st["ms_setup_kernel"] / max(st["setup_launches"], 1)
    timing_source = "eager frames (HIP events around each kernel)"
    if gst["setup_launches"] > 0 and gst["raster_launches"] > 0:
        # event nodes inside the replayed graph: the timed region itself
        ms_setup = gst["ms_setup_kernel"] / gst["setup_launches"]
        ms_raster = gst["ms_raster"] / gst["raster_launches"]
        timing_source = "graph replays in the timed region (HIP event nodes around k_setup and k_raster)"
    shaded = st["fragments_shaded"] / frames
    tested = st["fragments_tested"] / frames

    if args.workload == "spheres":
        bpt = algorithmic_bytes_per_triangle(scene.info)
    elif args.workload == "batched":
        bpt = 12.0 + 32.0 * 0.5          # u32 indices + ~1/2 vertex (32 B) per triangle of an icosphere
    else:
        bpt = 48.0
    achieved = tris * bpt / (ms_setup * 1e-3) / 1e9        # GB/s
    texel_bytes = 0.0 if args.workload == "random" else 16.0
    fbytes = frag_bytes(tested, shaded, texel_bytes)
    frag_gbs = fbytes / (ms_raster * 1e-3) / 1e9

    out = {
        "metric": "Mtri/s + shaded Gfrag/s at 1080p; fwd+bwd views/s vs plain fwd, 1 MI355X",
        "value": round(value, 3),
        "unit": "Mtri/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded UV spheres, value-noise texture)",
        "config": {
            "workload": ("textured UV-sphere scene: 64 instances x 130,560 indexed triangles, 1920x1080 RGBA8+R32F, "
                         "1024^2 RGBA8 bilinear REPEAT, Lambert, depth LESS, back-face cull")
            if args.workload == "spheres" else
            ("1000 displaced icospheres, log-normal triangle counts (median 2000), 20% alpha-blended, "
             "16 RGBA8 textures, 1920x1080, one batch" if args.workload == "batched" else
             "4096 random non-indexed triangles, 640x480 RGBA8+R32F, Gouraud, depth LESS, no cull"),
            "triangles_per_frame": tris,
            "W": scene.W, "H": scene.H,
            "raster_tile": {"config": scene.info["raster_config"][0], "tile_w": scene.info["raster_config"][1][0],
                            "tile_h": scene.info["raster_config"][1][1], "segments": scene.info["raster_config"][1][2]},
        },
        "frame_mode": "hipGraph replay of the captured frame (clear + draws), back to back",
        "eager_ms_per_frame": round(eager_s * 1e3 / args.steps, 4),
        "graph_device_ms_per_frame": round(gst["ms_total"] / max(gst["frames"], 1), 4),
        "mtri_per_s_stage": round(tris / ((ms_vertex + ms_bin) * 1e-3) / 1e6, 3),
        "shaded_gfrag_per_s": round(shaded / (ms_raster * 1e-3) / 1e9, 4),
        "fragments_shaded_per_frame": int(shaded),
        "fragments_tested_per_frame": int(tested),
        "triangles_binned_per_frame": int(st["triangles_binned"] / frames),
        "stage_ms": {"vertex_setup": round(ms_vertex, 4), "binning": round(ms_bin, 4), "raster": round(ms_raster, 4),
                     "setup_kernel": round(ms_setup, 4)},
        "kernel_timing": timing_source,
        "fragment_roofline": {"kernel": "k_raster", "bound": "hbm", "achieved": round(frag_gbs, 2), "peak": HBM_PEAK_GBS,
                              "unit": "GB/s", "frac": round(frag_gbs / HBM_PEAK_GBS, 5),
                              "bytes_per_frame": fbytes},
        "roofline": {
            "kernel": "k_setup",
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": measured_traffic(args.workload),
            "bytes_per_triangle": round(bpt, 3),
            "bytes_counted": "per submitted triangle (each instance's own index and attribute fetches)",
            "unique_input_bytes": unique_input_bytes(scene.info) if args.workload == "spheres" else None,
        },
    }
    if args.workload == "spheres" and not args.headline_only:
        _, vr = run_views(max(3, args.steps // 4), 1)
        out["differentiable"] = views_summary(vr)
        out["smoothed_coverage"] = run_smooth(max(2, args.steps // 8), 1)
        out["render_modes"] = run_render_modes(max(3, args.steps // 4), 1)
        out["demo_targets"] = run_demo_targets(max(3, args.steps // 2), 1)
        out["present_blit"] = run_blit()
        out["scene_file"] = run_scene_file(cpu=not args.no_cpu)
        out["random_640x480"] = run_workload_leg("random", max(3, args.steps // 2), 1)
        out["batched_1080p"] = run_workload_leg("batched", max(3, args.steps // 2), 1)
    if not args.no_cpu and args.workload == "spheres":
        out["cpu_baseline"] = run_cpu_baseline()
        if not args.headline_only:
            out["cpu_baselines"] = run_cpu_baselines()
    print(json.dumps(out), flush=True)


def unique_input_bytes(info):
    """The sphere frame's distinct input bytes per launch: the shared vertex
    and index buffers once plus the per-instance data (64 B a transform)."""
    return int(info["vertices_per_instance"] * 32 + info["triangles_per_instance"] * 12 + info["instances"] * 64)


def measured_traffic(workload):
    """HBM bytes per k_setup launch from the committed rocprofv3 counter runs
    (FETCH_SIZE and WRITE_SIZE scaled by a 256 MiB copy), or None."""
    if workload != "spheres":
        return None
    path = os.path.join(ROOT, "profiles", "k_setup_spheres_1080p_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("traffic_bytes_per_launch")


def views_summary(vr):
    f, fb = vr["fwd"], vr["fwd_bwd"]
    return {
        "workload": "32 views x 512^2 of a 20,480-triangle icosphere, per-vertex RGB + 512^2 RGBA32F bilinear texture, "
                    "L2 loss, fixed-visibility adjoint to vertex colours and texels",
        "fwd_views_per_s": round(f["views_per_s"], 1),
        "fwd_bwd_views_per_s": round(fb["views_per_s"], 1),
        "fwd_bwd_over_fwd_time": round(fb["ms"] / f["ms"], 3),
        "ms_fwd": round(f["ms"], 4),
        "ms_fwd_bwd": round(fb["ms"], 4),
        "device_ms_backward": round(fb["device_ms_backward"], 4),
        "device_ms_forward": round(f["device_ms_forward"], 4),
        "raster_tile": list(vr["raster_config"][1]),
    }


def main_views(args):
    sc, vr = run_views(args.steps, args.warmup, args.tune)
    fb = vr["fwd_bwd"]
    out = {
        "metric": "Mtri/s + shaded Gfrag/s at 1080p; fwd+bwd views/s vs plain fwd, 1 MI355X",
        "value": round(fb["views_per_s"], 2),
        "unit": "views/s (fwd+bwd)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(fb["ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded icosphere colours and texture, targets from a N(0, 0.1) perturbation)",
        "config": {"workload": "differentiable views", "views": sc.views, "size": sc.size,
                   "triangles": sc.n_triangles, "tex_size": sc.tex_size},
        "differentiable": views_summary(vr),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
