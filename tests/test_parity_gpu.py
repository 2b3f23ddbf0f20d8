"""GPU parity: the HIP draw path against the CPU oracle on the same seeded
inputs (coverage, depth-test outcome, per-pixel primitive order and FP32 depth
bit-exact; RGBA8 within 1 LSB; tested/shaded/binned counts exact)."""
import numpy as np
import pytest

from softlight_amd import api as A
from softlight_amd import scenes
from tests.parity import assert_parity, render_both

pytestmark = pytest.mark.gpu


def test_random_triangles_small(gpu_lib, oracle_lib):
    g, o = render_both(gpu_lib, oracle_lib, scenes.build_random_triangles, seed=3, ntri=64, W=96, H=80, debug=True)
    assert_parity(g, o)


def test_random_triangles_640x480(gpu_lib, oracle_lib):
    g, o = render_both(gpu_lib, oracle_lib, scenes.build_random_triangles, seed=0, ntri=4096, W=640, H=480, debug=True)
    rep = assert_parity(g, o)
    assert rep["fragments_tested"][0] > 90_000_000


@pytest.mark.parametrize("scissor,viewport", [((100, 50, 300, 250), None), ((-20, 30, 400, 500), (40, 20, 500, 400)),
                                               (None, (64, 48, 256, 192))])
def test_random_triangles_scissor_and_viewport(gpu_lib, oracle_lib, scissor, viewport):
    """Non-default scissor rectangles and viewports (SL_ViewportState.cpp:
    17-111): the scissor matrix, the clip against it and the viewport mapping
    bit-exact against the oracle."""
    def build(ctx, **kw):
        if scissor is not None:
            ctx.set_scissor(*scissor)
        if viewport is not None:
            ctx.set_viewport(*viewport)
        return scenes.build_random_triangles(ctx, **kw)
    g, o = render_both(gpu_lib, oracle_lib, build, seed=12, ntri=1500, W=640, H=480, debug=True)
    rep = assert_parity(g, o)
    assert rep["fragments_tested"][0] > 0


def test_spheres_reduced(gpu_lib, oracle_lib):
    g, o = render_both(gpu_lib, oracle_lib, scenes.build_sphere_scene, seed=1, n_instances=8, stacks=48, sectors=48,
                       W=480, H=270, tex_size=128, debug=True)
    assert_parity(g, o)


def test_spheres_1080p(gpu_lib, oracle_lib):
    g, o = render_both(gpu_lib, oracle_lib, scenes.build_sphere_scene, seed=1, debug=True)
    rep = assert_parity(g, o)
    assert rep["triangles_submitted"][0] == 64 * 130560


@pytest.mark.parametrize("config", range(12))
def test_every_raster_configuration(gpu_lib, oracle_lib, config):
    """Results do not depend on the tile configuration."""
    from softlight_amd import api

    outs = []
    for lib in (gpu_lib, oracle_lib):
        ctx = api.SLContext(lib)
        ctx.set_raster_config(config)
        s = scenes.build_random_triangles(ctx, seed=7, ntri=600, W=203, H=157, debug=True)
        ctx.reset_stats()
        s.render()
        o = s.read()
        o["stats"] = ctx.stats()
        outs.append(o)
    assert_parity(*outs)
    g, o = render_both(gpu_lib, oracle_lib, _spheres_cfg(config), seed=2, n_instances=6, stacks=24, sectors=24,
                       W=333, H=199, tex_size=64, debug=True)
    assert_parity(g, o)


@pytest.mark.parametrize("wg", [1, 2, 8])
def test_occupancy_cap_does_not_change_results(gpu_lib, oracle_lib, wg):
    """Capping the raster's resident workgroups per CU (extra LDS per
    workgroup) changes scheduling only."""
    def build(ctx, **kw):
        ctx.set_raster_occupancy(wg)
        return scenes.build_sphere_scene(ctx, **kw)
    g, o = render_both(gpu_lib, oracle_lib, build, seed=2, n_instances=6, stacks=24, sectors=24, W=333, H=199,
                       tex_size=64, debug=True)
    assert_parity(g, o)


def test_large_tile_falls_back_for_float_targets(gpu_lib, oracle_lib):
    """Configuration 11 (64x64) with an RGBA32F colour target and the
    visibility attachment runs on the 32x32 tile: visibility bit-exact and
    colour as the oracle's."""
    outs = []
    for lib in (gpu_lib, oracle_lib):
        ctx = A.SLContext(lib)
        ctx.set_raster_config(11)
        s = scenes.DiffScene(ctx, level=3, views=4, size=96, views_per_fbo=4, tex_size=64, edge_views=1)
        s.forward()
        ctx.finish()
        t = s.targets[0]
        outs.append((ctx.texture_read(t["vis"]), ctx.texture_read(t["color"])))
    (vg, cg), (vo, co) = outs
    np.testing.assert_array_equal(vg.view(np.uint32), vo.view(np.uint32))
    np.testing.assert_allclose(cg, co, atol=1e-6, rtol=0)


def _spheres_cfg(config):
    def build(ctx, **kw):
        ctx.set_raster_config(config)
        return scenes.build_sphere_scene(ctx, **kw)
    return build


def test_batched_meshes_reduced(gpu_lib, oracle_lib):
    """120 meshes, mixed opaque / alpha-blended materials, 16 textures, one batch."""
    g, o = render_both(gpu_lib, oracle_lib, scenes.build_batched_scene, seed=5, n_meshes=120, median_tris=400,
                       W=480, H=270, tex_size=64, debug=True)
    rep = assert_parity(g, o)
    assert g["scene"].info["blended"] > 0


def test_batched_meshes_1080p(gpu_lib, oracle_lib):
    """The full batched frame: 1000 meshes (log-normal triangle counts) at 1080p."""
    g, o = render_both(gpu_lib, oracle_lib, scenes.build_batched_scene, debug=True)
    assert_parity(g, o)


# ---------------------------------------------------------------------------
# The raster kernels the benchmark times: no debug attachments (so the
# non-debug k_raster instances run) and the tuned tile of each workload.
# ---------------------------------------------------------------------------
TUNED = {
    "random": ("random:640x480:4096", scenes.build_random_triangles, dict(seed=0, ntri=4096, W=640, H=480)),
    "spheres": ("spheres:1920x1080:8355840", scenes.build_sphere_scene, dict(seed=1)),
    "batched": ("batched:1920x1080:3399137", scenes.build_batched_scene, dict(seed=3, n_meshes=1000)),
}


def _tuned_builder(name):
    from softlight_amd import autotune
    signature, build, kw = TUNED[name]
    entry = autotune.cached_entry(signature)
    assert entry is not None, f"no tuned raster configuration cached for {signature}"

    def builder(ctx, **extra):
        autotune.apply(ctx, entry)           # tile, occupancy cap and bin capacity, as the benchmark runs
        return build(ctx, **kw, **extra)
    return builder, entry["config"]


@pytest.mark.parametrize("name", ["random", "spheres", "batched"])
def test_tuned_configuration_without_debug_buffers(gpu_lib, oracle_lib, name):
    """The benchmark's frames (C1 4096 random triangles at 640x480, C2 the
    1080p sphere frame, C3 the 1000-mesh batched frame) on their tuned raster
    tiles and without parity attachments: depth bit-exact, RGBA8 within 1 LSB,
    counters equal."""
    builder, cfg = _tuned_builder(name)
    g, o = render_both(gpu_lib, oracle_lib, builder, debug=False)
    assert "last" not in g
    assert g["scene"].ctx.raster_config()[0] == cfg
    rep = assert_parity(g, o)
    assert rep["fragments_shaded"][0] > 0


def test_batched_graph_replays_match_oracle(gpu_lib, oracle_lib):
    """C3 as the benchmark runs it: the 1000-mesh frame captured as a HIP
    graph on its tuned tile and replayed three times; before each replay the
    targets are overwritten with a junk pattern, and every replay's image
    equals the oracle's frame (depth bit-exact, RGBA8 within 1 LSB, counters
    equal per frame)."""
    from softlight_amd import api

    builder, _ = _tuned_builder("batched")
    ctx_o = api.SLContext(oracle_lib)
    so = builder(ctx_o, debug=False)
    ctx_o.reset_stats()
    so.render()
    o = so.read()
    o["stats"] = ctx_o.stats()

    ctx = api.SLContext(gpu_lib)
    s = builder(ctx, debug=False)
    s.render()
    ctx.frame_graph_capture()
    ctx.finish()
    junk_c = np.full((s.H, s.W, 4), 0x5A, dtype=np.uint8)
    junk_d = np.full((s.H, s.W), -0.25, dtype=np.float32)
    for _ in range(3):
        ctx.texture_set_texels(s.color, junk_c)
        ctx.texture_set_texels(s.depth, junk_d)
        ctx.reset_stats()
        ctx.frame_graph_launch()
        ctx.finish()
        g = s.read()
        g["stats"] = ctx.stats()
        assert g["stats"]["frames"] == 1
        assert_parity(g, o)


def test_frame_graph_replays_match(gpu_lib, oracle_lib):
    """A captured frame replayed three times gives the eager frame's and the
    oracle's result, and the counters count every replay."""
    from softlight_amd import api

    ctx = api.SLContext(gpu_lib)
    s = scenes.build_sphere_scene(ctx, seed=4, n_instances=10, stacks=40, sectors=40, W=400, H=240, tex_size=64,
                                  debug=True)
    s.render()
    ctx.frame_graph_capture()
    eager = s.read()
    ctx.reset_stats()
    for _ in range(3):
        ctx.frame_graph_launch()
    ctx.finish()
    st = ctx.stats()
    replay = s.read()
    assert st["frames"] == 3 and st["triangles_submitted"] == 3 * s.n_triangles
    for k in ("color", "depth", "last", "count"):
        assert np.array_equal(eager[k], replay[k]), k
    ctx_o = api.SLContext(oracle_lib)
    so = scenes.build_sphere_scene(ctx_o, seed=4, n_instances=10, stacks=40, sectors=40, W=400, H=240, tex_size=64,
                                   debug=True)
    so.render()
    o = so.read()
    assert np.array_equal(replay["depth"].view(np.uint32), o["depth"].view(np.uint32))
    assert np.array_equal(replay["last"], o["last"])
