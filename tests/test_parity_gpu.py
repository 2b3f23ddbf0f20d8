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


def test_spheres_reduced(gpu_lib, oracle_lib):
    g, o = render_both(gpu_lib, oracle_lib, scenes.build_sphere_scene, seed=1, n_instances=8, stacks=48, sectors=48,
                       W=480, H=270, tex_size=128, debug=True)
    assert_parity(g, o)


def test_spheres_1080p(gpu_lib, oracle_lib):
    g, o = render_both(gpu_lib, oracle_lib, scenes.build_sphere_scene, seed=1, debug=True)
    rep = assert_parity(g, o)
    assert rep["triangles_submitted"][0] == 64 * 130560


@pytest.mark.parametrize("config", range(10))
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
