"""GPU parity: the HIP draw path against the CPU oracle on the same seeded
inputs (coverage, depth-test outcome, per-pixel primitive order and FP32 depth
bit-exact; RGBA8 within 1 LSB; tested/shaded/binned counts exact)."""
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


@pytest.mark.parametrize("config", range(8))
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
