"""Depth rejection (sl_set_raster_depth_reject): a wave of a tile with 64 or
more records skips the filled triangles whose depth provably fails a LESS /
LESS_EQUAL test at every pixel it owns, and only counts their depth tests.
With it on, coverage, depth, per-pixel primitive order and the
tested/shaded/binned counts equal the oracle's exactly and equal the
output with it off: on random triangles (every tile ~2,300 records), on
every tile configuration, with depth-raising draws (GREATER, test off with
writes) between LESS draws of one batch, with heavy-tile splitting, blended
batches and R16F depth (where the variant rejects nothing)."""
import numpy as np
import pytest

from softlight_amd import api as A
from softlight_amd import scenes
from tests.parity import assert_parity, render_both

pytestmark = pytest.mark.gpu


def _reject(build, on=True, config=None, split=None):
    def b(ctx, **kw):
        ctx.set_raster_depth_reject(on)
        if config is not None:
            ctx.set_raster_config(config)
        if split is not None:
            ctx.set_raster_split(split)
        return build(ctx, **kw)
    return b


def _same(a, b):
    for k in ("color", "depth", "last", "count"):
        assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), k
    for k in ("triangles_binned", "fragments_tested", "fragments_shaded"):
        assert a["stats"][k] == b["stats"][k], k


def test_random_triangles_640x480(gpu_lib, oracle_lib):
    g, o = render_both(gpu_lib, oracle_lib, _reject(scenes.build_random_triangles), seed=0, ntri=4096, W=640, H=480,
                       debug=True)
    rep = assert_parity(g, o)
    assert rep["fragments_tested"][0] > 90_000_000


def test_matches_reject_off(gpu_lib, oracle_lib):
    kw = dict(seed=4, ntri=2000, W=320, H=240, debug=True)
    g_on, o = render_both(gpu_lib, oracle_lib, _reject(scenes.build_random_triangles), **kw)
    g_off, _ = render_both(gpu_lib, oracle_lib, _reject(scenes.build_random_triangles, on=False), **kw)
    assert_parity(g_on, o)
    _same(g_on, g_off)


@pytest.mark.parametrize("config", range(12))
def test_every_raster_configuration(gpu_lib, oracle_lib, config):
    g, o = render_both(gpu_lib, oracle_lib, _reject(scenes.build_random_triangles, config=config), seed=7, ntri=600,
                       W=203, H=157, debug=True)
    assert_parity(g, o)


def test_spheres_and_batched_meshes(gpu_lib, oracle_lib):
    g, o = render_both(gpu_lib, oracle_lib, _reject(scenes.build_sphere_scene), seed=1, n_instances=8, stacks=48,
                       sectors=48, W=480, H=270, tex_size=128, debug=True)
    assert_parity(g, o)
    g, o = render_both(gpu_lib, oracle_lib, _reject(scenes.build_batched_scene), seed=5, n_meshes=120,
                       median_tris=400, W=480, H=270, tex_size=64, debug=True)
    assert_parity(g, o)
    assert g["scene"].info["blended"] > 0


def test_blended_random_triangles(gpu_lib, oracle_lib):
    """Blended LESS draws are rejected too: a fragment that fails its depth
    test neither blends nor writes."""
    g, o = render_both(gpu_lib, oracle_lib, _reject(scenes.build_random_triangles), seed=8, ntri=1500, W=200, H=150,
                       debug=True, blend=A.SL_BLEND_ALPHA)
    assert_parity(g, o)


def test_half_depth_target(gpu_lib, oracle_lib):
    g, o = render_both(gpu_lib, oracle_lib, _reject(scenes.build_random_triangles), seed=9, ntri=1500, W=200, H=150,
                       debug=True, depth_type=A.SL_COLOR_R_HALF)
    assert_parity(g, o)


def _mixed_depth_functions(ctx, seed=11, ntri=1800, W=160, H=120, debug=True):
    """One batch of draws over the same random triangles: LESS, GREATER with
    writes, LESS, test off with writes, LESS_EQUAL, NOT_EQUAL with writes,
    LESS -- the raising draws land between LESS records of one staged batch."""
    rec = scenes.random_triangle_vertices(seed, ntri)
    vbo = ctx.create_vbo(rec)
    vao = ctx.create_vao(vbo, None, [(0, 16, A.VERTEX_DIMENSION_3, A.VERTEX_DATA_FLOAT),
                                     (12, 16, A.VERTEX_DIMENSION_4, A.VERTEX_DATA_BYTE)])
    color, depth, fbo, dbg = scenes.make_target(ctx, W, H, debug)
    funcs = [A.SL_DEPTH_TEST_LESS_THAN, A.SL_DEPTH_TEST_GREATER_THAN, A.SL_DEPTH_TEST_LESS_THAN,
             A.SL_DEPTH_TEST_OFF, A.SL_DEPTH_TEST_LESS_EQUAL, A.SL_DEPTH_TEST_NOT_EQUAL, A.SL_DEPTH_TEST_LESS_THAN]
    sizes = [500, 37, 400, 23, 500, 40, 300]
    assert sum(sizes) == ntri
    draws, first = [], 0
    for f, n in zip(funcs, sizes):
        shader = scenes.gouraud_shader(ctx, depth_test=f, depth_mask=A.SL_DEPTH_MASK_ON)
        draws.append(("draw_shader", A.sl_mesh(vao, 3 * first, 3 * n, A.RENDER_MODE_TRIANGLES, 0), shader))
        first += n
    return scenes.Scene(ctx, W, H, fbo, color, depth, None, draws, n_triangles=ntri, debug=dbg)


def test_depth_raising_draws_in_one_batch(gpu_lib, oracle_lib):
    g_on, o = render_both(gpu_lib, oracle_lib, _reject(_mixed_depth_functions))
    g_off, _ = render_both(gpu_lib, oracle_lib, _reject(_mixed_depth_functions, on=False))
    assert_parity(g_on, o)
    _same(g_on, g_off)
    assert g_on["stats"]["fragments_tested"] > 100_000


@pytest.mark.parametrize("part", [3, 64])
def test_with_heavy_tile_split(gpu_lib, oracle_lib, part):
    g, o = render_both(gpu_lib, oracle_lib, _reject(scenes.build_random_triangles, split=part), seed=6, ntri=1500,
                       W=200, H=150, debug=True)
    assert_parity(g, o)
