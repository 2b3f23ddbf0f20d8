"""Heavy-tile split of the raster (sl_set_raster_split): a tile's record list
cut into parts rasterised by separate workgroups and merged must give the
frame the unsplit raster and the oracle give -- depth, colour, per-pixel
last primitive and depth-pass count, counters -- for every depth-function
class that splits (LESS / LESS_EQUAL and GREATER / GREATER_EQUAL with depth
writes, any function without them), with ties, for points, lines and
wireframe triangles, with the visibility attachment, and on the full-size
benchmark frames.  Batches that do not split (blending, mixed functions)
are unchanged."""
import numpy as np
import pytest

from softlight_amd import api as A
from softlight_amd import scenes
from tests.geometry import build_screen_triangles, render
from tests.parity import assert_parity, render_both

pytestmark = pytest.mark.gpu


def _with_split(build, part_records):
    def b(ctx, **kw):
        ctx.set_raster_split(part_records)
        return build(ctx, **kw)
    return b


def _same(a, b):
    for k in ("color", "depth", "last", "count"):
        if k in a:
            assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), k
    for k in ("triangles_binned", "fragments_tested", "fragments_shaded"):
        assert a["stats"][k] == b["stats"][k], k


def _overlapping(ctx, n=400, W=96, H=64, seed=3, z="random", **kw):
    """n random triangles over a small target, so every tile holds far more
    records than a part."""
    rng = np.random.default_rng(seed)
    tris = [[tuple(int(v) for v in rng.integers(0, [W + 1, H + 1])) for _ in range(3)] for _ in range(n)]
    if z == "random":
        zs = list(rng.uniform(-0.9, 0.9, size=n).astype(np.float32))
    else:
        zs = [[float(z)] * 3] * n                     # every fragment at one depth: ties everywhere
    cols = rng.integers(0, 256, size=(3 * n, 4))
    return build_screen_triangles(ctx, tris, W, H, z=zs, colors=cols, **kw)


@pytest.mark.parametrize("dtest,dmask,z", [
    (A.SL_DEPTH_TEST_LESS_THAN, A.SL_DEPTH_MASK_ON, "random"),
    (A.SL_DEPTH_TEST_LESS_EQUAL, A.SL_DEPTH_MASK_ON, "random"),
    (A.SL_DEPTH_TEST_LESS_THAN, A.SL_DEPTH_MASK_ON, 0.25),
    (A.SL_DEPTH_TEST_LESS_EQUAL, A.SL_DEPTH_MASK_ON, 0.25),
    (A.SL_DEPTH_TEST_GREATER_THAN, A.SL_DEPTH_MASK_ON, "random"),
    (A.SL_DEPTH_TEST_GREATER_EQUAL, A.SL_DEPTH_MASK_ON, 0.25),
    (A.SL_DEPTH_TEST_LESS_THAN, A.SL_DEPTH_MASK_OFF, "random"),
    (A.SL_DEPTH_TEST_OFF, A.SL_DEPTH_MASK_OFF, "random"),
    (A.SL_DEPTH_TEST_NOT_EQUAL, A.SL_DEPTH_MASK_OFF, 0.25),
], ids=["less", "lequal", "less_ties", "lequal_ties", "greater", "gequal_ties", "less_nomask", "off_nomask",
        "notequal_nomask"])
@pytest.mark.parametrize("part", [1, 5])
def test_split_matches_unsplit_and_oracle(gpu_lib, oracle_lib, dtest, dmask, z, part):
    def build(ctx, **kw):
        s = _overlapping(ctx, z=z, depth_test=dtest, depth_mask=dmask)
        if dtest in (A.SL_DEPTH_TEST_GREATER_THAN, A.SL_DEPTH_TEST_GREATER_EQUAL):
            s.clear_depth = -1.0
        return s
    g_split, o = render_both(gpu_lib, oracle_lib, _with_split(build, part))
    g_whole, _ = render_both(gpu_lib, oracle_lib, _with_split(build, 0))
    assert_parity(g_split, o)
    _same(g_split, g_whole)
    assert g_split["stats"]["fragments_tested"] > 50_000


@pytest.mark.parametrize("part", [1, 3, 64])
@pytest.mark.parametrize("kind", ["points", "lines", "wire"])
def test_split_render_modes(gpu_lib, oracle_lib, kind, part):
    if kind == "points":
        build = lambda ctx, **kw: scenes.build_random_points(ctx, seed=2, npoints=30000, W=160, H=120, debug=True)
    elif kind == "lines":
        build = lambda ctx, **kw: scenes.build_random_lines(ctx, seed=2, nlines=3000, W=160, H=120, debug=True)
    else:
        build = lambda ctx, **kw: scenes.build_random_triangles(ctx, seed=2, ntri=600, W=160, H=120, debug=True, wire=True)
    g, o = render_both(gpu_lib, oracle_lib, _with_split(build, part))
    assert_parity(g, o)


@pytest.mark.parametrize("config", [0, 2, 5, 7, 9, 10, 11])
def test_split_on_every_tile_shape(gpu_lib, oracle_lib, config):
    def build(ctx, **kw):
        ctx.set_raster_config(config)
        ctx.set_raster_split(4)
        return scenes.build_sphere_scene(ctx, seed=2, n_instances=6, stacks=24, sectors=24, W=333, H=199, tex_size=64,
                                         debug=True)
    g, o = render_both(gpu_lib, oracle_lib, build)
    assert_parity(g, o)


def test_split_with_visibility_and_gradients(gpu_lib, oracle_lib):
    """The differentiable views split: visibility bit-exact, gradients as the oracle's."""
    outs = []
    for lib, part in ((oracle_lib, 0), (gpu_lib, 2)):
        ctx = A.SLContext(lib)
        ctx.set_raster_split(part)
        s = scenes.DiffScene(ctx, level=4, views=4, size=128, views_per_fbo=4, tex_size=64, edge_views=1)
        s.make_targets()
        s.forward(); s.loss_grad_images(); s.backward()
        outs.append((s.visibility(), s.gradients()))
    (vo, (gco, gto)), (vg, (gcg, gtg)) = outs
    for a, b in zip(vo, vg):
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    for ref, got in ((gco, gcg), (gto, gtg)):
        assert float(np.abs(ref - got).max()) <= 1e-4 * float(np.abs(ref).max())


def test_blended_batches_do_not_split(gpu_lib, oracle_lib):
    """The batched frame (alpha-blended materials) keeps one workgroup per
    tile whatever the split setting, and matches the oracle."""
    g, o = render_both(gpu_lib, oracle_lib, _with_split(scenes.build_batched_scene, 1), seed=5, n_meshes=120,
                       median_tris=400, W=480, H=270, tex_size=64, debug=True)
    assert_parity(g, o)


@pytest.mark.parametrize("name,build,kw", [
    ("random", scenes.build_random_triangles, dict(seed=0, ntri=4096, W=640, H=480)),
    ("spheres", scenes.build_sphere_scene, dict(seed=1)),
])
@pytest.mark.parametrize("part", [16, 64])
def test_split_full_size_frames(gpu_lib, oracle_lib, name, build, kw, part):
    g, o = render_both(gpu_lib, oracle_lib, _with_split(build, part), debug=True, **kw)
    assert_parity(g, o)


def test_split_graph_replay(gpu_lib, oracle_lib):
    """A captured frame with split tiles replays to the oracle's image."""
    ctx = A.SLContext(gpu_lib)
    ctx.set_raster_split(8)
    s = scenes.build_sphere_scene(ctx, seed=4, n_instances=10, stacks=40, sectors=40, W=400, H=240, tex_size=64, debug=True)
    s.render()
    ctx.frame_graph_capture()
    ctx.finish()
    for _ in range(3):
        ctx.frame_graph_launch()
    ctx.finish()
    g = s.read()
    co = A.SLContext(oracle_lib)
    so = scenes.build_sphere_scene(co, seed=4, n_instances=10, stacks=40, sectors=40, W=400, H=240, tex_size=64, debug=True)
    so.render()
    o = so.read()
    assert np.array_equal(g["depth"].view(np.uint32), o["depth"].view(np.uint32))
    assert np.array_equal(g["last"], o["last"])
    assert np.array_equal(g["count"], o["count"])
    assert int(np.abs(g["color"].astype(int) - o["color"].astype(int)).max()) <= 1
