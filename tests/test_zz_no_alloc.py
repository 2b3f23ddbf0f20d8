"""sl_draw_static, the draw entry point that allocates nothing: its frames
equal sl_draw's (followed by sl_finish) bit for bit -- colour, depth, the
per-pixel last primitive and depth-pass count, and the frame counters -- on
two calls in a row, the second with a smaller input (and a 30-draw frame
drawn one call at a time), and a frame whose
buffers exceed the fixed scratch array is refused with SL_ERR_CAPACITY
before any pixel is touched."""
import numpy as np
import pytest

from softlight_amd import api as A
from softlight_amd import scenes

pytestmark = pytest.mark.gpu

STATS = ("triangles_submitted", "triangles_binned", "tile_pairs", "fragments_tested", "fragments_shaded", "frames")


def _frame(lib, build, static, **kw):
    """One frame: clear, then every draw of the scene through sl_draw
    (executed at the read) or sl_draw_static (each executed at once)."""
    ctx = A.SLContext(lib)
    s = build(ctx, **kw)
    ctx.reset_stats()
    ctx.clear_framebuffer(s.fbo, 0, s.clear_rgba, s.clear_depth)
    for kind, mesh, n in s.draws:
        shader = n if kind == "draw_shader" else s.shader
        assert kind in ("draw", "draw_shader")
        if static:
            ctx.draw_static(mesh, shader, s.fbo)
        else:
            ctx.draw(mesh, shader, s.fbo)
    out = s.read()
    st = ctx.stats()
    out["stats"] = {k: st[k] for k in STATS}
    out["draws"] = len(s.draws)
    ctx.close()
    return out


def _same(a, b):
    for k in ("color", "depth", "last", "count"):
        assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), k
    for k in STATS:
        if k != "frames":
            assert a["stats"][k] == b["stats"][k], (k, a["stats"], b["stats"])


def test_static_draw_matches_draw_on_two_calls(gpu_lib):
    """A larger frame, then a smaller one (the array keeps the earlier call's
    contents), then a frame of 30 draws with mixed opaque and alpha-blended
    materials drawn one sl_draw_static at a time (one frame per call) against
    the same draws batched by sl_draw."""
    cases = [(scenes.build_random_triangles, dict(seed=5, ntri=2000, W=320, H=240, debug=True)),
             (scenes.build_random_triangles, dict(seed=6, ntri=300, W=160, H=96, debug=True)),
             (scenes.build_batched_scene, dict(seed=5, n_meshes=30, median_tris=300, W=320, H=180, tex_size=64,
                                               debug=True))]
    for build, kw in cases:
        ref = _frame(gpu_lib, build, False, **kw)
        got = _frame(gpu_lib, build, True, **kw)
        assert ref["stats"]["fragments_shaded"] > 0
        assert ref["stats"]["frames"] == 1 and got["stats"]["frames"] == ref["draws"]
        _same(ref, got)


def test_static_draw_refuses_frames_beyond_the_array(gpu_lib):
    """600,000 small triangles need about 600k setup records (144 B + 48 B of
    varyings each): more than the 64 MiB array.  The call returns
    SL_ERR_CAPACITY and leaves the targets as they were."""
    from tests.test_capacity_gpu import small_triangles
    ctx = A.SLContext(gpu_lib)
    s = small_triangles(ctx, n=600_000)
    ctx.clear_framebuffer(s.fbo, 0, (0.25, 0.5, 0.75, 1.0), 0.5)
    ctx.finish()
    before = s.read()
    kind, mesh, _ = s.draws[0]
    with pytest.raises(A.SLError, match=r"\(-105\)"):
        ctx.draw_static(mesh, s.shader, s.fbo)
    after = s.read()
    for k in ("color", "depth", "last", "count"):
        assert np.array_equal(before[k], after[k]), k
    # the same context still draws through sl_draw, and sl_draw_static again
    # afterwards on a frame that fits
    ctx.draw(mesh, s.shader, s.fbo)
    ctx.finish()
    small = A.sl_mesh(mesh.vaoId, 0, 3000, mesh.mode, 0)
    ctx.draw_static(small, s.shader, s.fbo)
    ctx.close()
