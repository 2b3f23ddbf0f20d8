"""The vertex pass (k_vertex): an instanced indexed draw with at least 2^16
(instance, vertex) entries and at most two vertices per triangle has its clip
positions computed once per (instance, vertex) before setup; setup, emitting
and clipping read them.  Results must equal the oracle's (which runs the
position program per triangle corner) bit for bit -- also with indices past
the vertex buffer (those corners run the program themselves), with a vertex
buffer rewritten between replays of a captured frame, and on the full-size
sphere frame (tests/test_parity_gpu.py)."""
import numpy as np
import pytest

from softlight_amd import api as A
from softlight_amd import scenes
from tests.parity import assert_parity, render_both

pytestmark = pytest.mark.gpu

# 64 instances of a 40 x 40 UV sphere: 1681 x 64 = 107,584 entries (>= 2^16),
# 3,120 triangles per instance: the draw takes the vertex pass
KW = dict(seed=3, n_instances=64, stacks=40, sectors=40, W=480, H=270, tex_size=64, debug=True)


def test_qualifying_draw_matches_oracle(gpu_lib, oracle_lib):
    g, o = render_both(gpu_lib, oracle_lib, scenes.build_sphere_scene, **KW)
    rep = assert_parity(g, o)
    assert rep["triangles_submitted"][0] == 64 * 3120


def test_indices_past_the_vertex_buffer(gpu_lib, oracle_lib):
    """Every 97th index points past the vertex buffer: the pass covers ids
    below the buffer's vertex count, the rest read zeros in the program."""
    def build(ctx, **kw):
        s = scenes.build_sphere_scene(ctx, **kw)
        verts, idx, _ = s.info["_keep"]
        bad = idx.copy()
        bad[::97] = verts.shape[0] + (np.arange(bad[::97].size) % 7).astype(bad.dtype)
        ibo = ctx.create_ibo(bad)
        mesh = s.draws[0][1]
        ctx.vao_set_index_buffer(mesh.vaoId, ibo)
        s.info["_keep2"] = bad
        return s
    g, o = render_both(gpu_lib, oracle_lib, build, **KW)
    assert_parity(g, o)


def test_graph_replay_after_vertex_update(gpu_lib, oracle_lib):
    """A captured frame recomputes the positions on every replay: rewrite the
    vertex buffer (shrunk spheres), replay, compare with the oracle drawing
    the rewritten buffer."""
    def shrink(ctx, s):
        verts = s.info["_keep"][0].copy()
        verts[:, 0:3] *= 0.6
        ctx.vbo_assign(s.info["vbo"], verts)

    ctx = A.SLContext(gpu_lib)
    s = scenes.build_sphere_scene(ctx, **KW)
    s.render()
    ctx.frame_graph_capture()
    ctx.finish()
    shrink(ctx, s)
    for _ in range(2):
        ctx.frame_graph_launch()
    ctx.finish()
    g = s.read()
    co = A.SLContext(oracle_lib)
    so = scenes.build_sphere_scene(co, **KW)
    shrink(co, so)
    so.render()
    o = so.read()
    assert np.array_equal(g["depth"].view(np.uint32), o["depth"].view(np.uint32))
    assert np.array_equal(g["last"], o["last"])
    assert np.array_equal(g["count"], o["count"])
    assert int(np.abs(g["color"].astype(int) - o["color"].astype(int)).max()) <= 1
