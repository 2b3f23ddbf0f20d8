"""A frame that outgrows the bin buffers: the counting pass reports the
overflow, the host grows the record, bin-entry and clip-job buffers to the
requirement (+25%) and runs the setup again before anything is drawn.  The
result must be the oracle's, on the first frame of a fresh context (which
overflows the initial capacities) and on the next one (which fits)."""
import numpy as np
import pytest

from softlight_amd import api as A
from softlight_amd import scenes
from tests.geometry import ndc_for_pixel
from tests.parity import assert_parity

pytestmark = pytest.mark.gpu


def small_triangles(ctx, n=300_000, W=256, H=256, seed=3):
    """n right triangles of 3 pixels a side at random positions: more records
    than the initial record capacity (2^18), depth test off, random colours."""
    rng = np.random.default_rng(seed)
    x = rng.integers(0, W - 3, size=n)
    y = rng.integers(0, H - 3, size=n)
    X = np.stack([x, x + 3, x], axis=1).reshape(-1)
    Y = np.stack([y, y, y + 3], axis=1).reshape(-1)
    ndcx = np.array([ndc_for_pixel(int(v), W) for v in range(W + 1)], dtype=np.float32)
    ndcy = np.array([ndc_for_pixel(int(v), H) for v in range(H + 1)], dtype=np.float32)
    rec = np.zeros((3 * n, 16), dtype=np.uint8)
    pos = np.stack([ndcx[X], ndcy[Y], np.zeros(3 * n, np.float32)], axis=1).astype(np.float32)
    rec[:, 0:12] = pos.view(np.uint8).reshape(-1, 12)
    rec[:, 12:16] = rng.integers(0, 256, size=(3 * n, 4), dtype=np.uint8)
    return scenes.build_from_vertex_records(ctx, rec, W, H, debug=True,
                                            shader_kw=dict(cull=A.SL_CULL_OFF, depth_test=A.SL_DEPTH_TEST_OFF))


def test_frame_outgrowing_the_bin_capacity(gpu_lib, oracle_lib):
    outs = []
    for lib in (gpu_lib, oracle_lib):
        ctx = A.SLContext(lib)
        s = small_triangles(ctx)
        frames = []
        for _ in range(2):
            ctx.reset_stats()
            s.render()
            out = s.read()
            out["stats"] = ctx.stats()
            frames.append(out)
        outs.append(frames)
        ctx.close()
    (g0, g1), (o0, o1) = outs
    assert g0["stats"]["triangles_binned"] > (1 << 18)
    assert_parity(g0, o0)
    assert_parity(g1, o1)
