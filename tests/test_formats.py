"""Target formats beyond RGBA8 / RGBA32F colour and R32F depth: R16F depth
(the demos' depth format) and 8-bit / float colour of one to three
components (SL_Framebuffer.cpp:42-80, 268-370; SL_Context.cpp:934-962).

What pins the oracle, and the mistake each check catches:

  half depth        depth texels equal numpy's float32 -> float16 rounding
                    (round to nearest even) of the same draw's R32F depth;
                    the clear value is (half)(float)depth: truncation or a
                    wrong rounding mode
  half depth test   a fragment whose z lies between a float depth and its
                    rounded half passes against the half texel only: the test
                    reading the float z written instead of the stored half
  n-component u8    R8 / RG8 / RGB8 texels are the first n bytes of the RGBA8
                    result of the same draw (itself pinned against trunc(c*255));
                    float R / RG / RGB likewise against RGBA32F
  blend read        a blend reads a three-component texel as (r, r, r, max)
                    and a one/two-component texel with min for the missing
                    colour components and max alpha (assign_alpha_pixel):
                    additive blending of a black source onto a cleared RGB8
                    target turns g and b into r, on RG8 and R8 changes nothing
GPU (marked gpu): the same checks, and bit-exact parity with the oracle on the
sphere frame in the demos' RGB8 + R16F configuration, random triangles on
every new format, and the blended batched scene on RGB8.
"""
from __future__ import annotations

import numpy as np
import pytest

from softlight_amd import api as A
from softlight_amd import scenes
from tests.geometry import render, vertex_records
from tests.parity import assert_parity, render_both

BACKENDS = [pytest.param("oracle", id="oracle"), pytest.param("gpu", id="gpu", marks=pytest.mark.gpu)]
U8 = [(A.SL_COLOR_R_8U, 1), (A.SL_COLOR_RG_8U, 2), (A.SL_COLOR_RGB_8U, 3)]
FL = [(A.SL_COLOR_R_FLOAT, 1), (A.SL_COLOR_RG_FLOAT, 2), (A.SL_COLOR_RGB_FLOAT, 3)]


@pytest.fixture
def make_ctx(request):
    def _make(kind):
        return A.SLContext(request.getfixturevalue("oracle_lib" if kind == "oracle" else "gpu_lib"))
    return _make


def tri_scene(ctx, W, H, tris, z=None, colors=None, color_type=A.SL_COLOR_RGBA_8U, depth_type=A.SL_COLOR_R_FLOAT,
              depth_test=A.SL_DEPTH_TEST_OFF, blend=A.SL_BLEND_OFF):
    rec = vertex_records(tris, W, H, z=z, colors=colors)
    return scenes.build_from_vertex_records(ctx, rec, W, H, debug=True, color_type=color_type, depth_type=depth_type,
                                            shader_kw=dict(cull=A.SL_CULL_OFF, depth_test=depth_test,
                                                           depth_mask=A.SL_DEPTH_MASK_ON, blend=blend))


TRIS = [[(3, 2), (60, 17), (21, 45)], [(50, 40), (5, 30), (62, 3)]]
ZS = [[-0.73, 0.41, 0.127], [0.33, -0.58, 0.91]]


@pytest.mark.parametrize("kind", BACKENDS)
def test_half_depth_is_the_rounded_float_depth(kind, make_ctx):
    ctx = make_ctx(kind)
    W, H = 64, 48
    ref = render(ctx, tri_scene(ctx, W, H, TRIS, z=ZS))
    s = tri_scene(ctx, W, H, TRIS, z=ZS, depth_type=A.SL_COLOR_R_HALF)
    s.clear_depth = 0.3
    half = render(ctx, s)
    assert half["depth"].dtype == np.float16
    covered = ref["count"] > 0
    assert covered.sum() > 500
    np.testing.assert_array_equal(half["depth"][covered].view(np.uint16),
                                  ref["depth"][covered].astype(np.float16).view(np.uint16))
    np.testing.assert_array_equal(half["depth"][~covered], np.float16(np.float32(0.3)))
    np.testing.assert_array_equal(half["count"], ref["count"])


@pytest.mark.parametrize("kind", BACKENDS)
def test_half_depth_test_reads_the_stored_half(kind, make_ctx):
    ctx = make_ctx(kind)
    W, H = 32, 32
    # a rounds up in half; b lies between a and half(a)
    a = np.float32(0.30004)
    ha = np.float32(np.float16(a))
    assert ha > a
    b = np.float32((float(a) + float(ha)) / 2)
    assert a < b < ha
    tri = [(1, 1), (30, 2), (4, 29)]
    cols = [(255, 0, 0, 255)] * 3 + [(0, 255, 0, 255)] * 3
    out = {}
    for dt in (A.SL_COLOR_R_FLOAT, A.SL_COLOR_R_HALF):
        s = tri_scene(ctx, W, H, [tri, tri], z=[[float(a)] * 3, [float(b)] * 3], colors=cols, depth_type=dt,
                      depth_test=A.SL_DEPTH_TEST_LESS_THAN)
        out[dt] = render(ctx, s)
    inside = out[A.SL_COLOR_R_FLOAT]["count"] > 0
    # float target: b > a fails; half target: b < half(a) passes
    f, h = out[A.SL_COLOR_R_FLOAT]["color"][inside], out[A.SL_COLOR_R_HALF]["color"][inside]
    assert np.all(f[:, 0] >= 250) and np.all(f[:, 1] <= 5)
    assert np.all(h[:, 1] >= 250) and np.all(h[:, 0] <= 5)
    assert np.all(out[A.SL_COLOR_R_HALF]["count"][inside] == 2)


@pytest.mark.parametrize("kind", BACKENDS)
@pytest.mark.parametrize("fmt,n", U8 + FL)
def test_n_component_texels(kind, make_ctx, fmt, n):
    ctx = make_ctx(kind)
    W, H = 64, 48
    rng = np.random.default_rng(4)
    cols = rng.integers(0, 256, (6, 4)).astype(np.uint8)
    full = A.SL_COLOR_RGBA_8U if (fmt, n) in U8 else A.SL_COLOR_RGBA_FLOAT
    out = []
    for t in (full, fmt):
        s = tri_scene(ctx, W, H, TRIS, z=ZS, colors=cols, color_type=t, depth_test=A.SL_DEPTH_TEST_LESS_THAN)
        s.clear_rgba = (0.25, 0.5, 0.75, 1.0)
        out.append(render(ctx, s))
    ref, got = out
    # drawn and cleared pixels alike: the RGBA result's first n components
    assert (ref["count"] > 0).sum() > 500
    np.testing.assert_array_equal(got["color"].reshape(H, W, n), ref["color"].reshape(H, W, 4)[..., :n])


@pytest.mark.parametrize("kind", BACKENDS)
@pytest.mark.parametrize("fmt,n", U8 + FL)
def test_blend_reads_missing_components(kind, make_ctx, fmt, n):
    ctx = make_ctx(kind)
    W, H = 32, 32
    tri = [(1, 1), (30, 2), (4, 29)]
    s = tri_scene(ctx, W, H, [tri], colors=[(0, 0, 0, 255)] * 3, color_type=fmt, blend=A.SL_BLEND_ADDITIVE)
    s.clear_rgba = (0.2, 0.6, 0.8, 1.0)
    out = render(ctx, s)
    c = out["color"].reshape(H, W, n)
    inside = out["count"] > 0
    outside = ~inside
    assert inside.sum() > 200
    if fmt in (A.SL_COLOR_R_8U, A.SL_COLOR_RG_8U, A.SL_COLOR_RGB_8U):
        cleared = np.array([int(0.2 * 255), int(0.6 * 255), int(0.8 * 255)])[:n]
        # the read value r = 51 * (1/255) in float32, written back as trunc(r * 255)
        r = np.float32(cleared[0]) * np.float32(0.00392156862745)
        back = int(np.float32(r) * np.float32(255.0))
        expect = np.array([back] * 3)[:n] if n == 3 else np.array([back, cleared[1] if n > 1 else 0])[:n]
        if n == 2:
            g = np.float32(cleared[1]) * np.float32(0.00392156862745)
            expect = np.array([back, int(np.float32(g) * np.float32(255.0))])
    else:
        cleared = np.array([0.2, 0.6, 0.8], dtype=np.float32)[:n]
        expect = np.array([cleared[0]] * 3, dtype=np.float32)[:n] if n == 3 else cleared
    np.testing.assert_array_equal(c[outside], np.broadcast_to(cleared.astype(c.dtype), c[outside].shape))
    np.testing.assert_array_equal(c[inside], np.broadcast_to(np.asarray(expect).astype(c.dtype), c[inside].shape))


# ---------------------------------------------------------------------------
@pytest.mark.gpu
def test_gpu_spheres_rgb8_half_depth(gpu_lib, oracle_lib):
    g, o = render_both(gpu_lib, oracle_lib, scenes.build_sphere_scene, seed=1, n_instances=8, stacks=48, sectors=48,
                       W=640, H=360, tex_size=128, debug=True, color_type=A.SL_COLOR_RGB_8U,
                       depth_type=A.SL_COLOR_R_HALF)
    assert_parity(g, o)


@pytest.mark.gpu
@pytest.mark.parametrize("fmt", [A.SL_COLOR_R_8U, A.SL_COLOR_RG_8U, A.SL_COLOR_RGB_8U, A.SL_COLOR_R_FLOAT,
                                 A.SL_COLOR_RG_FLOAT, A.SL_COLOR_RGB_FLOAT])
@pytest.mark.parametrize("blend", [A.SL_BLEND_OFF, A.SL_BLEND_ALPHA])
def test_gpu_random_triangles_formats(gpu_lib, oracle_lib, fmt, blend):
    g, o = render_both(gpu_lib, oracle_lib, scenes.build_random_triangles, seed=9, ntri=800, W=320, H=240,
                       debug=True, color_type=fmt, depth_type=A.SL_COLOR_R_HALF, blend=blend)
    assert_parity(g, o, float_tol=0.0)


@pytest.mark.gpu
def test_gpu_batched_rgb8(gpu_lib, oracle_lib):
    g, o = render_both(gpu_lib, oracle_lib, scenes.build_batched_scene, seed=5, n_meshes=120, median_tris=400,
                       W=640, H=360, n_textures=4, tex_size=64, debug=True, color_type=A.SL_COLOR_RGB_8U,
                       depth_type=A.SL_COLOR_R_HALF)
    assert_parity(g, o)


# ---------------------------------------------------------------------------
# packed formats (SL_ColorCompressed.hpp): fields r, g, b, a from the lowest bit
# ---------------------------------------------------------------------------
PACKED = {A.SL_COLOR_RGB_332: (3, 3, 2, 0), A.SL_COLOR_RGB_565: (5, 6, 5, 0), A.SL_COLOR_RGBA_5551: (5, 5, 5, 1),
          A.SL_COLOR_RGBA_4444: (4, 4, 4, 4), A.SL_COLOR_RGBA_1010102: (10, 10, 10, 2)}


def fields(word, bits):
    out, shift = [], 0
    for b in bits:
        out.append((int(word) >> shift) & ((1 << b) - 1) if b else None)
        shift += b
    return out


@pytest.mark.parametrize("kind", BACKENDS)
@pytest.mark.parametrize("fmt", list(PACKED))
def test_packed_texels(kind, make_ctx, fmt):
    """Drawn texels: each field = trunc(c * max) of the float colour the same
    draw writes to an RGBA32F target; cleared texels: trunc(clear * max) in
    double."""
    ctx = make_ctx(kind)
    W, H = 64, 48
    bits = PACKED[fmt]
    rng = np.random.default_rng(6)
    cols = rng.integers(0, 256, (6, 4)).astype(np.uint8)
    out = []
    for t in (A.SL_COLOR_RGBA_FLOAT, fmt):
        s = tri_scene(ctx, W, H, TRIS, z=ZS, colors=cols, color_type=t, depth_test=A.SL_DEPTH_TEST_LESS_THAN)
        s.clear_rgba = (0.3, 0.55, 0.8, 0.6)
        out.append(render(ctx, s))
    ref, got = out
    covered = ref["count"] > 0
    assert covered.sum() > 500
    words = got["color"].reshape(H, W)
    fc = ref["color"].reshape(H, W, 4)
    ys, xs = np.nonzero(covered)
    for y, x in list(zip(ys, xs))[::7]:
        f = fields(words[y, x], bits)
        for k, b in enumerate(bits):
            if b:
                assert f[k] == int(np.float32(fc[y, x, k]) * np.float32((1 << b) - 1)) % (1 << b), (x, y, k)
    cy, cx = np.nonzero(~covered)
    f = fields(words[cy[0], cx[0]], bits)
    for k, (b, c) in enumerate(zip(bits, (0.3, 0.55, 0.8, 0.6))):
        if b:
            assert f[k] == int(c * ((1 << b) - 1)), (k, f)


@pytest.mark.parametrize("kind", BACKENDS)
@pytest.mark.parametrize("fmt", [A.SL_COLOR_RGB_565, A.SL_COLOR_RGBA_4444])
def test_packed_blend_decodes_the_texel(kind, make_ctx, fmt):
    """Additive blend of an opaque black source: the texel decodes to
    field / max, blends, clamps and encodes back: the cleared colour fields
    survive unchanged, alpha saturates."""
    ctx = make_ctx(kind)
    W, H = 32, 32
    tri = [(1, 1), (30, 2), (4, 29)]
    s = tri_scene(ctx, W, H, [tri], colors=[(0, 0, 0, 255)] * 3, color_type=fmt, blend=A.SL_BLEND_ADDITIVE)
    s.clear_rgba = (0.3, 0.55, 0.8, 0.6)
    o = render(ctx, s)
    words = o["color"].reshape(H, W)
    inside = o["count"] > 0
    assert inside.sum() > 200
    bits = PACKED[fmt]
    want = []
    for k, (b, c) in enumerate(zip(bits, (0.3, 0.55, 0.8, 0.6))):
        if b and k == 3:
            want.append((1 << b) - 1)          # alpha: 0.6 + 1 * 1 clamps to 1
        elif b:
            m = (1 << b) - 1
            v = int(c * m)
            # decode v / m in float32, add 0, encode: trunc((v / m) * m)
            want.append(int(np.float32(np.float32(v) / np.float32(m)) * np.float32(m)))
        else:
            want.append(None)
    for y, x in zip(*np.nonzero(inside)):
        assert [v for v in fields(words[y, x], bits)] == want


@pytest.mark.gpu
@pytest.mark.parametrize("fmt", list(PACKED))
def test_gpu_packed_parity(gpu_lib, oracle_lib, fmt):
    g, o = render_both(gpu_lib, oracle_lib, scenes.build_random_triangles, seed=13, ntri=600, W=320, H=240,
                       debug=True, color_type=fmt, blend=A.SL_BLEND_ALPHA)
    np.testing.assert_array_equal(g["color"], o["color"])
    np.testing.assert_array_equal(g["count"], o["count"])
