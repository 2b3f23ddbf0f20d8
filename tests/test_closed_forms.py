"""Closed-form, exact-arithmetic and invariant checks of the draw path.

These pin the CPU oracle (oracle/sl_oracle.cpp), which has no reference
golden vectors to lean on (the reference's tests assert return codes only and
its LightSky math submodules are absent), and run the same checks on the HIP
library (marked gpu).  What each test pins and the mistake it is built to
catch:

  span coverage        exact rational span rule vs rendered coverage: wrong
                       edge pairing, top row included, round vs truncate
  right-triangle count n(n+1)/2 closed form; all 6 vertex orders and both
                       windings give the same set
  rectangle pair       two triangles sharing a diagonal cover a 64x32
                       rectangle exactly once (no gap, no double cover)
  sub-pixel reject     bbox narrower than 1 px in x or y draws nothing
  depth plane          depth at covered pixel corners vs exact barycentric
                       interpolation of z: wrong plane, sign or operand order
  perspective colour   varyings vs FP64 perspective-correct interpolation
                       with bc/w: using w instead of 1/w, missing divide
  depth functions      all 7 compare functions vs Python operators
  last writer          depth test off: per-pixel last primitive = max covering
                       index from the exact coverage
  order invariance     unique depths + LESS: result independent of order
  culling              CCW kept / CW culled under BACK, the reverse under FRONT
  blending             4 blend modes vs FP64 formulas of SL_Framebuffer.cpp:324-342
  bilinear sampler     RGBA32F texture sampling vs FP64 bilinear with the
                       reference's W/H neighbour clamp and REPEAT wrap
  clear                trunc(c*255) clear values
  threads              multithreaded oracle == single-threaded oracle
"""
from __future__ import annotations

import itertools
import math

import numpy as np
import pytest

from softlight_amd import api as A
from softlight_amd import scenes
from tests.geometry import (build_screen_triangles, exact_barycentric, exact_coverage, render,
                            vertex_records)

BACKENDS = [pytest.param("oracle", id="oracle"), pytest.param("gpu", id="gpu", marks=pytest.mark.gpu)]


@pytest.fixture
def make_ctx(request):
    def _make(kind):
        if kind == "oracle":
            import oracle
            return A.SLContext(oracle.load())
        import softlight_amd
        return A.SLContext(softlight_amd.load_library())
    return _make


def coverage_set(count):
    ys, xs = np.nonzero(count)
    return set(zip(xs.tolist(), ys.tolist()))


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("kind", BACKENDS)
def test_right_triangle_count_closed_form(kind, make_ctx):
    ctx = make_ctx(kind)
    n = 16
    base = [(4, 4), (4 + n, 4), (4, 4 + n)]
    sets = []
    for perm in itertools.permutations(base):
        s = build_screen_triangles(ctx, [list(perm)], 40, 40)
        out = render(ctx, s)
        assert int(out["count"].sum()) == n * (n + 1) // 2
        assert out["stats"]["fragments_tested"] == n * (n + 1) // 2
        sets.append(coverage_set(out["count"]))
    assert all(s == sets[0] for s in sets)
    # rows [4, 20): row y covers x in [4, 4 + n - (y - 4)); the top row y = 20 is never drawn
    expect = {(x, y) for y in range(4, 4 + n) for x in range(4, 4 + n - (y - 4))}
    assert sets[0] == expect


@pytest.mark.parametrize("kind", BACKENDS)
def test_rectangle_pair_covers_once(kind, make_ctx):
    ctx = make_ctx(kind)
    a, b = 64, 32
    tris = [[(0, 0), (a, 0), (a, b)], [(0, 0), (a, b), (0, b)]]
    s = build_screen_triangles(ctx, tris, 80, 48)
    out = render(ctx, s)
    cnt = out["count"]
    assert np.all(cnt[:b, :a] == 1)
    assert cnt.sum() == a * b


@pytest.mark.parametrize("kind", BACKENDS)
def test_subpixel_triangles_rejected(kind, make_ctx):
    ctx = make_ctx(kind)
    tris = [[(5, 5), (5, 9), (5, 12)],      # zero width
            [(2, 7), (9, 7), (20, 7)],      # zero height
            [(10, 10), (30, 10), (20, 11)]]  # height 1: exactly one row
    s = build_screen_triangles(ctx, tris, 40, 40)
    out = render(ctx, s)
    assert out["stats"]["triangles_binned"] == 1
    ys = np.nonzero(out["count"])[0]
    assert set(ys.tolist()) <= {10}


@pytest.mark.parametrize("kind", BACKENDS)
def test_random_small_triangles_match_exact_span_rule(kind, make_ctx):
    ctx = make_ctx(kind)
    rng = np.random.default_rng(5)
    W, H = 48, 40
    for trial in range(60):
        tri = [tuple(int(v) for v in rng.integers(0, [W + 1, H + 1])) for _ in range(3)]
        s = build_screen_triangles(ctx, [tri], W, H)
        out = render(ctx, s)
        got = coverage_set(out["count"])
        exp, amb = exact_coverage(tri, W, H)
        assert (got ^ exp) <= amb, (trial, tri, sorted(got ^ exp)[:10])


@pytest.mark.parametrize("kind", BACKENDS)
def test_depth_plane_matches_exact_interpolation(kind, make_ctx):
    ctx = make_ctx(kind)
    W, H = 96, 64
    tri = [(3, 2), (90, 17), (21, 61)]
    zs = [-0.75, 0.5, 0.125]
    s = build_screen_triangles(ctx, [tri], W, H, z=[zs], depth_test=A.SL_DEPTH_TEST_OFF)
    out = render(ctx, s)
    ys, xs = np.nonzero(out["count"])
    assert xs.size > 1500
    err = 0.0
    for x, y in zip(xs.tolist(), ys.tolist()):
        b = exact_barycentric(tri, x, y)
        ez = float(sum(bi * zi for bi, zi in zip(b, (np_f32(z) for z in zs))))
        err = max(err, abs(float(out["depth"][y, x]) - ez))
    assert err < 2e-5, err


def np_f32(v):
    from fractions import Fraction
    return Fraction(float(np.float32(v)))


@pytest.mark.parametrize("kind", BACKENDS)
def test_perspective_correct_colour(kind, make_ctx):
    """Gouraud colour with w != 1 against FP64 perspective-correct interpolation."""
    ctx = make_ctx(kind)
    W, H = 80, 60
    tri = [(2, 3), (77, 9), (30, 57)]
    ws = [0.5, 4.0, 1.5]
    cols = np.array([[255, 0, 0, 255], [0, 255, 0, 128], [0, 0, 255, 0]], dtype=np.uint8)
    rec = vertex_records([tri], W, H, colors=cols)
    pos = rec[:, 0:12].copy().view(np.float32).reshape(3, 3)
    clip = np.zeros((3, 4), dtype=np.float32)
    for k in range(3):
        clip[k, :3] = pos[k] * np.float32(ws[k])
        clip[k, 3] = ws[k]
    # records with a 4-component position
    rec4 = np.zeros((3, 20), dtype=np.uint8)
    rec4[:, 0:16] = clip.view(np.uint8).reshape(3, 16)
    rec4[:, 16:20] = cols
    vbo = ctx.create_vbo(rec4)
    vao = ctx.create_vao(vbo, None, [(0, 20, A.VERTEX_DIMENSION_4, A.VERTEX_DATA_FLOAT),
                                     (16, 20, A.VERTEX_DIMENSION_4, A.VERTEX_DATA_BYTE)])
    color, depth, fbo, dbg = scenes.make_target(ctx, W, H, True)
    sh = scenes.gouraud_shader(ctx, depth_test=A.SL_DEPTH_TEST_OFF)
    s = scenes.Scene(ctx, W, H, fbo, color, depth, sh, [("draw", A.sl_mesh(vao, 0, 3, A.RENDER_MODE_TRIANGLES, 0), 1)], debug=dbg)
    out = render(ctx, s)
    ys, xs = np.nonzero(out["count"])
    assert xs.size > 1000
    cf = cols.astype(np.float64) * float(np.float32(0.00392156862745))
    worst = 0
    for x, y in zip(xs.tolist(), ys.tolist()):
        b = [float(v) for v in exact_barycentric(tri, x, y)]
        t = [b[k] / ws[k] for k in range(3)]
        st = sum(t)
        bc = [v / st for v in t]
        exp = (bc[0] * cf[0] + bc[1] * cf[1] + bc[2] * cf[2]) * 255.0
        # float -> u8 keeps the low byte of the truncated int (x86 semantics),
        # so compare modulo 256
        exp_u8 = np.trunc(exp).astype(np.int64) & 0xFF
        got = out["color"][y, x].astype(np.int64)
        dist = np.minimum((got - exp_u8) % 256, (exp_u8 - got) % 256)
        worst = max(worst, int(dist.max()))
    assert worst <= 1.0, worst


@pytest.mark.parametrize("kind", BACKENDS)
@pytest.mark.parametrize("func,op", [
    (A.SL_DEPTH_TEST_OFF, lambda a, b: True), (A.SL_DEPTH_TEST_LESS_THAN, lambda a, b: a < b),
    (A.SL_DEPTH_TEST_LESS_EQUAL, lambda a, b: a <= b), (A.SL_DEPTH_TEST_GREATER_THAN, lambda a, b: a > b),
    (A.SL_DEPTH_TEST_GREATER_EQUAL, lambda a, b: a >= b), (A.SL_DEPTH_TEST_EQUAL, lambda a, b: a == b),
    (A.SL_DEPTH_TEST_NOT_EQUAL, lambda a, b: a != b)])
def test_depth_functions(kind, make_ctx, func, op):
    ctx = make_ctx(kind)
    W, H = 32, 32
    tri = [(0, 0), (32, 0), (0, 32)]
    for za, zb in ((0.0, 0.0), (0.25, -0.5), (-0.5, 0.25)):
        s = build_screen_triangles(ctx, [tri, tri], W, H, z=[za, zb], depth_test=func,
                                   colors=[[10, 0, 0, 255]] * 3 + [[0, 20, 0, 255]] * 3)
        s.clear_depth = 0.0
        out = render(ctx, s)
        cnt = out["count"]
        covered = cnt > 0
        d = out["depth"]
        # first triangle against the cleared depth 0.0, second against the result
        da = 0.0
        first = op(np.float32(za), np.float32(da))
        db = za if first else da
        second = op(np.float32(zb), np.float32(db))
        ys, xs = np.nonzero(covered | (np.arange(W)[None, :] + np.arange(H)[:, None] < 31))
        expect_count = int(first) + int(second)
        assert np.all(cnt[ys, xs] == expect_count), (func, za, zb)
        exp_depth = zb if second else (za if first else 0.0)
        if expect_count:
            assert np.all(d[ys, xs] == np.float32(exp_depth))


@pytest.mark.parametrize("kind", BACKENDS)
def test_last_writer_is_max_covering_index(kind, make_ctx):
    ctx = make_ctx(kind)
    rng = np.random.default_rng(9)
    W, H = 40, 32
    tris = [[tuple(int(v) for v in rng.integers(0, [W + 1, H + 1])) for _ in range(3)] for _ in range(24)]
    s = build_screen_triangles(ctx, tris, W, H)
    out = render(ctx, s)
    last = np.zeros((H, W), dtype=np.int64)
    amb_all = set()
    for i, tri in enumerate(tris):
        cov, amb = exact_coverage(tri, W, H)
        amb_all |= amb
        for (x, y) in cov:
            last[y, x] = i + 1
    mism = {(x, y) for y, x in zip(*np.nonzero(last != out["last"]))}
    assert mism <= amb_all, sorted(mism)[:10]


@pytest.mark.parametrize("kind", BACKENDS)
def test_less_is_order_independent_with_unique_depths(kind, make_ctx):
    ctx = make_ctx(kind)
    rng = np.random.default_rng(13)
    W, H = 64, 48
    n = 30
    tris = [[tuple(int(v) for v in rng.integers(0, [W + 1, H + 1])) for _ in range(3)] for _ in range(n)]
    zs = list(np.linspace(-0.9, 0.9, n).astype(np.float32))
    cols = rng.integers(0, 256, size=(n, 4)).repeat(3, axis=0)
    outs = []
    for order in (np.arange(n), rng.permutation(n)):
        s = build_screen_triangles(ctx, [tris[i] for i in order], W, H, z=[zs[i] for i in order],
                                   colors=cols.reshape(n, 3, 4)[order].reshape(-1, 4),
                                   depth_test=A.SL_DEPTH_TEST_LESS_THAN)
        outs.append(render(ctx, s))
    assert np.array_equal(outs[0]["color"], outs[1]["color"])
    assert np.array_equal(outs[0]["depth"].view(np.uint32), outs[1]["depth"].view(np.uint32))


@pytest.mark.parametrize("kind", BACKENDS)
def test_back_face_culling(kind, make_ctx):
    ctx = make_ctx(kind)
    ccw = [(2, 2), (30, 2), (2, 30)]       # counter-clockwise with y up
    cw = [(2, 2), (2, 30), (30, 2)]
    for cull, keep_ccw, keep_cw in ((A.SL_CULL_BACK_FACE, True, False), (A.SL_CULL_FRONT_FACE, False, True),
                                    (A.SL_CULL_OFF, True, True)):
        for tri, keep in ((ccw, keep_ccw), (cw, keep_cw)):
            s = build_screen_triangles(ctx, [tri], 32, 32, cull=cull)
            out = render(ctx, s)
            assert (out["count"].sum() > 0) == keep, (cull, tri)


@pytest.mark.parametrize("kind", BACKENDS)
@pytest.mark.parametrize("mode", [A.SL_BLEND_ALPHA, A.SL_BLEND_PREMULTIPLED_ALPHA, A.SL_BLEND_ADDITIVE, A.SL_BLEND_SCREEN])
def test_blend_modes_vs_fp64(kind, make_ctx, mode):
    ctx = make_ctx(kind)
    W, H = 16, 16
    src = np.array([200, 100, 40, 96], dtype=np.uint8)
    clear = (0.2, 0.6, 0.9, 0.5)
    s = build_screen_triangles(ctx, [[(0, 0), (16, 0), (0, 16)]], W, H, colors=[src] * 3, blend=mode)
    s.clear_rgba = clear
    out = render(ctx, s)
    k = float(np.float32(0.00392156862745))
    sv = src.astype(np.float64) * k
    d = np.floor(np.array(clear) * 255.0) * k
    a = sv[3]
    if mode == A.SL_BLEND_ALPHA:
        da = (1 - a) * d[3] + a
        rgb = (sv[:3] * a + d[:3] * ((1 - a) * d[3])) / da
        exp = np.array([*rgb, da])
    elif mode == A.SL_BLEND_PREMULTIPLED_ALPHA:
        exp = d * (1 - a) + sv
    elif mode == A.SL_BLEND_ADDITIVE:
        exp = sv * a + d
    else:
        exp = sv * a + d * (1 - a)
    exp = np.trunc(np.clip(exp, 0, 1) * 255.0)
    ys, xs = np.nonzero(out["count"])
    got = out["color"][ys, xs].astype(np.float64)
    assert np.max(np.abs(got - exp[None, :])) <= 1.0


@pytest.mark.parametrize("kind", BACKENDS)
def test_bilinear_sampler_vs_fp64(kind, make_ctx):
    """COLOR_TEXTURED program on an RGBA32F texture: output = vertex colour (1)
    x bilinear REPEAT sample, checked against FP64 bilinear filtering with the
    reference's x+1 -> min(x+1, W) neighbour (texel W aliases the next row)."""
    ctx = make_ctx(kind)
    W, H = 48, 40
    TW, TH = 5, 3
    tex = np.arange(TW * TH * 4, dtype=np.float32).reshape(TH, TW, 4) / 64.0 + 0.125
    texh = ctx.create_texture(A.SL_COLOR_RGBA_FLOAT, TW, TH, 1, tex)
    # a screen-filling quad with uv spanning [-0.5, 1.25] x [-0.25, 1.0]
    corners = [(0, 0), (W, 0), (W, H), (0, H)]
    uvs = [(-0.5, -0.25), (1.25, -0.25), (1.25, 1.0), (-0.5, 1.0)]
    verts = []
    for (X, Y), (u, v) in zip(corners, uvs):
        x = (2.0 * X + 1.0) / W - 1.0 if X < W else 1.0
        y = (2.0 * Y + 1.0) / H - 1.0 if Y < H else 1.0
        verts.append([x, y, 0.0, 1.0, 1.0, 1.0, u, v])
    idx = [0, 1, 2, 0, 2, 3]
    vb = np.array([verts[i] for i in idx], dtype=np.float32)
    vbo = ctx.create_vbo(vb)
    vao = ctx.create_vao(vbo, None, [(0, 32, A.VERTEX_DIMENSION_3, A.VERTEX_DATA_FLOAT),
                                     (12, 32, A.VERTEX_DIMENSION_3, A.VERTEX_DATA_FLOAT),
                                     (24, 32, A.VERTEX_DIMENSION_2, A.VERTEX_DATA_FLOAT)])
    u = A.sl_std_uniforms()
    u.mvp[:] = np.eye(4, dtype=np.float32).reshape(-1).tolist()
    u.model[:] = np.eye(4, dtype=np.float32).reshape(-1).tolist()
    u.texture = texh
    u.instance_vbo = A.SL_NO_RESOURCE
    ubo = ctx.create_ubo(u)
    vs = A.sl_vertex_shader(2, A.SL_CULL_OFF, A.SL_PROGRAM_COLOR_TEXTURED)
    fs = A.sl_fragment_shader(2, 1, A.SL_BLEND_OFF, A.SL_DEPTH_TEST_OFF, A.SL_DEPTH_MASK_ON, A.SL_PROGRAM_COLOR_TEXTURED)
    sh = ctx.create_shader(vs, fs, ubo)
    color, depth, fbo, dbg = scenes.make_target(ctx, W, H, True, A.SL_COLOR_RGBA_FLOAT)
    s = scenes.Scene(ctx, W, H, fbo, color, depth, sh, [("draw", A.sl_mesh(vao, 0, 6, A.RENDER_MODE_TRIANGLES, 0), 1)], debug=dbg)
    out = render(ctx, s)
    flat = np.zeros(((TH + 2) * TW, 4))
    flat[:TW * TH] = tex.reshape(-1, 4)

    def texel(x, y):
        return flat[x + TW * y]

    ys, xs = np.nonzero(out["count"])
    assert xs.size > 1500
    worst = 0.0
    for x, y in zip(xs.tolist(), ys.tolist()):
        # affine uv over the quad (w = 1): interpolate in screen space
        uu = -0.5 + 1.75 * x / W
        vv = -0.25 + 1.25 * y / H
        wu = (1.0 if uu < 0 else 0.0) + math.fmod(uu, 1.0)
        wv = (1.0 if vv < 0 else 0.0) + math.fmod(vv, 1.0)
        xf, yf = wu * TW, wv * TH
        x0, y0 = int(xf), int(yf)
        x1, y1 = min(x0 + 1, TW), min(y0 + 1, TH)
        dx, dy = xf - x0, yf - y0
        exp = (texel(x0, y0) * (1 - dx) * (1 - dy) + texel(x0, y1) * (1 - dx) * dy
               + texel(x1, y0) * dx * (1 - dy) + texel(x1, y1) * dx * dy)
        # pixels whose uv lands on a texel boundary may pick either side
        if min(abs(xf - round(xf)), abs(yf - round(yf))) < 1e-3:
            continue
        worst = max(worst, float(np.max(np.abs(out["color"][y, x] - exp))))
    assert worst < 2e-3, worst


@pytest.mark.parametrize("kind", BACKENDS)
def test_clear_values(kind, make_ctx):
    ctx = make_ctx(kind)
    s = build_screen_triangles(ctx, [[(0, 0), (1, 0), (0, 1)]], 24, 20)
    s.clear_rgba = (0.5, 0.25, 1.0, 0.999)
    s.clear_depth = 0.375
    out = render(ctx, s)
    free = out["count"] == 0
    assert free.sum() >= 24 * 20 - 1
    assert np.all(out["color"][free] == np.array([127, 63, 255, 254], dtype=np.uint8))
    assert np.all(out["depth"][free] == np.float32(0.375))


def test_oracle_threads_match_single_thread():
    import oracle
    outs = []
    for n in (1, 4):
        ctx = A.SLContext(oracle.load())
        ctx.num_threads(n)
        s = scenes.build_random_triangles(ctx, seed=2, ntri=512, W=160, H=120, debug=True)
        outs.append(render(ctx, s))
    a, b = outs
    assert np.array_equal(a["color"], b["color"])
    assert np.array_equal(a["depth"].view(np.uint32), b["depth"].view(np.uint32))
    assert np.array_equal(a["last"], b["last"])
    assert a["stats"]["fragments_shaded"] == b["stats"]["fragments_shaded"]
