"""Independent checks of oracle functions that the span-rule, depth and
blending tests leave open (each also runs on the HIP library, marked gpu):

  RGBA8 bilinear sampler   a Lambert draw with ambient 1, diffuse 0 and tint 1
                           writes the sampler's own u8 result (u8 -> float ->
                           u8 is the identity in float32); it must be one of the
                           truncations of 255 * (FP64 bilinear of texels/255)
                           at the pixel's exact uv, with the reference's W/H
                           neighbour clamp (SL_Sampler.hpp:186-218).  Catches
                           rounding instead of truncation, a wrapped instead of
                           clamped neighbour, swapped weights.
  Lambert program          trunc(255 * min(albedo * (ambient + diffuse *
                           max(0, -n.l)) * tint, 1)) per colour channel and
                           trunc(255 * albedo.a * tint.a), evaluated in FP64 on
                           both candidate albedos.  Catches a flipped light
                           direction, a missing normalisation or clamp, alpha
                           lit or tinted wrongly.
  nearest sampler          the texel at (uint)(W * wrap(u)), (uint)(H * wrap(v))
                           in ordered layout (SL_Sampler.hpp:122-148).
  triangle clipper         triangles straddling each of the six clip planes
                           (and corners of several): the covered pixels and
                           per-pixel visit counts (depth test off) equal the
                           exact-rational Sutherland-Hodgman clip of the
                           triangle, perspective-divided, snapped, and drawn as
                           the reverse fan (SL_TriProcessor.cpp:370-519) under
                           the exact span rule; depth and perspective-correct
                           colour of the last covering fan triangle match FP64.
                           Catches a wrong intersection parameter, a wrong
                           plane or sign, a dropped vertex (plane and fan
                           order are shown not to be observable).
  scissor                  a screen-covering triangle under a scissor rectangle
                           covers the rectangle (up to the one-pixel band where
                           its clipped corners land on integers) and nothing
                           outside; triangles inside the rectangle draw exactly
                           as without a scissor (SL_ViewportState.cpp:17-111).
"""
from __future__ import annotations

import itertools
import math
from fractions import Fraction as Fr

import numpy as np
import pytest

from softlight_amd import api as A
from softlight_amd import scenes
from tests.geometry import build_screen_triangles, exact_coverage, ndc_for_pixel, render

BACKENDS = [pytest.param("oracle", id="oracle"), pytest.param("gpu", id="gpu", marks=pytest.mark.gpu)]
C8 = float(np.float32(0.00392156862745))        # color_cast<float, uint8_t> factor (SL_Color.hpp:666-669)


@pytest.fixture
def make_ctx(request):
    def _make(kind):
        if kind == "oracle":
            import oracle
            return A.SLContext(oracle.load())
        import softlight_amd
        return A.SLContext(softlight_amd.load_library())
    return _make


# ---------------------------------------------------------------------------
# textured quads through SL_PROGRAM_LAMBERT_TEXTURED
# ---------------------------------------------------------------------------
QUAD_UV = [(-0.5, -0.25), (1.25, -0.25), (1.25, 1.0), (-0.5, 1.0)]


def quad_uv(x, y, W, H):
    """Exact uv at pixel corner (x, y) of the screen-filling quad (w = 1, so
    uv is affine in screen space)."""
    return Fr(-1, 2) + Fr(7, 4) * Fr(x, W), Fr(-1, 4) + Fr(5, 4) * Fr(y, H)


def textured_quad(ctx, W, H, tex, normal=(0.0, 0.0, 1.0), light=(0.0, 0.0, -1.0), ambient=(1.0, 1.0, 1.0),
                  diffuse=(0.0, 0.0, 0.0), tint=(1.0, 1.0, 1.0, 1.0), flags=0, color_type=A.SL_COLOR_RGBA_8U):
    corners = [(0, 0), (W, 0), (W, H), (0, H)]
    verts = []
    for (X, Y), (u, v) in zip(corners, QUAD_UV):
        verts.append([ndc_for_pixel(X, W), ndc_for_pixel(Y, H), 0.0, *normal, u, v])
    vb = np.array([verts[i] for i in (0, 1, 2, 0, 2, 3)], dtype=np.float32)
    vbo = ctx.create_vbo(vb)
    vao = ctx.create_vao(vbo, None, [(0, 32, A.VERTEX_DIMENSION_3, A.VERTEX_DATA_FLOAT),
                                     (12, 32, A.VERTEX_DIMENSION_3, A.VERTEX_DATA_FLOAT),
                                     (24, 32, A.VERTEX_DIMENSION_2, A.VERTEX_DATA_FLOAT)])
    u = A.sl_std_uniforms()
    u.mvp[:] = np.eye(4, dtype=np.float32).reshape(-1).tolist()
    u.model[:] = np.eye(4, dtype=np.float32).reshape(-1).tolist()
    u.light_dir[:] = [*light, 0.0]
    u.ambient[:] = [*ambient, 0.0]
    u.diffuse[:] = [*diffuse, 0.0]
    u.tint[:] = list(tint)
    u.texture = tex
    u.instance_vbo = A.SL_NO_RESOURCE
    u.flags = flags
    ubo = ctx.create_ubo(u)
    vs = A.sl_vertex_shader(2, A.SL_CULL_OFF, A.SL_PROGRAM_LAMBERT_TEXTURED)
    fs = A.sl_fragment_shader(2, 1, A.SL_BLEND_OFF, A.SL_DEPTH_TEST_OFF, A.SL_DEPTH_MASK_ON, A.SL_PROGRAM_LAMBERT_TEXTURED)
    sh = ctx.create_shader(vs, fs, ubo)
    color, depth, fbo, dbg = scenes.make_target(ctx, W, H, True, color_type)
    return scenes.Scene(ctx, W, H, fbo, color, depth, sh,
                        [("draw", A.sl_mesh(vao, 0, 6, A.RENDER_MODE_TRIANGLES, 0), 1)], debug=dbg)


def wrap(u):
    """SL_WrapModeRepeat on an exact value: (u < 0) + (u - trunc(u))."""
    return (1 if u < 0 else 0) + (u - math.trunc(u))


def bilinear_u8_candidates(texels, TW, TH, u, v, delta=Fr(1, 200)):
    """Per channel, the set of u8 values the sampler may return at exact
    (u, v): the truncations of 255 * FP64 bilinear of texels * (1/255) with
    x1 = min(x0 + 1, W), y1 = min(y0 + 1, H) in ordered layout (texel W of a
    row is texel 0 of the next; rows past H read the zeroed padding), allowing
    for `delta` of float error.  None when uv lies too close to a texel edge
    (where float rounding may pick other taps)."""
    xf, yf = wrap(u) * TW, wrap(v) * TH
    if min(abs(xf - round(xf)), abs(yf - round(yf))) < Fr(1, 500):
        return None
    x0, y0 = math.floor(xf), math.floor(yf)
    x1, y1 = min(x0 + 1, TW), min(y0 + 1, TH)
    dx, dy = float(xf - x0), float(yf - y0)
    flat = np.zeros(((TH + 2) * TW, 4))
    flat[:TW * TH] = texels.reshape(-1, 4).astype(np.float64) * C8

    def t(x, y):
        return flat[x + TW * y]
    s = (t(x0, y0) * (1 - dx) * (1 - dy) + t(x0, y1) * (1 - dx) * dy + t(x1, y0) * dx * (1 - dy)
         + t(x1, y1) * dx * dy) * 255.0
    d = float(delta)
    return [sorted({int(math.trunc(c - d)), int(math.trunc(c + d))}) for c in s]


@pytest.mark.parametrize("kind", BACKENDS)
def test_rgba8_bilinear_sampler_vs_fp64(kind, make_ctx):
    ctx = make_ctx(kind)
    W, H, TW, TH = 64, 48, 7, 5
    rng = np.random.default_rng(21)
    texels = rng.integers(0, 256, size=(TH, TW, 4), dtype=np.uint8)
    tex = ctx.create_texture(A.SL_COLOR_RGBA_8U, TW, TH, 1, texels)
    out = render(ctx, textured_quad(ctx, W, H, tex))
    ys, xs = np.nonzero(out["count"])
    assert xs.size == W * H - W - H + 1 or xs.size > 0.9 * W * H
    checked = unambiguous = 0
    for x, y in zip(xs.tolist(), ys.tolist()):
        cand = bilinear_u8_candidates(texels, TW, TH, *quad_uv(x, y, W, H))
        if cand is None:
            continue
        got = out["color"][y, x].tolist()
        for c in range(4):
            assert got[c] in cand[c], (x, y, c, got, cand)
        checked += 1
        unambiguous += all(len(cc) == 1 for cc in cand)
    assert checked > 0.8 * xs.size
    assert unambiguous > 0.9 * checked


def lambert_candidates(alb_cands, n, light, ambient, diffuse, tint, delta=1e-3):
    """Candidate RGBA8 outputs of the Lambert program for every candidate
    sampled albedo (u8 per channel), in FP64."""
    nn = np.asarray(n, np.float64)
    nn = nn / np.linalg.norm(nn)
    lam = max(0.0, -float(np.dot(nn, np.asarray(light, np.float64))))
    outs = []
    for c in range(4):
        vals = set()
        for s in alb_cands[c]:
            a = s * C8
            if c < 3:
                v = min(a * (ambient[c] + diffuse[c] * lam) * tint[c], 1.0) * 255.0
            else:
                v = a * tint[3] * 255.0
            vals |= {int(math.trunc(v - delta)), int(math.trunc(v + delta))}
        outs.append(sorted(vals))
    return outs


@pytest.mark.parametrize("kind", BACKENDS)
@pytest.mark.parametrize("case", [
    # light from the front, green tint pushes channels over 1 (the clamp)
    dict(n=(0.3, -0.2, 0.9), light=(-0.5, 0.4, -0.7), ambient=(0.1, 0.2, 0.3), diffuse=(0.7, 0.6, 0.5),
         tint=(0.9, 1.4, 0.8, 0.7)),
    # light from behind: lambert term clamped to 0, ambient only
    dict(n=(0.0, 0.0, 2.0), light=(0.0, 0.6, 0.8), ambient=(0.45, 0.3, 0.15), diffuse=(0.9, 0.9, 0.9),
         tint=(1.0, 1.0, 1.0, 0.5)),
])
def test_lambert_program_closed_form(kind, make_ctx, case):
    ctx = make_ctx(kind)
    W, H, TW, TH = 48, 40, 6, 4
    rng = np.random.default_rng(22)
    texels = rng.integers(0, 256, size=(TH, TW, 4), dtype=np.uint8)
    tex = ctx.create_texture(A.SL_COLOR_RGBA_8U, TW, TH, 1, texels)
    L = np.asarray(case["light"], np.float64)
    L = tuple(float(v) for v in (L / np.linalg.norm(L)).astype(np.float32))
    out = render(ctx, textured_quad(ctx, W, H, tex, normal=case["n"], light=L, ambient=case["ambient"],
                                    diffuse=case["diffuse"], tint=case["tint"]))
    ys, xs = np.nonzero(out["count"])
    checked = clamped = 0
    for x, y in zip(xs.tolist(), ys.tolist()):
        alb = bilinear_u8_candidates(texels, TW, TH, *quad_uv(x, y, W, H))
        if alb is None:
            continue
        exp = lambert_candidates(alb, case["n"], L, case["ambient"], case["diffuse"], case["tint"])
        got = out["color"][y, x].tolist()
        for c in range(4):
            assert got[c] in exp[c], (x, y, c, got, exp, alb)
        checked += 1
        clamped += got[1] == 255
    assert checked > 0.8 * xs.size
    if case["tint"][1] > 1.0:
        assert clamped > 0


def nearest_texel(texels, TW, TH, u, v):
    """The texel sl_sample_nearest<..., REPEAT> returns at exact (u, v):
    (floor(W * wrap(u)), floor(H * wrap(v))) in ordered layout, x = W being
    texel 0 of the next row and row H the zeroed padding; None within 1/500
    of a texel edge."""
    xf, yf = wrap(u) * TW, wrap(v) * TH
    if min(abs(xf - round(xf)), abs(yf - round(yf))) < Fr(1, 500):
        return None
    flat = np.zeros(((TH + 2) * TW, texels.shape[-1]), dtype=texels.dtype)
    flat[:TW * TH] = texels.reshape(TW * TH, -1)
    return flat[math.floor(xf) + TW * math.floor(yf)]


@pytest.mark.parametrize("kind", BACKENDS)
def test_nearest_sampler_rgba8(kind, make_ctx):
    """Lambert with ambient 1, diffuse 0, tint 1 and SL_UNIFORM_SAMPLE_NEAREST
    writes the sampled texel unchanged."""
    ctx = make_ctx(kind)
    W, H, TW, TH = 64, 48, 7, 5
    rng = np.random.default_rng(23)
    texels = rng.integers(0, 256, size=(TH, TW, 4), dtype=np.uint8)
    tex = ctx.create_texture(A.SL_COLOR_RGBA_8U, TW, TH, 1, texels)
    out = render(ctx, textured_quad(ctx, W, H, tex, flags=A.SL_UNIFORM_SAMPLE_NEAREST))
    ys, xs = np.nonzero(out["count"])
    checked = 0
    for x, y in zip(xs.tolist(), ys.tolist()):
        exp = nearest_texel(texels, TW, TH, *quad_uv(x, y, W, H))
        if exp is None:
            continue
        assert np.array_equal(out["color"][y, x], exp), (x, y, out["color"][y, x], exp)
        checked += 1
    assert checked > 0.9 * xs.size
    # the bilinear draw of the same quad differs (the flag reaches the sampler)
    bil = render(ctx, textured_quad(ctx, W, H, tex))
    assert np.count_nonzero(np.any(bil["color"] != out["color"], axis=-1)) > 0.5 * xs.size


@pytest.mark.parametrize("kind", BACKENDS)
def test_nearest_sampler_rgba32f(kind, make_ctx):
    """COLOR_TEXTURED with white vertex colour and SL_UNIFORM_SAMPLE_NEAREST
    writes the RGBA32F texel (times the interpolated white, 1 within a few
    ulp; neighbouring texels differ by 1/64 or more)."""
    ctx = make_ctx(kind)
    W, H, TW, TH = 48, 40, 5, 3
    texels = (np.arange(TW * TH * 4, dtype=np.float32).reshape(TH, TW, 4) / 64.0 + 0.125).astype(np.float32)
    texh = ctx.create_texture(A.SL_COLOR_RGBA_FLOAT, TW, TH, 1, texels)
    corners = [(0, 0), (W, 0), (W, H), (0, H)]
    verts = [[ndc_for_pixel(X, W), ndc_for_pixel(Y, H), 0.0, 1.0, 1.0, 1.0, u, v] for (X, Y), (u, v) in zip(corners, QUAD_UV)]
    vbo = ctx.create_vbo(np.array([verts[i] for i in (0, 1, 2, 0, 2, 3)], dtype=np.float32))
    vao = ctx.create_vao(vbo, None, [(0, 32, A.VERTEX_DIMENSION_3, A.VERTEX_DATA_FLOAT),
                                     (12, 32, A.VERTEX_DIMENSION_3, A.VERTEX_DATA_FLOAT),
                                     (24, 32, A.VERTEX_DIMENSION_2, A.VERTEX_DATA_FLOAT)])
    u = A.sl_std_uniforms()
    u.mvp[:] = np.eye(4, dtype=np.float32).reshape(-1).tolist()
    u.model[:] = np.eye(4, dtype=np.float32).reshape(-1).tolist()
    u.texture = texh
    u.instance_vbo = A.SL_NO_RESOURCE
    u.flags = A.SL_UNIFORM_SAMPLE_NEAREST
    ubo = ctx.create_ubo(u)
    vs = A.sl_vertex_shader(2, A.SL_CULL_OFF, A.SL_PROGRAM_COLOR_TEXTURED)
    fs = A.sl_fragment_shader(2, 1, A.SL_BLEND_OFF, A.SL_DEPTH_TEST_OFF, A.SL_DEPTH_MASK_ON, A.SL_PROGRAM_COLOR_TEXTURED)
    sh = ctx.create_shader(vs, fs, ubo)
    color, depth, fbo, dbg = scenes.make_target(ctx, W, H, True, A.SL_COLOR_RGBA_FLOAT)
    s = scenes.Scene(ctx, W, H, fbo, color, depth, sh, [("draw", A.sl_mesh(vao, 0, 6, A.RENDER_MODE_TRIANGLES, 0), 1)], debug=dbg)
    out = render(ctx, s)
    ys, xs = np.nonzero(out["count"])
    checked = 0
    for x, y in zip(xs.tolist(), ys.tolist()):
        exp = nearest_texel(texels, TW, TH, *quad_uv(x, y, W, H))
        if exp is None:
            continue
        assert np.allclose(out["color"][y, x], exp, rtol=1e-6, atol=0), (x, y, out["color"][y, x], exp)
        checked += 1
    assert checked > 0.9 * xs.size


# ---------------------------------------------------------------------------
# the 6-plane clipper (SL_TriProcessor.cpp:370-519) in exact arithmetic
# ---------------------------------------------------------------------------
PLANES = [(1, 0, 0, 1), (-1, 0, 0, 1), (0, 1, 0, 1), (0, -1, 0, 1), (0, 0, 1, 1), (0, 0, -1, 1)]


def exact_clip(verts, attrs, wrong_t=False, planes=PLANES):
    """Sutherland-Hodgman of a clip-space polygon against the six planes in
    the reference's order; a vertex is kept when dot(plane, v) >= 0 and an
    edge crossing adds p0 + t (p1 - p0), t = t0 / (t0 - t1), for positions
    and attributes alike.  Exact rationals.  `wrong_t` and `planes` plant
    mistakes (t measured from the other end, another plane order) for the
    test that shows the model tells them apart."""
    poly = list(zip(verts, attrs))
    for pl in planes:
        out = []
        if not poly:
            return []
        p0, a0 = poly[-1]
        t0 = sum(c * v for c, v in zip(pl, p0))
        for p1, a1 in poly:
            t1 = sum(c * v for c, v in zip(pl, p1))
            if (t0 >= 0) != (t1 >= 0):
                t = (t1 / (t1 - t0)) if wrong_t else (t0 / (t0 - t1))
                out.append((tuple(u + t * (w - u) for u, w in zip(p0, p1)), tuple(u + t * (w - u) for u, w in zip(a0, a1))))
            if t1 >= 0:
                out.append((p1, a1))
            p0, a0, t0 = p1, a1, t1
        poly = out
    return poly


def snap_candidates(value, size):
    """Screen coordinate max(floor((ndc + 1) * size / 2), 0) of an exact NDC
    value, with both floors where the value lies within 1e-4 of an integer
    (clipped vertices on the +x / +y planes land exactly on W / H)."""
    s = (value + 1) * Fr(size, 2)
    eps = Fr(1, 10000)
    return sorted({max(math.floor(s - eps), 0), max(math.floor(s + eps), 0)})


def random_straddling_triangle(rng, planes):
    """A clip-space triangle (float32 values, w > 0) with one or two vertices
    outside each plane in `planes`; no vertex lies within 1e-3 of any plane."""
    while True:
        vs = []
        for _ in range(3):
            w = float(rng.uniform(0.4, 2.0))
            vs.append([*(float(v) * w for v in rng.uniform(-0.8, 0.8, size=3)), w])
        for pl in planes:
            axis = next(a for a in range(3) if pl[a] != 0)
            sign = pl[axis]
            for i in rng.choice(3, size=int(rng.integers(1, 3)), replace=False):
                # outside: sign * v[axis] + w < 0
                vs[i][axis] = -sign * vs[i][3] * float(rng.uniform(1.1, 2.5))
        vs = [[float(np.float32(c)) for c in v] for v in vs]
        margin = min(abs(sum(c * q for c, q in zip(pl, v))) for pl in PLANES for v in vs)
        if margin > 1e-3:
            return vs


def fan_model(tri, colors, W, H, forward=False, **mistake):
    """All admissible renderings of one clipped triangle: for every choice of
    the ambiguous vertex snaps, the fan triangles (reverse order; `forward`
    plants the opposite order) as integer screen triangles with their exact
    z (NDC), 1/w and colour at each vertex."""
    verts = [tuple(Fr(c) for c in v) for v in tri]
    attrs = [tuple(Fr(float(np.float32(c * C8))) for c in col) for col in colors]
    poly = exact_clip(verts, attrs, **mistake)
    if len(poly) < 3:
        return []
    opts = []
    for (x, y, z, w), a in poly:
        xs = snap_candidates(x / w, W)
        ys = snap_candidates(y / w, H)
        opts.append([((X, Y), z / w, 1 / w, a) for X in xs for Y in ys])
    models = []
    for choice in itertools.islice(itertools.product(*opts), 256):
        order = range(len(choice) - 2) if forward else range(len(choice) - 3, -1, -1)
        fan = [(choice[0], choice[i + 1], choice[i + 2]) for i in order]
        models.append(fan)
    return models


def coverage_of(fan, W, H):
    count, amb = {}, set()
    for t in fan:
        cov, a = exact_coverage([v[0] for v in t], W, H)
        amb |= a
        for p in cov:
            count[p] = count.get(p, 0) + 1
    return count, amb


def tri_bary(t, x, y):
    (x0, y0), (x1, y1), (x2, y2) = (v[0] for v in t)
    den = Fr((x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0))
    b1 = Fr((x - x0) * (y2 - y0) - (x2 - x0) * (y - y0)) / den
    b2 = Fr((x1 - x0) * (y - y0) - (x - x0) * (y1 - y0)) / den
    return 1 - b1 - b2, b1, b2


def clip_records(tris, colors):
    rec = np.zeros((len(tris) * 3, 20), dtype=np.uint8)
    pos = np.asarray(tris, dtype=np.float32).reshape(-1, 4)
    rec[:, 0:16] = pos.view(np.uint8).reshape(-1, 16)
    rec[:, 16:20] = np.asarray(colors, dtype=np.uint8).reshape(-1, 4)
    return rec


def draw_clip_triangles(ctx, tris, colors, W, H):
    rec = clip_records(tris, colors)
    vbo = ctx.create_vbo(rec)
    vao = ctx.create_vao(vbo, None, [(0, 20, A.VERTEX_DIMENSION_4, A.VERTEX_DATA_FLOAT),
                                     (16, 20, A.VERTEX_DIMENSION_4, A.VERTEX_DATA_BYTE)])
    color, depth, fbo, dbg = scenes.make_target(ctx, W, H, True)
    sh = scenes.gouraud_shader(ctx, depth_test=A.SL_DEPTH_TEST_OFF)
    s = scenes.Scene(ctx, W, H, fbo, color, depth, sh, [("draw", A.sl_mesh(vao, 0, 3 * len(tris), A.RENDER_MODE_TRIANGLES, 0), 1)],
                     debug=dbg)
    return render(ctx, s)


PLANE_SETS = [[pl] for pl in PLANES] + [[PLANES[0], PLANES[2]], [PLANES[1], PLANES[3], PLANES[4]], [PLANES[0], PLANES[5]]]


def pixel_last_colour_depth(fan, x, y, W, H):
    """Exact depth and RGBA8 colour (mod 256, perspective-correct) the last
    fan triangle covering pixel (x, y) gives it, or None."""
    for t in reversed(fan):
        if (x, y) in exact_coverage([v[0] for v in t], W, H)[0]:
            b = tri_bary(t, x, y)
            z = sum(bi * v[1] for bi, v in zip(b, t))
            hw = [bi * v[2] for bi, v in zip(b, t)]
            sw = sum(hw)
            col = [float(sum(h / sw * v[3][c] for h, v in zip(hw, t)) * 255) for c in range(4)]
            return float(z), np.trunc(np.array(col)).astype(np.int64) & 0xFF
    return None


def match_model(models, count_img, depth_img, color_img, W, H):
    """The first admissible rendering whose covered pixels and visit counts
    equal the image's (span-end ambiguous pixels excepted) and whose depth
    (within 2e-5) and colour (within 1 LSB) agree at every other covered
    pixel; returns (fan, pixels checked) or (None, 0)."""
    got = {(int(x), int(y)): int(count_img[y, x]) for y, x in zip(*np.nonzero(count_img))}
    for fan in models:
        cnt, amb = coverage_of(fan, W, H)
        if {p for p in set(cnt) | set(got) if cnt.get(p, 0) != got.get(p, 0)} - amb:
            continue
        ok, n = True, 0
        for (x, y) in cnt:
            if (x, y) in amb or not (0 <= x < W and 0 <= y < H):
                continue
            z, exp = pixel_last_colour_depth(fan, x, y, W, H)
            g = color_img[y, x].astype(np.int64)
            if abs(float(depth_img[y, x]) - z) >= 2e-5 or int(np.minimum((g - exp) % 256, (exp - g) % 256).max()) > 1:
                ok = False
                break
            n += 1
        if ok:
            return fan, n
    return None, 0


@pytest.mark.parametrize("kind", BACKENDS)
@pytest.mark.parametrize("which", range(len(PLANE_SETS)), ids=lambda i: "planes" + "".join(
    "+-"[PLANES.index(p) % 2] + "xyz"[PLANES.index(p) // 2] for p in PLANE_SETS[i]))
def test_clipped_triangles_match_exact_clipping(kind, make_ctx, which):
    ctx = make_ctx(kind)
    W, H = 64, 48
    rng = np.random.default_rng(100 + which)
    checked = 0
    for trial in range(6):
        tri = random_straddling_triangle(rng, PLANE_SETS[which])
        colors = rng.integers(0, 256, size=(3, 4))
        out = draw_clip_triangles(ctx, [tri], colors, W, H)
        models = fan_model(tri, colors, W, H)
        if not models:
            assert out["count"].sum() == 0, (trial, tri)
            continue
        fan, n = match_model(models, out["count"], out["depth"], out["color"], W, H)
        assert fan is not None, (trial, tri)
        assert np.all(out["last"][out["count"] > 0] == 1)
        checked += n
    assert checked > 200


def test_clipper_model_has_power():
    """A rendering made with a planted mistake in the intersection parameter
    (t measured from the other end of the edge) fails to match the exact
    model of the reference's clipper, so an oracle or kernel with that
    mistake would fail the test above.  Run on the oracle's renderings with
    the mistake planted in the model instead: the correct model matches
    every rendering, the wrong one misses.  The plane order and the fan
    order are not observable at this bar: any order of the six half-spaces
    yields the same convex polygon (only where the fan's diagonals run
    changes, and perspective-correct colours agree within 1 LSB across a
    diagonal), and fan triangles sharing an edge never cover a pixel twice
    under the span rule, so the order of the fan only matters where
    snapping folds the polygon; both models are kept here to show it."""
    import oracle
    W, H = 64, 48
    rng = np.random.default_rng(7)
    ctx = A.SLContext(oracle.load())
    misses = {"t": 0, "order": 0, "forward": 0}
    total = 0
    for trial in range(18):
        tri = random_straddling_triangle(rng, [PLANES[trial % 6], PLANES[(trial + 3) % 6]])
        colors = rng.integers(0, 256, size=(3, 4))
        out = draw_clip_triangles(ctx, [tri], colors, W, H)
        models = fan_model(tri, colors, W, H)
        if not models or out["count"].sum() < 20:
            continue
        total += 1
        assert match_model(models, out["count"], out["depth"], out["color"], W, H)[0] is not None
        for key, kw in (("t", dict(wrong_t=True)), ("order", dict(planes=PLANES[::-1])), ("forward", dict(forward=True))):
            wrong = fan_model(tri, colors, W, H, **kw)
            if not wrong or match_model(wrong, out["count"], out["depth"], out["color"], W, H)[0] is None:
                misses[key] += 1
    assert total >= 8
    assert misses["t"] == total, misses
    assert misses["order"] == 0 and misses["forward"] == 0, misses


# ---------------------------------------------------------------------------
# scissor (SL_ViewportState.cpp:17-111)
# ---------------------------------------------------------------------------
def screen_covering_triangle(ctx, W, H):
    """One triangle whose clip to the screen (or to a scissor rectangle) is
    the whole rectangle: NDC (-1, -1), (3, -1), (-1, 3)."""
    ndc = np.array([[-1.0, -1.0, 0.0], [3.0, -1.0, 0.0], [-1.0, 3.0, 0.0]], dtype=np.float32)
    rec = np.zeros((3, 16), dtype=np.uint8)
    rec[:, 0:12] = ndc.view(np.uint8).reshape(3, 12)
    rec[:, 12:16] = 255
    return scenes.build_from_vertex_records(ctx, rec, W, H, debug=True,
                                            shader_kw=dict(cull=A.SL_CULL_OFF, depth_test=A.SL_DEPTH_TEST_OFF))


@pytest.mark.parametrize("kind", BACKENDS)
@pytest.mark.parametrize("rect", [None, (10, 7, 30, 20), (0, 0, 64, 48), (-6, 5, 25, 60), (40, 30, 100, 100)])
def test_scissor_rectangle_closed_form(kind, make_ctx, rect):
    """The clipped polygon is the rectangle; its clipped corners land on
    integers, where float rounding may snap them one pixel lower, so the
    drawn set lies between the rectangle less its last column and row and the
    rectangle grown by one column and row at its low sides."""
    ctx = make_ctx(kind)
    W, H = 64, 48
    if rect is not None:
        ctx.set_scissor(*rect)
    out = render(ctx, screen_covering_triangle(ctx, W, H))
    sx, sy, sw, sh = rect if rect is not None else (0, 0, W, H)
    x0, y0 = max(sx, 0), max(sy, 0)
    x1, y1 = min(sx + sw, W), min(sy + sh, H)
    cnt = out["count"]
    may = np.zeros((H, W), bool)
    may[max(y0 - 1, 0):y1, max(x0 - 1, 0):x1] = True
    must = np.zeros((H, W), bool)
    must[y0:y1 - 1, x0:x1 - 1] = True
    assert np.all(cnt[~may] == 0), rect
    assert np.all(cnt[must] == 1), rect
    assert cnt.max() <= 1


@pytest.mark.parametrize("kind", BACKENDS)
def test_scissor_leaves_inside_triangles_unchanged(kind, make_ctx):
    """Triangles inside the scissor rectangle (vertices at pixel centres) draw
    bit-identically with and without the scissor: its matrix and viewport
    cancel for them."""
    W, H = 80, 60
    rng = np.random.default_rng(31)
    tris = [[tuple(int(v) for v in rng.integers([12, 9], [60, 45])) for _ in range(3)] for _ in range(20)]
    zs = list(rng.uniform(-0.9, 0.9, size=20).astype(np.float32))
    cols = rng.integers(0, 256, size=(60, 4))
    outs = []
    for rect in (None, (10, 8, 55, 45)):
        ctx = make_ctx(kind)
        if rect:
            ctx.set_scissor(*rect)
        s = build_screen_triangles(ctx, tris, W, H, z=zs, depth_test=A.SL_DEPTH_TEST_LESS_THAN,
                                   colors=cols)
        outs.append(render(ctx, s))
    a, b = outs
    assert a["count"].sum() > 500
    for k in ("color", "last", "count"):
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(a["depth"].view(np.uint32), b["depth"].view(np.uint32))
