"""Smoothed span coverage: the differentiable-coverage mode (no reference
counterpart).  What pins the oracle and what each check catches:

  sigma = 0            the smoothed image is the colour attachment bit for bit
                       (oracle and GPU): the hard limit of the model
  closed form          one triangle and a two-triangle quad at known integer
                       screen positions: every pixel of the smoothed image
                       against the model evaluated in float64 from the exact
                       (rational) edge positions -- the leak of the right
                       end of the left triangle into the first uncovered
                       pixel, w = exp(-(x + 1 - e) / sigma), the leak of the
                       background into a triangle's first pixel, internal
                       edges between two colours, untouched interior pixels
                       (catches a wrong window, direction, sign or sample
                       position)
  adjoint identity     the smoothing is linear in the colour image:
                       <G, S(C)> = <S^T(G), C> for random C, G
  finite differences   dL/d(screen x, y) of every vertex of the quad against
                       central differences of the forward through the
                       screen-space offsets (vertices moved off the pixel
                       grid in x so no span end sits on a window edge; the
                       row through the bottom vertices carries no loss)
GPU (marked gpu): the same closed forms and sigma = 0 check, and images, image
gradients and position gradients against the oracle on a 4-view icosphere.
"""
from __future__ import annotations

import math
from fractions import Fraction as Fr

import numpy as np
import pytest

from softlight_amd import api as A
from softlight_amd import scenes
from tests.geometry import build_screen_triangles

BACKENDS = [pytest.param("oracle", id="oracle"), pytest.param("gpu", id="gpu", marks=pytest.mark.gpu)]

W = H = 20
TRI = [(2, 1), (13, 1), (2, 15)]
QUAD = [[(2, 1), (13, 1), (2, 15)], [(13, 1), (13, 15), (2, 15)]]
C1 = (64, 128, 192, 255)
C2 = (200, 40, 90, 160)


@pytest.fixture
def make_ctx(request):
    def _make(kind):
        lib = request.getfixturevalue("oracle_lib" if kind == "oracle" else "gpu_lib")
        return A.SLContext(lib)
    return _make


def _flat(tris, colors):
    return [c for c, t in zip(colors, tris) for _ in range(3)]


def render_tris(ctx, tris, colors):
    s = build_screen_triangles(ctx, tris, W, H, colors=_flat(tris, colors), color_type=A.SL_COLOR_RGBA_FLOAT)
    vis = ctx.create_texture(A.SL_COLOR_RGBA_FLOAT, W, H)
    ctx.attach_visibility(s.fbo, vis)
    s.render()
    ctx.finish()
    return s, vis


def smooth(ctx, s, sigma):
    out = ctx.create_texture(A.SL_COLOR_RGBA_FLOAT, W, H)
    ctx.smooth_coverage(s.fbo, out, sigma)
    return ctx.texture_read(out)


def u8f(c):
    return np.array([np.float32(v) * np.float32(0.00392156862745) for v in c], dtype=np.float64)


def edge_x(p, q, y):
    (x0, y0), (x1, y1) = p, q
    return Fr(x0) + Fr(x1 - x0) * Fr(y - y0, y1 - y0)


def expected_quad(sigma, tris, colors):
    """The model in float64 from exact edge positions, for the fixtures above
    (left edge x = 2, the shared diagonal, and for the quad a right edge x = 13)."""
    img = np.zeros((H, W, 4))
    cols = [u8f(c) for c in colors]
    diag = (TRI[1], TRI[2])
    for y in range(1, 15):
        e = edge_x(*diag, y)
        xr = math.trunc(e)
        covered1 = range(2, xr)
        for x in covered1:
            img[y, x] = cols[0]
        if len(tris) == 2:
            for x in range(xr, 13):
                img[y, x] = cols[1]
        hard = img[y].copy()
        if len(covered1) > 0:
            # background leaks into the first pixel of the left triangle (e = 2)
            w = math.exp(-(3 - 2) / sigma)
            img[y, 2] = (1 - w) * hard[2] + w * hard[1]
            # the diagonal: right end of the left triangle leaks into pixel xr
            if xr < W:
                w = math.exp(-((xr + 1) - float(e)) / sigma)
                img[y, xr] = (1 - w) * hard[xr] + w * hard[xr - 1]
        elif len(tris) == 2 and xr <= 2:
            # no left-triangle pixel: the right triangle starts at xr with the background before it
            w = math.exp(-((xr + 1) - float(e)) / sigma)
            img[y, xr] = (1 - w) * hard[xr] + w * hard[xr - 1]
        if len(tris) == 2:
            # right edge x = 13 of the right triangle leaks into pixel 13
            w = math.exp(-(14 - 13) / sigma)
            img[y, 13] = (1 - w) * hard[13] + w * hard[12]
    return img


@pytest.mark.parametrize("kind", BACKENDS)
def test_sigma_zero_is_the_hard_image(kind, make_ctx):
    ctx = make_ctx(kind)
    s, _ = render_tris(ctx, QUAD, [C1, C2])
    hard = ctx.texture_read(s.color)
    out = smooth(ctx, s, 0.0)
    np.testing.assert_array_equal(out.view(np.uint32), hard.view(np.uint32))
    assert not np.array_equal(smooth(ctx, s, 0.5), hard)


@pytest.mark.parametrize("kind", BACKENDS)
@pytest.mark.parametrize("sigma", [0.5, 0.1])
@pytest.mark.parametrize("shape", ["triangle", "quad"])
def test_closed_form(kind, make_ctx, sigma, shape):
    ctx = make_ctx(kind)
    tris, colors = ([TRI], [C1]) if shape == "triangle" else (QUAD, [C1, C2])
    s, _ = render_tris(ctx, tris, colors)
    out = smooth(ctx, s, sigma).astype(np.float64)
    exp = expected_quad(sigma, tris, colors)
    # row 1 holds vertices exactly on pixel samples: its span ends are integers
    # that float rounding may put on either side of a window edge
    out[1] = exp[1] = 0
    np.testing.assert_allclose(out, exp, atol=2e-6, rtol=0)


def test_adjoint_identity(oracle_lib):
    ctx = A.SLContext(oracle_lib)
    s, _ = render_tris(ctx, QUAD, [C1, C2])
    rng = np.random.default_rng(3)
    C = rng.uniform(-1, 1, (H, W, 4)).astype(np.float32)
    G = rng.uniform(-1, 1, (H, W, 4)).astype(np.float32)
    ctx.texture_set_texels(s.color, C)
    out = smooth(ctx, s, 0.3)
    g = ctx.create_texture(A.SL_COLOR_RGBA_FLOAT, W, H, 1, G)
    gi = ctx.create_texture(A.SL_COLOR_RGBA_FLOAT, W, H)
    ctx.smooth_coverage_backward(s.fbo, g, 0.3, gi)
    lhs = float(np.sum(G.astype(np.float64) * out))
    rhs = float(np.sum(ctx.texture_read(gi).astype(np.float64) * C))
    assert abs(lhs - rhs) <= 1e-5 * max(1.0, abs(lhs))


def test_position_gradients_vs_finite_differences(oracle_lib):
    ctx = A.SLContext(oracle_lib)
    s, _ = render_tris(ctx, QUAD, [C1, C2])
    nv = 6
    base = np.zeros((1, nv, 2), dtype=np.float32)
    base[..., 0] = 0.3                      # off the pixel grid: no span end on a window edge
    off = ctx.create_vbo(base)
    grad = ctx.create_vbo(np.zeros((1, nv, 2), dtype=np.float32))
    ctx.attach_positions(s.fbo, grad, off, nv)
    rng = np.random.default_rng(5)
    G = rng.uniform(-1, 1, (H, W, 4)).astype(np.float32)
    # row 1 runs through the two bottom vertices: moving one of them in y tilts
    # the bottom edge through the row's samples (not differentiable there)
    G[1] = 0
    g = ctx.create_texture(A.SL_COLOR_RGBA_FLOAT, W, H, 1, G)
    gi = ctx.create_texture(A.SL_COLOR_RGBA_FLOAT, W, H)
    sigma = 0.5
    ctx.smooth_coverage_backward(s.fbo, g, sigma, gi)
    ctx.finish()
    analytic = ctx.vbo_read(grad, nv * 8, dtype=np.float32).reshape(nv, 2).astype(np.float64)

    def loss(o):
        ctx.vbo_assign(off, o)
        return float(np.sum(G.astype(np.float64) * smooth(ctx, s, sigma)))

    h = 1e-2
    fd = np.zeros((nv, 2))
    for v in range(nv):
        for a in range(2):
            p = base.copy(); p[0, v, a] += h
            m = base.copy(); m[0, v, a] -= h
            fd[v, a] = (loss(p) - loss(m)) / (2 * h)
    scale = np.abs(fd).max()
    assert scale > 0.1
    np.testing.assert_allclose(analytic, fd, atol=2e-3 * scale, rtol=0)


@pytest.mark.gpu
@pytest.mark.parametrize("sigma", [0.0, 0.5, 0.05])
def test_gpu_matches_oracle_on_views(oracle_lib, gpu_lib, sigma):
    outs = []
    for lib in (oracle_lib, gpu_lib):
        ctx = A.SLContext(lib)
        sc = scenes.DiffScene(ctx, level=3, views=4, size=96, views_per_fbo=4, tex_size=64, edge_views=1)
        sc.forward()
        t = sc.targets[0]
        nv = sc.n_vertices
        grad = ctx.create_vbo(np.zeros((len(t["views"]), nv, 2), dtype=np.float32))
        ctx.attach_positions(t["fbo"], grad, A.SL_INVALID_ID, nv)
        out = ctx.create_texture(A.SL_COLOR_RGBA_FLOAT, t["W"], t["H"])
        ctx.smooth_coverage(t["fbo"], out, sigma)
        rng = np.random.default_rng(9)
        G = rng.uniform(-1, 1, (t["H"], t["W"], 4)).astype(np.float32)
        g = ctx.create_texture(A.SL_COLOR_RGBA_FLOAT, t["W"], t["H"], 1, G)
        gi = ctx.create_texture(A.SL_COLOR_RGBA_FLOAT, t["W"], t["H"])
        ctx.smooth_coverage_backward(t["fbo"], g, sigma, gi)
        ctx.finish()
        outs.append(dict(img=ctx.texture_read(out), gi=ctx.texture_read(gi), hard=ctx.texture_read(t["color"]),
                         pg=ctx.vbo_read(grad, len(t["views"]) * nv * 8, dtype=np.float32)))
    o, g = outs
    np.testing.assert_array_equal(o["hard"].view(np.uint32), g["hard"].view(np.uint32))
    if sigma == 0.0:
        np.testing.assert_array_equal(g["img"].view(np.uint32), g["hard"].view(np.uint32))
    np.testing.assert_allclose(g["img"], o["img"], atol=1e-5, rtol=0)
    np.testing.assert_allclose(g["gi"], o["gi"], atol=1e-5, rtol=0)
    scale = max(1.0, float(np.abs(o["pg"]).max()))
    np.testing.assert_allclose(g["pg"], o["pg"], atol=1e-4 * scale, rtol=0)
    if sigma > 0.0:
        assert np.abs(o["pg"]).max() > 0


# ---------------------------------------------------------------------------
# The batched smoothed-coverage workload (8 meshes of 5k..60k triangles, the
# sigma sweep and the sigma -> 0 check), GPU against the oracle
# ---------------------------------------------------------------------------
def _smooth_workload(lib, sigma, threads=8, **kw):
    ctx = A.SLContext(lib)
    ctx.num_threads(threads)
    args = dict(seed=6, meshes=8, min_tris=5000, max_tris=60000, views=8, size=512, views_per_fbo=8, tex_size=512)
    args.update(kw)
    sc = scenes.SmoothScene(ctx, **args)
    sc.make_targets(sigma)
    sc.forward(sigma)
    sc.backward(sigma)
    ctx.finish()
    out = dict(hard=[], vis=[], smooth=[], gimg=[], pg=[])
    for t in sc.targets:
        out["hard"].append(ctx.texture_read(t["color"]))
        out["vis"].append(ctx.texture_read(t["vis"]))
        out["smooth"].append(ctx.texture_read(t["smooth"]))
        out["gimg"].append(ctx.texture_read(t["grad_image"]))
        out["pg"].append(ctx.vbo_read(t["pos_grad"], len(t["views"]) * len(sc.meshes) * sc.n_vertices_max * 8,
                                      dtype=np.float32))
    out["gcol"] = [ctx.vbo_read(g, sc.pos[m].shape[0] * 16, dtype=np.float32) for m, g in enumerate(sc.grad_vbos)]
    out["gtex"] = ctx.texture_read(sc.grad_tex)
    out["tris"] = sc.mesh_tris
    ctx.close()
    return out


def _assert_smooth_parity(o, g, sigma):
    for a, b in zip(o["hard"], g["hard"]):
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    for a, b in zip(o["vis"], g["vis"]):
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    for a, b, h in zip(o["smooth"], g["smooth"], g["hard"]):
        if sigma == 0.0:
            np.testing.assert_array_equal(b.view(np.uint32), h.view(np.uint32))
        np.testing.assert_allclose(b, a, atol=1e-5, rtol=0)
    for a, b in zip(o["gimg"], g["gimg"]):
        np.testing.assert_allclose(b, a, atol=1e-5, rtol=0)
    for key in ("pg", "gcol"):
        for a, b in zip(o[key], g[key]):
            scale = max(1e-6, float(np.abs(a).max()))
            np.testing.assert_allclose(b, a, atol=1e-4 * scale, rtol=0)
    scale = float(np.abs(o["gtex"]).max())
    np.testing.assert_allclose(g["gtex"], o["gtex"], atol=1e-4 * scale, rtol=0)


@pytest.mark.gpu
@pytest.mark.parametrize("sigma", [0.0, 1e-2, 1e-3, 1e-4])
def test_gpu_matches_oracle_on_batched_smooth_workload(oracle_lib, gpu_lib, sigma):
    """8 meshes (log-uniform 5k..60k triangles) seen by 8 views at 512x512:
    hard image and visibility bit-exact, smoothed image and image gradient
    within 1e-5, position, colour and texel gradients within 1e-4 of the
    largest, against the oracle; sigma = 0 returns the hard image bit for bit."""
    o = _smooth_workload(oracle_lib, sigma)
    g = _smooth_workload(gpu_lib, sigma)
    assert min(o["tris"]) >= 5000 * 0.8 and max(o["tris"]) <= 60000 * 1.2
    _assert_smooth_parity(o, g, sigma)


@pytest.mark.gpu
def test_gpu_matches_oracle_on_full_smooth_workload(oracle_lib, gpu_lib):
    """The full configuration: 64 views at 512x512 of the 8 meshes, sigma = 1e-2."""
    o = _smooth_workload(oracle_lib, 1e-2, views=64)
    g = _smooth_workload(gpu_lib, 1e-2, views=64)
    _assert_smooth_parity(o, g, 1e-2)
