"""Blit (SL_Context::blit, SL_BlitProcessor::blit_nearest): closed forms on the
oracle and the GPU, and GPU against the oracle on random regions.

  same-size blit     = the source flipped vertically (rows are read bottom-up,
                       SL_BlitProcessor.cpp:643-644)
  2x magnification   out[y][x] = in[H-1-floor(y/2)][floor(x/2)] (16.16 steps)
  conversions        RGBA32F -> RGBA8 = low byte of trunc(c*255); RGBA8 ->
                       RGBA32F = c * 0.00392156862745f; R32F -> RGBA8 =
                       (0, 0, trunc(r*255), 255)
  regions            bit-exact GPU vs oracle, including the reference's use of
                       absolute destination x for the source column
"""
from __future__ import annotations

import numpy as np
import pytest

from softlight_amd import api as A
from softlight_amd import scenes

BACKENDS = [pytest.param("oracle", id="oracle"), pytest.param("gpu", id="gpu", marks=pytest.mark.gpu)]


@pytest.fixture
def make_ctx(request):
    def _make(kind):
        lib = request.getfixturevalue("oracle_lib" if kind == "oracle" else "gpu_lib")
        return A.SLContext(lib)
    return _make


def rgba8(seed, w, h):
    return scenes.uniform_u8(seed, w * h * 4).reshape(h, w, 4)


def rgbaf(seed, w, h):
    return scenes.uniform(seed, w * h * 4, -0.25, 1.25).reshape(h, w, 4).astype(np.float32)


@pytest.mark.parametrize("backend", BACKENDS)
def test_same_size_blit_flips_rows(make_ctx, backend):
    ctx = make_ctx(backend)
    src = rgba8(1, 37, 23)
    a = ctx.create_texture(A.SL_COLOR_RGBA_8U, 37, 23, 1, src)
    b = ctx.create_texture(A.SL_COLOR_RGBA_8U, 37, 23)
    ctx.blit(b, a)
    np.testing.assert_array_equal(ctx.texture_read(b), src[::-1])


@pytest.mark.parametrize("backend", BACKENDS)
def test_double_size_blit(make_ctx, backend):
    ctx = make_ctx(backend)
    W, H = 19, 11
    src = rgba8(2, W, H)
    a = ctx.create_texture(A.SL_COLOR_RGBA_8U, W, H, 1, src)
    b = ctx.create_texture(A.SL_COLOR_RGBA_8U, 2 * W, 2 * H)
    ctx.blit(b, a)
    y = np.arange(2 * H)[:, None]
    x = np.arange(2 * W)[None, :]
    np.testing.assert_array_equal(ctx.texture_read(b), src[H - 1 - y // 2, x // 2])


@pytest.mark.parametrize("backend", BACKENDS)
def test_blit_conversions(make_ctx, backend):
    ctx = make_ctx(backend)
    W, H = 16, 9
    f = rgbaf(3, W, H)
    u = rgba8(4, W, H)
    r = scenes.uniform(5, W * H, 0.0, 1.0).reshape(H, W).astype(np.float32)
    tf = ctx.create_texture(A.SL_COLOR_RGBA_FLOAT, W, H, 1, f)
    tu = ctx.create_texture(A.SL_COLOR_RGBA_8U, W, H, 1, u)
    tr = ctx.create_texture(A.SL_COLOR_R_FLOAT, W, H, 1, r)
    o8 = ctx.create_texture(A.SL_COLOR_RGBA_8U, W, H)
    of = ctx.create_texture(A.SL_COLOR_RGBA_FLOAT, W, H)

    def u8_of(v):
        t = np.trunc(v.astype(np.float32) * np.float32(255.0))
        t = np.where((t >= -2147483648.0) & (t < 2147483648.0), t, -2147483648.0)
        return (t.astype(np.int64) & 0xFF).astype(np.uint8)

    ctx.blit(o8, tf)
    np.testing.assert_array_equal(ctx.texture_read(o8), u8_of(f)[::-1])
    ctx.blit(of, tu)
    np.testing.assert_array_equal(ctx.texture_read(of), (u.astype(np.float32) * np.float32(0.00392156862745))[::-1])
    ctx.blit(o8, tr)
    exp = np.zeros((H, W, 4), np.uint8)
    exp[..., 2] = u8_of(r)
    exp[..., 3] = 255
    np.testing.assert_array_equal(ctx.texture_read(o8), exp[::-1])


def test_blit_rejects_unsupported(oracle_lib):
    ctx = A.SLContext(oracle_lib)
    a = ctx.create_texture(A.SL_COLOR_R_32U, 4, 4)
    b = ctx.create_texture(A.SL_COLOR_RGBA_8U, 4, 4)
    with pytest.raises(A.SLError):
        ctx.blit(b, a)
    with pytest.raises(A.SLError):
        ctx.blit(b, b)


REGIONS = [
    ((0, 0, 64, 48), (0, 0, 100, 70)),       # whole source, larger target region
    ((8, 4, 40, 30), (10, 5, 90, 60)),       # sub-rectangles (absolute x quirk)
    ((0, 0, 64, 48), (0, 0, 130, 200)),      # rows past the target height
    ((20, 10, 64, 48), (50, 60, 100, 70)),   # source rows wrapping below zero are skipped
]


@pytest.mark.gpu
@pytest.mark.parametrize("fmt", ["u8_u8", "f_u8", "u8_f", "f_f", "r_u8"])
@pytest.mark.parametrize("region", range(len(REGIONS)))
def test_gpu_blit_matches_oracle(oracle_lib, gpu_lib, fmt, region):
    src_kind, dst_kind = fmt.split("_")
    kinds = {"u8": A.SL_COLOR_RGBA_8U, "f": A.SL_COLOR_RGBA_FLOAT, "r": A.SL_COLOR_R_FLOAT}
    W, H, OW, OH = 64, 48, 100, 70
    if src_kind == "u8":
        src = rgba8(7, W, H)
    elif src_kind == "f":
        src = rgbaf(7, W, H)
    else:
        src = scenes.uniform(7, W * H, -0.1, 1.1).reshape(H, W).astype(np.float32)
    outs = []
    for lib in (oracle_lib, gpu_lib):
        ctx = A.SLContext(lib)
        a = ctx.create_texture(kinds[src_kind], W, H, 1, src)
        b = ctx.create_texture(kinds[dst_kind], OW, OH)
        s, d = REGIONS[region]
        ctx.blit(b, a, s, d)
        outs.append(ctx.texture_read(b))
    np.testing.assert_array_equal(outs[0].view(np.uint8), outs[1].view(np.uint8))
