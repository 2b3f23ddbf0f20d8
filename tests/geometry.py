"""Test-side helpers: exact (rational) statements of the reference's coverage
rule and interpolation, and builders for screen-space test triangles.

The coverage rule restated here is geometric, not a copy of the oracle's
float code: for a triangle with integer (snapped) vertices, row y in
[minY, maxY) covers x in [min(trunc(xa), trunc(xb)), max(trunc(xa), trunc(xb)))
where xa, xb are the exact x positions of the triangle's two edges at height y
(SL_ScanlineBounds.hpp:137-166,310-322; SL_TriRasterizer.cpp:826-916).
"""
from __future__ import annotations

import math
from fractions import Fraction as Fr

import numpy as np

from softlight_amd import api as A
from softlight_amd import scenes


def ndc_for_pixel(X: int, size: int) -> float:
    """NDC coordinate whose viewport snap floor(fma(x+1, size/2, 0)) is X
    (X == size maps to NDC 1.0, the edge of the clip volume)."""
    if X >= size:
        return 1.0
    return (2.0 * X + 1.0) / size - 1.0


def vertex_records(tris, W, H, z=None, w=None, colors=None) -> np.ndarray:
    """Records {f32 x, y, z; u8 rgba} for triangles given in integer screen
    coordinates [(X0, Y0), (X1, Y1), (X2, Y2)] (w = 1 unless given; clip
    coordinates are scaled by w so the snapped position is unchanged)."""
    n = len(tris)
    rec = np.zeros((n * 3, 16), dtype=np.uint8)
    pos = np.zeros((n * 3, 3), dtype=np.float32)
    for t, tri in enumerate(tris):
        for k, (X, Y) in enumerate(tri):
            zz = 0.0 if z is None else z[t][k] if np.ndim(z[t]) else z[t]
            pos[t * 3 + k] = (ndc_for_pixel(X, W), ndc_for_pixel(Y, H), zz)
    rec[:, 0:12] = pos.view(np.uint8).reshape(-1, 12)
    if colors is None:
        rec[:, 12:16] = 255
    else:
        rec[:, 12:16] = np.asarray(colors, dtype=np.uint8).reshape(-1, 4)
    return rec


def build_screen_triangles(ctx, tris, W, H, z=None, colors=None, depth_test=A.SL_DEPTH_TEST_OFF,
                           depth_mask=A.SL_DEPTH_MASK_ON, cull=A.SL_CULL_OFF, blend=A.SL_BLEND_OFF,
                           color_type=A.SL_COLOR_RGBA_8U):
    rec = vertex_records(tris, W, H, z=z, colors=colors)
    return scenes.build_from_vertex_records(
        ctx, rec, W, H, debug=True, color_type=color_type,
        shader_kw=dict(cull=cull, depth_test=depth_test, depth_mask=depth_mask, blend=blend))


def edge_x_at(p, q, y):
    """Exact x of segment p-q at height y (None if horizontal)."""
    (x0, y0), (x1, y1) = p, q
    if y0 == y1:
        return None
    return Fr(x0) + Fr(x1 - x0) * Fr(y - y0, y1 - y0)


def exact_coverage(tri, W, H, margin=Fr(1, 1000)):
    """Covered pixels of one snapped triangle under the span rule, as a set, and
    the set of 'ambiguous' pixels whose row end lies within `margin` of an
    integer (where float rounding of the edge position may move the end)."""
    xs = [p[0] for p in tri]
    ys = [p[1] for p in tri]
    if max(xs) - min(xs) < 1 or max(ys) - min(ys) < 1:
        return set(), set()
    covered, amb = set(), set()
    for y in range(min(ys), max(ys)):
        hits = []
        for a, b in ((0, 1), (1, 2), (2, 0)):
            p, q = tri[a], tri[b]
            lo, hi = sorted((p[1], q[1]))
            if lo <= y <= hi:
                x = edge_x_at(p, q, y)
                if x is not None:
                    hits.append(x)
        if len(hits) < 2:
            continue
        xa, xb = min(hits), max(hits)
        ends = (math.trunc(xa), math.trunc(xb))
        x0, x1 = min(ends), max(ends)
        near = any(abs(v - round(v)) < margin for v in (xa, xb))
        for x in range(max(x0, 0), min(x1, W)):
            if 0 <= y < H:
                covered.add((x, y))
        if near:
            for x in (x0 - 1, x0, x1 - 1, x1):
                amb.add((x, y))
    return covered, amb


def exact_barycentric(tri, x, y):
    """Exact screen-space barycentrics of pixel corner (x, y)."""
    (x0, y0), (x1, y1), (x2, y2) = tri
    den = Fr((x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0))
    b1 = Fr((x - x0) * (y2 - y0) - (x2 - x0) * (y - y0)) / den
    b2 = Fr((x1 - x0) * (y - y0) - (x - x0) * (y1 - y0)) / den
    return 1 - b1 - b2, b1, b2


def render(ctx, scene):
    ctx.reset_stats()
    scene.render()
    out = scene.read()
    out["stats"] = ctx.stats()
    return out


# ---------------------------------------------------------------------------
# wireframe rows (SL_TriRasterizer.cpp:116-202)
# ---------------------------------------------------------------------------
INT32_MIN = -2 ** 31
_M32 = 0xFFFFFFFF


def i32(v: int) -> int:
    """v wrapped to a two's-complement int32."""
    v &= _M32
    return v - (1 << 32) if v >= (1 << 31) else v


def scan_sorted(tri):
    """The vertices ordered by y descending with the scanline setup's three
    conditional swaps (ties keep input order): v0 top, v1 middle, v2 bottom
    (SL_ScanlineBounds.hpp:143-166)."""
    a, b, c = (tuple(p) for p in tri)
    if a[1] < b[1]:
        a, b = b, a
    if a[1] < c[1]:
        a, c = c, a
    if b[1] < c[1]:
        b, c = c, b
    return a, b, c


def line_x(p, q, t):
    """Exact x of the line through p and q at height t (None if horizontal)."""
    if p[1] == q[1]:
        return None
    return Fr(p[0]) + Fr(q[0] - p[0]) * Fr(t - p[1], q[1] - p[1])


def exact_span(tri, t, margin=Fr(1, 1000)):
    """(xMin, xMax, near) of the span the scanline setup gives at height t,
    also outside the triangle's rows: the long edge v0-v2 and the short edge
    v1-v2 below v1, v0-v1 from v1 up, both extended as lines; ends truncate
    toward zero, a horizontal edge gives INT32_MIN (its slope is infinite).
    `near` flags an end within `margin` of an integer."""
    v0, v1, v2 = scan_sorted(tri)
    lo = line_x(v0, v2, t)
    hi = line_x(v1, v2, t) if t < v1[1] else line_x(v0, v1, t)
    ends = [INT32_MIN if e is None else math.trunc(e) for e in (lo, hi)]
    near = any(e is not None and abs(e - round(e)) < margin for e in (lo, hi))
    return min(ends), max(ends), near


def wire_gap(a: int, b: int) -> int:
    dd = i32(a - b)
    ad = i32(-dd) if dd < 0 else dd
    return 1 if ad < 1 else ad


def wire_row_visits(tri, y):
    """Pixels the wireframe raster visits on row y, in visit order (a pixel
    may appear twice), and whether an end of this row's or the row above's
    span lies near an integer."""
    xMin0, xMax0, n0 = exact_span(tri, y)
    xMin1, xMax1, n1 = exact_span(tri, y - 1)
    d0, d1 = wire_gap(xMin0, xMin1), wire_gap(xMax0, xMax1)
    visits = []
    ix, x = 0, xMin0
    while (x & _M32) < (xMax0 & _M32):
        if ix == d0:
            x = max(xMin0, i32(xMax0 - d1 - 1))
        else:
            visits.append(x)
        ix += 1
        x += 1
    return visits, (n0 or n1)


def wire_visits(tri):
    """{(x, y): visits} of one snapped triangle and the set of rows with a
    near-integer span end (where float rounding may move an end)."""
    xs = [p[0] for p in tri]
    ys = [p[1] for p in tri]
    counts, amb = {}, set()
    if max(xs) - min(xs) < 1 or max(ys) - min(ys) < 1:
        return counts, amb
    for y in range(min(ys), max(ys)):
        vis, near = wire_row_visits(tri, y)
        if near:
            amb.add(y)
        for x in vis:
            counts[(x, y)] = counts.get((x, y), 0) + 1
    return counts, amb


# ---------------------------------------------------------------------------
# points and lines
# ---------------------------------------------------------------------------
def ndc_records(ndc, colors=None) -> np.ndarray:
    """Records {f32 x, y, z; u8 rgba} for vertices given directly in NDC (w = 1)."""
    pos = np.asarray(ndc, dtype=np.float32).reshape(-1, 3)
    rec = np.zeros((pos.shape[0], 16), dtype=np.uint8)
    rec[:, 0:12] = pos.view(np.uint8).reshape(-1, 12)
    rec[:, 12:16] = 255 if colors is None else np.asarray(colors, dtype=np.uint8).reshape(-1, 4)
    return rec


def pixel_records(pts, W, H, z=None, colors=None) -> np.ndarray:
    """Records for vertices at integer screen positions (X, Y) (snapped to
    exactly those, X == W and Y == H included)."""
    ndc = [(ndc_for_pixel(X, W), ndc_for_pixel(Y, H), 0.0 if z is None else z[i]) for i, (X, Y) in enumerate(pts)]
    return ndc_records(ndc, colors)


def bresenham_pixels(p, q):
    """Pixels of the segment between integer points p and q, one per step of
    its major axis: in each column (row, for a steep segment) the pixel whose
    centre line is nearest the ideal segment, a tie going to the row (column)
    of the end with the smaller major coordinate (p first when equal).  The
    major axis is y only when |dy| > |dx|.  Returned in walk order."""
    (x1, y1), (x2, y2) = p, q
    steep = abs(x1 - x2) < abs(y1 - y2)
    if steep:
        x1, y1, x2, y2 = y1, x1, y2, x2
    if x1 > x2:
        x1, y1, x2, y2 = x2, y2, x1, y1
    dx, dy = x2 - x1, y2 - y1
    out = []
    for k in range(dx + 1):
        # nearest row to the ideal y1 + k*dy/dx, half-way rounding toward y1
        q_ = Fr(k * abs(dy), dx) if dx else Fr(0)
        n = math.ceil(q_ - Fr(1, 2))
        y = y1 + (n if dy >= 0 else -n)
        out.append((y, x1 + k) if steep else (x1 + k, y))
    return out
