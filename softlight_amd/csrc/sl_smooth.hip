// sl_smooth.hip -- smoothed coverage: the differentiable-coverage mode of the
// draw path (no reference counterpart; the reference renders forward only).
//
// Model.  Pixel (x, y) samples the point (x + 1, y): the span rule covers it
// when e_min < x + 1 <= e_max at row y (SL_ScanlineBounds.hpp:310-322).
// Where the visibility winners of two neighbouring pixels differ and an edge
// of one of the winners separates their samples, the colour before the edge
// leaks into the pixel past it with the weight
//
//   w(d) = 2 s(-d / sigma) = 1 - d / sqrt(d^2 + sigma^2),
//
// s(u) = (1 + u / sqrt(1 + u^2)) / 2 the algebraic sigmoid and d > 0 the
// perpendicular distance of the receiving pixel's sample from the edge's
// line.  w is 1 on the edge, so the image stays continuous where a pixel's
// winner changes, and goes to 0 for every d > 0 as sigma -> 0; its tail falls
// as (sigma / d)^2 / 2, so the weight and its derivative -sigma^2 / r^3 stay
// well inside float range for sigma far below the pixel spacing.  Two
// separable passes over the hard colour image C:
//
//   X(x, y) = (1 - wx) C(x, y) + wx C(x - 1, y)
//     the edge between the samples of x - 1 and x is a row-span end in
//     [x, x + 1): the right end of the left pixel's winner if it lies there,
//     else the left end of the right pixel's (the span rule puts every tie on
//     the left pixel, so the receiver's distance is never 0);
//   O(x, y) = X(x, y) + sum wy (X(x, y') - X(x, y))
//     the edge between the samples of rows y and y + 1 in column x + 1 is a
//     column-span end v in [y, y + 1]: the high end of row y's winner if it
//     lies there, else the low end of row y + 1's.  Row y receives (distance
//     v - y) unless its sample lies on the edge (v = y); then row y + 1
//     receives (distance 1).
//
// Which edge bounds a winner is decided on the positions as rasterised (the
// raster's own span arithmetic), so the choice agrees with the hard winners;
// the weights use the positions plus the screen-space offsets, which makes
// them smooth functions of the offsets.  A clipped winner is the polygon the
// clipper leaves: its row spans are the union of the fan pieces the raster
// drew, its column spans come from its edges, and a vertex the clipper made
// has no offset and takes no position gradient (the edge's other end does).
// sigma = 0 returns the hard image bit for bit.
//
// Gradients (winners fixed): the image gradient is the transposed filter
// (column pass, then row pass); a boundary's weight gets
// dL/dw = <G(receiver), In(source) - In(receiver)> (In = C for the row pass,
// X for the column pass), dL/dd = dL/dw * -sigma^2 / r^3, and d flows to the
// screen positions of the edge's two vertices.  Every pixel recomputes the
// boundaries it needs from the visibility attachment (the winner's triangle
// is re-fetched and re-transformed), so nothing but the attachment is stored.
#include "sl_device.h"
#include "sl_stages.h"
#include "sl_smooth.h"

namespace slg {

__device__ __forceinline__ uint32_t find_saved_draw(const GradDesc* D, uint32_t n, uint32_t seq)
{
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1)
    {
        const uint32_t mid = (lo + hi) >> 1;
        if (D[mid].d.seqBase < seq) lo = mid;
        else hi = mid;
    }
    return lo;
}

// A vertex of a winner: rasterised position, position plus offset, vertex id.
struct SVert
{
    float rx, ry;             // snapped screen position (what the raster drew)
    float sx, sy;             // plus the vertex's offset
    uint32_t vid;
    bool orig;                // a vertex of the triangle (not made by the clipper)
};

constexpr int kPolyMax = 9;

// A pixel's winner: its triangle, or (clipped) the polygon the clipper left.
struct Shape
{
    bool ok;
    bool clipped;
    uint32_t k;               // saved draw
    int n;                    // polygon size (clipped)
    SVert t[3];               // the triangle (unclipped)
    SVert poly[kPolyMax];     // the polygon (clipped)
};

__device__ __forceinline__ SVert make_vert(const SmoothParams& P, uint32_t k, float4 clip, const DrawDesc& D, uint32_t vid,
                                           bool orig)
{
    const float4 sp = divide_and_snap(clip, D.vp);
    SVert v;
    v.rx = sp.x;
    v.ry = sp.y;
    float ox = 0.f, oy = 0.f;
    if (orig && P.posOffset != nullptr && vid < P.numVerts)
    {
        const float2 o = P.posOffset[(uint64_t)k * P.numVerts + vid];
        ox = o.x;
        oy = o.y;
    }
    v.sx = sp.x + ox;
    v.sy = sp.y + oy;
    v.vid = vid;
    v.orig = orig;
    return v;
}

__device__ __forceinline__ Shape winner_shape(const SmoothParams& P, uint32_t seq)
{
    Shape S;
    S.ok = false;
    S.clipped = false;
    S.n = 3;
    if (seq == 0u) return S;
    const uint32_t k = find_saved_draw(P.draws, P.numDraws, seq);
    const DrawDesc& D = P.draws[k].d;
    if (seq <= D.seqBase) return S;
    const uint64_t t = (uint64_t)(seq - 1u - D.seqBase);
    if (t >= D.numTris) return S;
    const uint64_t inst = t / D.trisPerInstance;
    const uint64_t ti = t - inst * D.trisPerInstance;
    const uint64_t e = D.elementBegin + 3u * ti;
    uint32_t vid[3] = {(uint32_t)e, (uint32_t)(e + 1), (uint32_t)(e + 2)};
    if (D.indexed)
    {
#pragma unroll
        for (int j = 0; j < 3; ++j) vid[j] = (uint32_t)fetch_index(D, e + j);
    }
    const bool all3[3] = {true, true, true};
    float4 pos[3];
    vertex_positions3(D, P.uniforms, vid, inst, all3, pos);
#pragma unroll
    for (int j = 0; j < 3; ++j) pos[j] = mat_mul(D.scissor, pos[j]);
    const int status = clip_status(pos[0], pos[1], pos[2]);
    S.k = k;
    if (status == 3)
    {
#pragma unroll
        for (int j = 0; j < 3; ++j) S.t[j] = make_vert(P, k, pos[j], D, vid[j], true);
        S.ok = true;
        return S;
    }
    if (status != 1) return S;
    float4 pb[kPolyMax], tb[kPolyMax];
    const PolyBuf<float4> poly{pb, 1}, tmp{tb, 1};
    const PolyBuf<Vary4> none{nullptr, 0};
    const int n = clip_polygon<false>(pos, nullptr, 0, poly, none, tmp, none);
    if (n < 3) return S;
    for (int i = 0; i < n; ++i)
    {
        // the clipper copies the vertices it keeps unchanged
        const float4 q = pb[i];
        int from = -1;
#pragma unroll
        for (int j = 2; j >= 0; --j)
            if (__float_as_uint(q.x) == __float_as_uint(pos[j].x) && __float_as_uint(q.y) == __float_as_uint(pos[j].y) &&
                __float_as_uint(q.z) == __float_as_uint(pos[j].z) && __float_as_uint(q.w) == __float_as_uint(pos[j].w))
                from = j;
        const uint32_t v = from == 0 ? vid[0] : (from == 1 ? vid[1] : (from == 2 ? vid[2] : 0u));
        S.poly[i] = make_vert(P, k, q, D, v, from >= 0);
    }
    S.n = n;
    S.clipped = true;
    S.ok = true;
    return S;
}

// A span end and the edge it lies on.
struct End
{
    float e;
    SVert p, q;
};

// One triangle's span ends at row y on the rasterised positions, with
// SL_ScanlineBounds::init's vertex order and :310-322's arithmetic.
__device__ __forceinline__ void tri_row_ends(SVert a, SVert b, SVert c, float yf, End& left, End& right)
{
    SVert tmp;
    if (a.ry < b.ry) { tmp = a; a = b; b = tmp; }
    if (a.ry < c.ry) { tmp = a; a = c; c = tmp; }
    if (b.ry < c.ry) { tmp = b; b = c; c = tmp; }
    const float d0 = yf - a.ry, d1 = yf - b.ry;
    const float alpha = d0 * rcpf(c.ry - a.ry);
    End lo, hi;
    lo.e = ((c.rx - a.rx) * alpha) + a.rx;
    lo.p = a;
    lo.q = c;
    if (d1 < 0.f)
    {
        hi.e = (((c.rx - b.rx) * rcpf(c.ry - b.ry)) * d1) + b.rx;
        hi.p = b;
        hi.q = c;
    }
    else
    {
        hi.e = (((b.rx - a.rx) * rcpf(b.ry - a.ry)) * d0) + a.rx;
        hi.p = a;
        hi.q = b;
    }
    const bool loIsMax = lo.e >= hi.e;
    right = loIsMax ? lo : hi;
    left = loIsMax ? hi : lo;
}

// The winner's row span at y; a clipped polygon's is the union of the fan
// pieces (0, j + 1, j + 2) the raster drew whose rows hold y.
__device__ __forceinline__ bool row_ends(const Shape& S, int y, End& left, End& right)
{
    const float yf = (float)y;
    if (!S.clipped)
    {
        tri_row_ends(S.t[0], S.t[1], S.t[2], yf, left, right);
        return true;
    }
    bool any = false;
    for (int j = S.n - 3; j >= 0; --j)
    {
        const SVert a = S.poly[0], b = S.poly[j + 1], c = S.poly[j + 2];
        const float minX = fminf(fminf(a.rx, b.rx), c.rx), maxX = fmaxf(fmaxf(a.rx, b.rx), c.rx);
        const float minY = fminf(fminf(a.ry, b.ry), c.ry), maxY = fmaxf(fmaxf(a.ry, b.ry), c.ry);
        if ((maxX - minX < 1.f) || (maxY - minY < 1.f)) continue;   // push_bin's reject
        if (y < f2i(minY) || y >= f2i(maxY)) continue;
        End l, r;
        tri_row_ends(a, b, c, yf, l, r);
        if (!any || l.e < left.e) left = l;
        if (!any || r.e > right.e) right = r;
        any = true;
    }
    return any;
}

// The crossing of an edge with the line b = c, along a:
// Pa + ((Qa - Pa) (c - Pb)) / (Qb - Pb), exact for vertices on the pixel
// lattice whenever the crossing is.
__device__ __forceinline__ float edge_crossing(float pa, float pb, float qa, float qb, float c)
{
    return pa + __fdiv_rn(__fmul_rn(qa - pa, c - pb), qb - pb);
}

__device__ __forceinline__ void col_edge(const SVert& p, const SVert& q, float u, bool& any, End& low, End& high)
{
    if (p.rx == q.rx || u < fminf(p.rx, q.rx) || u > fmaxf(p.rx, q.rx)) return;
    const float v = edge_crossing(p.ry, p.rx, q.ry, q.rx, u);
    if (!any || v < low.e) { low.e = v; low.p = p; low.q = q; }
    if (!any || v > high.e) { high.e = v; high.p = p; high.q = q; }
    any = true;
}

// The winner's column span at u: the lowest and highest crossing of its
// edges with the vertical line x = u, edges in polygon order.
__device__ __forceinline__ bool col_ends(const Shape& S, float u, End& low, End& high)
{
    bool any = false;
    if (!S.clipped)
    {
        col_edge(S.t[0], S.t[1], u, any, low, high);
        col_edge(S.t[1], S.t[2], u, any, low, high);
        col_edge(S.t[2], S.t[0], u, any, low, high);
        return any;
    }
    for (int i = 0; i < S.n; ++i) col_edge(S.poly[i], S.poly[i + 1 == S.n ? 0 : i + 1], u, any, low, high);
    return any;
}

// A smoothed boundary: the edge with the offsets in pass coordinates -- a
// along the pass (x for the row pass, y for the column pass), b across it --
// and the receiving sample (a = sample, b = cross); the distance along a is
// dir (sample - crossing).
struct Bnd
{
    bool ok;
    bool gp, gq;              // the end is a vertex of the triangle (has an offset and a gradient)
    bool recvLow;             // column pass: row y receives (else row y + 1)
    float sample, cross, dir;
    float pa, pb, qa, qb;
    uint32_t k, vp, vq;
};

__device__ __forceinline__ Bnd make_bnd(uint32_t k, const End& E, bool rowPass, float sample, float cross, float dir)
{
    Bnd B;
    B.ok = true;
    B.gp = E.p.orig;
    B.gq = E.q.orig;
    B.recvLow = false;
    B.sample = sample;
    B.cross = cross;
    B.dir = dir;
    B.pa = rowPass ? E.p.sx : E.p.sy;
    B.pb = rowPass ? E.p.sy : E.p.sx;
    B.qa = rowPass ? E.q.sx : E.q.sy;
    B.qb = rowPass ? E.q.sy : E.q.sx;
    B.k = k;
    B.vp = E.p.vid;
    B.vq = E.q.vid;
    return B;
}

__device__ __forceinline__ Bnd no_bnd()
{
    Bnd B;
    B.ok = false;
    B.gp = B.gq = false;
    B.recvLow = false;
    B.sample = B.cross = B.dir = 0.f;
    B.pa = B.pb = B.qa = B.qb = 0.f;
    B.k = B.vp = B.vq = 0u;
    return B;
}

__device__ __forceinline__ uint32_t vis_seq(const SmoothParams& P, int x, int y)
{
    return __float_as_uint(P.vis[(uint64_t)y * P.W + x].x);
}

// The row-pass boundary between pixels x - 1 and x of row y: pixel x receives.
__device__ __forceinline__ Bnd boundary_x(const SmoothParams& P, int x, int y)
{
    if (x <= 0 || x >= P.W || y < 0 || y >= P.H) return no_bnd();
    const uint32_t sl = vis_seq(P, x - 1, y), sr = vis_seq(P, x, y);
    if (sl == sr) return no_bnd();
    const float lo = (float)x, hi = (float)(x + 1), yf = (float)y;
    End l, r;
    if (sl != 0u)
    {
        const Shape L = winner_shape(P, sl);
        if (L.ok && row_ends(L, y, l, r) && r.e >= lo && r.e < hi) return make_bnd(L.k, r, true, hi, yf, 1.f);
    }
    if (sr != 0u)
    {
        const Shape R = winner_shape(P, sr);
        if (R.ok && row_ends(R, y, l, r) && l.e >= lo && l.e < hi) return make_bnd(R.k, l, true, hi, yf, 1.f);
    }
    return no_bnd();
}

__device__ __forceinline__ Bnd col_bnd(uint32_t k, const End& E, int y, int x)
{
    const float yl = (float)y, u = (float)(x + 1);
    const bool low = E.e > yl;
    Bnd B = make_bnd(k, E, false, low ? yl : yl + 1.f, u, low ? -1.f : 1.f);
    B.recvLow = low;
    return B;
}

// The column-pass boundary between rows y and y + 1 of column x (sample
// x + 1): row y receives unless its sample lies on the edge.
__device__ __forceinline__ Bnd boundary_y(const SmoothParams& P, int x, int y)
{
    if (x < 0 || x >= P.W || y < 0 || y + 1 >= P.H) return no_bnd();
    const uint32_t sa = vis_seq(P, x, y), sb = vis_seq(P, x, y + 1);
    if (sa == sb) return no_bnd();
    const float lo = (float)y, hi = (float)(y + 1), u = (float)(x + 1);
    End l, h;
    if (sa != 0u)
    {
        const Shape A = winner_shape(P, sa);
        if (A.ok && col_ends(A, u, l, h) && h.e >= lo && h.e <= hi) return col_bnd(A.k, h, y, x);
    }
    if (sb != 0u)
    {
        const Shape B = winner_shape(P, sb);
        if (B.ok && col_ends(B, u, l, h) && l.e >= lo && l.e <= hi) return col_bnd(B.k, l, y, x);
    }
    return no_bnd();
}

// Signed perpendicular distance of the receiving sample from the edge's line.
__device__ __forceinline__ float bnd_distance(const Bnd& B)
{
    // sample - crossing, measured from P so that no large coordinates cancel
    const float daxis = B.dir * ((B.sample - B.pa) - __fdiv_rn(__fmul_rn(B.qa - B.pa, B.cross - B.pb), B.qb - B.pb));
    const float ea = B.qa - B.pa, eb = B.qb - B.pb;
    const float len = __fsqrt_rn(fmaf(ea, ea, __fmul_rn(eb, eb)));
    return daxis * __fdiv_rn(fabsf(eb), len);
}

// w(d) = 1 - d / r, r = sqrt(d^2 + sigma^2); for d >= 0 written as
// sigma^2 / (r (r + d)) so that it keeps its precision where it is small.
__device__ __forceinline__ float smooth_weight(float d, float sigma, float& r)
{
    const float s2 = __fmul_rn(sigma, sigma);
    r = __fsqrt_rn(fmaf(d, d, s2));
    return d >= 0.f ? __fdiv_rn(s2, __fmul_rn(r, r + d)) : 1.f - __fdiv_rn(d, r);
}

__device__ __forceinline__ float bnd_weight(const Bnd& B, float sigma)
{
    float r;
    return B.ok ? smooth_weight(bnd_distance(B), sigma, r) : 0.f;
}

// d(distance)/d(Pa, Pb, Qa, Qb): crossing ea = Pa + (Qa - Pa) t,
// t = (c - Pb) / (Qb - Pb); d = dir (sample - ea) |Qb - Pb| / |Q - P|.
__device__ __forceinline__ float bnd_partials(const Bnd& B, float out[4])
{
    const float inv = 1.f / (B.qb - B.pb);
    const float t = (B.cross - B.pb) * inv, s = (B.qa - B.pa) * inv;
    const float da = B.dir * ((B.sample - B.pa) - (B.qa - B.pa) * t);
    const float ePa = B.dir * (1.f - t), eQa = B.dir * t, ePb = -B.dir * (1.f - t) * s, eQb = -B.dir * t * s;
    const float ea = B.qa - B.pa, eb = B.qb - B.pb;
    const float len2 = ea * ea + eb * eb, len = sqrtf(len2), len3 = len2 * len;
    const float g = fabsf(eb) / len, sg = eb > 0.f ? 1.f : -1.f;
    const float gQa = -fabsf(eb) * ea / len3, gPa = -gQa;
    const float gQb = sg * ea * ea / len3, gPb = -gQb;
    out[0] = -ePa * g + da * gPa;
    out[1] = -ePb * g + da * gPb;
    out[2] = -eQa * g + da * gQa;
    out[3] = -eQb * g + da * gQb;
    return da * g;
}

__device__ __forceinline__ float4 lerp_px(float4 c, float4 l, float w)
{
    const float u = 1.f - w;
    return make_float4((u * c.x) + (w * l.x), (u * c.y) + (w * l.y), (u * c.z) + (w * l.z), (u * c.w) + (w * l.w));
}

__device__ __forceinline__ float4 add_scaled_diff(float4 o, float w, float4 a, float4 b)
{
    return make_float4(o.x + w * (a.x - b.x), o.y + w * (a.y - b.y), o.z + w * (a.z - b.z), o.w + w * (a.w - b.w));
}

// The boundary pass: each pixel's row-pass boundary (x - 1 | x) and
// column-pass boundary (y | y + 1), as weights (NaN: none; the column
// weight's sign bit: row y receives).  Every pass below reads them, so each
// boundary's winners are fetched once instead of by every pixel that uses it.
__global__ __launch_bounds__(256) void k_smooth_bounds(SmoothParams P)
{
    const uint64_t n = (uint64_t)P.W * P.H;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    {
        const int x = (int)(i % (uint64_t)P.W), y = (int)(i / (uint64_t)P.W);
        const Bnd Bx = boundary_x(P, x, y);
        const Bnd By = boundary_y(P, x, y);
        const float wx = Bx.ok ? bnd_weight(Bx, P.sigma) : __builtin_nanf("");
        float wy = __builtin_nanf("");
        if (By.ok)
        {
            const float w = bnd_weight(By, P.sigma);
            wy = By.recvLow ? -w : w;
        }
        P.bounds[i] = make_float2(wx, wy);
    }
}

// A stored boundary: present, its weight and (column pass) whether row y receives.
struct StoredBnd
{
    bool ok, recvLow;
    float w;
};

__device__ __forceinline__ StoredBnd stored_x(const SmoothParams& P, uint64_t i)
{
    const float v = P.bounds[i].x;
    return StoredBnd{!isnan(v), false, v};
}

__device__ __forceinline__ StoredBnd stored_y(const SmoothParams& P, uint64_t i)
{
    const float v = P.bounds[i].y;
    return StoredBnd{!isnan(v), (bool)signbit(v), fabsf(v)};
}

// X(x, y): the row pass at one pixel.
__device__ __forceinline__ float4 row_pass(const SmoothParams& P, int x, int y)
{
    const uint64_t i = (uint64_t)y * P.W + x;
    const StoredBnd B = stored_x(P, i);
    return B.ok ? lerp_px(P.color[i], P.color[i - 1], B.w) : P.color[i];
}

__global__ __launch_bounds__(256) void k_smooth(SmoothParams P)
{
    const uint64_t n = (uint64_t)P.W * P.H;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    {
        const int x = (int)(i % (uint64_t)P.W), y = (int)(i / (uint64_t)P.W);
        if (!(P.sigma > 0.f))
        {
            P.out[i] = P.color[i];
            continue;
        }
        const float4 X = row_pass(P, x, y);
        float4 o = X;
        if (y > 0)
        {
            // rows y - 1, y: row y receives when the tie rule says so
            const StoredBnd Bd = stored_y(P, i - P.W);
            if (Bd.ok && !Bd.recvLow) o = add_scaled_diff(o, Bd.w, row_pass(P, x, y - 1), X);
        }
        const StoredBnd Bu = stored_y(P, i);   // rows y, y + 1
        if (Bu.ok && Bu.recvLow) o = add_scaled_diff(o, Bu.w, row_pass(P, x, y + 1), X);
        P.out[i] = o;
    }
}

// dL/dX at (x, y): the column pass transposed.
__device__ __forceinline__ float4 grad_x_image(const SmoothParams& P, int x, int y)
{
    const uint64_t i = (uint64_t)y * P.W + x;
    const float4 g = P.grad[i];
    float4 gx = g;
    if (y > 0)
    {
        const StoredBnd Bd = stored_y(P, i - P.W);
        if (Bd.ok)
        {
            if (Bd.recvLow) gx = add_scaled_diff(gx, Bd.w, P.grad[i - P.W], make_float4(0.f, 0.f, 0.f, 0.f));   // this row is the source
            else gx = add_scaled_diff(gx, -Bd.w, g, make_float4(0.f, 0.f, 0.f, 0.f));                          // this row receives
        }
    }
    const StoredBnd Bu = stored_y(P, i);
    if (Bu.ok)
    {
        if (Bu.recvLow) gx = add_scaled_diff(gx, -Bu.w, g, make_float4(0.f, 0.f, 0.f, 0.f));
        else gx = add_scaled_diff(gx, Bu.w, P.grad[i + P.W], make_float4(0.f, 0.f, 0.f, 0.f));
    }
    return gx;
}

__device__ __forceinline__ float dot4f(float4 a, float4 b)
{
    return ((a.x * b.x) + (a.y * b.y)) + ((a.z * b.z) + (a.w * b.w));
}

__device__ __forceinline__ void scatter_position(const SmoothParams& P, const Bnd& B, float dLdd, bool rowPass)
{
    if (dLdd == 0.f) return;
    float p[4];
    bnd_partials(B, p);
    const int ax = rowPass ? 0 : 1, bx = rowPass ? 1 : 0;
    if (B.gp && B.vp < P.numVerts)
    {
        float* d = reinterpret_cast<float*>(P.posGrad + (uint64_t)B.k * P.numVerts + B.vp);
        unsafeAtomicAdd(d + ax, dLdd * p[0]);
        unsafeAtomicAdd(d + bx, dLdd * p[1]);
    }
    if (B.gq && B.vq < P.numVerts)
    {
        float* d = reinterpret_cast<float*>(P.posGrad + (uint64_t)B.k * P.numVerts + B.vq);
        unsafeAtomicAdd(d + ax, dLdd * p[2]);
        unsafeAtomicAdd(d + bx, dLdd * p[3]);
    }
}

// dL/dd from dL/dw: w = 1 - d / r gives dw/dd = -sigma^2 / r^3.
__device__ __forceinline__ float dL_dd(const Bnd& B, float sigma, float dLdw)
{
    float r;
    (void)smooth_weight(bnd_distance(B), sigma, r);
    const float q = sigma / r;
    return dLdw * (-(q * q) / r);
}

__global__ __launch_bounds__(256) void k_smooth_backward(SmoothParams P)
{
    const uint64_t n = (uint64_t)P.W * P.H;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    {
        const int x = (int)(i % (uint64_t)P.W), y = (int)(i / (uint64_t)P.W);
        if (!(P.sigma > 0.f))
        {
            P.out[i] = P.grad[i];
            continue;
        }
        // dL/dC = the row pass transposed over dL/dX
        const float4 gx = grad_x_image(P, x, y);
        const StoredBnd Sx = stored_x(P, i);
        const float wx = Sx.ok ? Sx.w : 0.f;
        float4 gc = make_float4((1.f - wx) * gx.x, (1.f - wx) * gx.y, (1.f - wx) * gx.z, (1.f - wx) * gx.w);
        if (x + 1 < P.W)
        {
            const StoredBnd Sn = stored_x(P, i + 1);
            if (Sn.ok) gc = add_scaled_diff(gc, Sn.w, grad_x_image(P, x + 1, y), make_float4(0.f, 0.f, 0.f, 0.f));
        }
        P.out[i] = gc;
        if (P.posGrad == nullptr) continue;
        // this pixel's row-pass boundary (x - 1 | x) and the column-pass
        // boundary (y | y + 1), rebuilt for their edges' vertices
        if (Sx.ok)
        {
            const Bnd Bx = boundary_x(P, x, y);
            const float dLdw = dot4f(gx, add_scaled_diff(make_float4(0.f, 0.f, 0.f, 0.f), 1.f, P.color[i - 1], P.color[i]));
            scatter_position(P, Bx, dL_dd(Bx, P.sigma, dLdw), true);
        }
        if (stored_y(P, i).ok)
        {
            const Bnd By = boundary_y(P, x, y);
            const float4 X0 = row_pass(P, x, y), X1 = row_pass(P, x, y + 1);
            const float dLdw = By.recvLow ? dot4f(P.grad[i], add_scaled_diff(make_float4(0.f, 0.f, 0.f, 0.f), 1.f, X1, X0))
                                          : dot4f(P.grad[i + P.W], add_scaled_diff(make_float4(0.f, 0.f, 0.f, 0.f), 1.f, X0, X1));
            scatter_position(P, By, dL_dd(By, P.sigma, dLdw), false);
        }
    }
}

static unsigned smooth_grid(const SmoothParams& p)
{
    const uint64_t n = (uint64_t)p.W * p.H;
    uint64_t g = (n + 255) / 256;
    return (unsigned)(g > 65536 ? 65536 : (g == 0 ? 1 : g));
}

hipError_t launch_smooth(const SmoothParams& p, hipStream_t s)
{
    (void)hipGetLastError();   // drop a stale error from an earlier call
    if (p.W <= 0 || p.H <= 0) return hipSuccess;
    if (p.sigma > 0.f) hipLaunchKernelGGL(k_smooth_bounds, dim3(smooth_grid(p)), dim3(256), 0, s, p);
    hipLaunchKernelGGL(k_smooth, dim3(smooth_grid(p)), dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_smooth_backward(const SmoothParams& p, hipStream_t s)
{
    (void)hipGetLastError();   // drop a stale error from an earlier call
    if (p.W <= 0 || p.H <= 0) return hipSuccess;
    if (p.sigma > 0.f) hipLaunchKernelGGL(k_smooth_bounds, dim3(smooth_grid(p)), dim3(256), 0, s, p);
    hipLaunchKernelGGL(k_smooth_backward, dim3(smooth_grid(p)), dim3(256), 0, s, p);
    return hipGetLastError();
}

}  // namespace slg
