// sl_smooth.hip -- smoothed span coverage: the differentiable-coverage mode
// of the draw path (no reference counterpart; the reference renders forward
// only).
//
// Model.  A pixel x of row y is covered by a triangle when its span ends e_min,
// e_max at that row (the float values SL_ScanlineBounds::step computes before
// truncation, SL_ScanlineBounds.hpp:310-322) satisfy e_min < x + 1 <= e_max,
// i.e. the pixel's sample sits at u = x + 1.  Where the hard winners of two
// neighbouring pixels x-1 and x differ and a span end of one of them lies
// between their samples, e in [x, x + 1), that boundary is smeared to the
// right with an exponential kernel of scale sigma:
//
//   out(x) = (1 - w) * C(x) + w * C(x-1),   w = exp(-(x + 1 - e) / sigma)
//
// where C is the hard colour image.  The boundary is the right end (e_max) of
// the left pixel's triangle when that lies in the window, else the left end
// (e_min) of the right pixel's triangle.  w is continuous in e (it reaches 1
// as e reaches the sample, where the hard winner of x changes) and goes to 0
// for every e in the window as sigma -> 0, ties included, so sigma = 0 gives
// the hard image bit for bit.  Boundaries of clipped triangles are not
// smoothed.
//
// Gradients (winners fixed): the image gradient of the hard colours is the
// transposed filter, and the boundary position gets
//   dL/de = sum_ch G(x) (C(x-1) - C(x)) * w / sigma,
// which flows to the screen positions of the two vertices of the edge that
// defines e (e = Px + (Qx - Px) t, t = (y - Py) / (Qy - Py)).
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

// The span of a pixel's winning triangle at one row, with what the gradient
// of each end needs.
struct WinnerSpan
{
    bool ok;                  // an unclipped triangle
    uint32_t k;               // saved draw
    uint32_t vid[3];
    float sx[3], sy[3];       // screen positions (snapped, plus the offsets)
    float lo, hi;             // span ends at the row
    int loP, loQ, hiP, hiQ;   // vertex slots of the edges that give lo and hi
};

__device__ __forceinline__ WinnerSpan winner_span(const SmoothParams& P, uint32_t seq, int y)
{
    WinnerSpan S;
    S.ok = false;
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
    if (clip_status(pos[0], pos[1], pos[2]) != 3) return S;
    S.k = k;
#pragma unroll
    for (int j = 0; j < 3; ++j)
    {
        const float4 sp = divide_and_snap(pos[j], D.vp);
        float ox = 0.f, oy = 0.f;
        if (P.posOffset != nullptr && vid[j] < P.numVerts)
        {
            const float2 o = P.posOffset[(uint64_t)k * P.numVerts + vid[j]];
            ox = o.x;
            oy = o.y;
        }
        S.sx[j] = sp.x + ox;
        S.sy[j] = sp.y + oy;
        S.vid[j] = vid[j];
    }
    // SL_ScanlineBounds::init: vertices by y descending, with the reference's swaps
    int a = 0, b = 1, c = 2, tmp;
    if (S.sy[a] < S.sy[b]) { tmp = a; a = b; b = tmp; }
    if (S.sy[a] < S.sy[c]) { tmp = a; a = c; c = tmp; }
    if (S.sy[b] < S.sy[c]) { tmp = b; b = c; c = tmp; }
    const float ax = S.sx[a], ay = S.sy[a], bx = S.sx[b], by = S.sy[b], cx = S.sx[c], cy = S.sy[c];
    // SL_ScanlineBounds::step, the raster's own arithmetic
    const float yf = (float)y;
    const float d0 = yf - ay, d1 = yf - by;
    const float alpha = d0 * rcpf(cy - ay);
    S.lo = ((cx - ax) * alpha) + ax;
    S.loP = a;
    S.loQ = c;
    if (d1 < 0.f)
    {
        S.hi = (((cx - bx) * rcpf(cy - by)) * d1) + bx;
        S.hiP = b;
        S.hiQ = c;
    }
    else
    {
        S.hi = (((bx - ax) * rcpf(by - ay)) * d0) + ax;
        S.hiP = a;
        S.hiQ = b;
    }
    S.ok = true;
    return S;
}

// The boundary between pixels x-1 and x of row y, if their hard winners differ
// and a span end lies in [x, x + 1): its position and the edge that gives it.
struct Boundary
{
    bool ok;
    float e;
    WinnerSpan S;
    int p, q;                 // vertex slots of the edge
};

__device__ __forceinline__ Boundary boundary_at(const SmoothParams& P, const float4* visRow, int x, int y)
{
    Boundary B;
    B.ok = false;
    if (x <= 0 || x >= P.W) return B;
    const uint32_t sl = __float_as_uint(visRow[x - 1].x), sr = __float_as_uint(visRow[x].x);
    if (sl == sr) return B;
    if (sl != 0u)
    {
        const WinnerSpan L = winner_span(P, sl, y);
        if (L.ok)
        {
            const bool loIsMax = L.lo >= L.hi;
            const float em = loIsMax ? L.lo : L.hi;
            if (em >= (float)x && em < (float)(x + 1))
            {
                B.ok = true;
                B.e = em;
                B.S = L;
                B.p = loIsMax ? L.loP : L.hiP;
                B.q = loIsMax ? L.loQ : L.hiQ;
                return B;
            }
        }
    }
    if (sr != 0u)
    {
        const WinnerSpan R = winner_span(P, sr, y);
        if (R.ok)
        {
            const bool loIsMax = R.lo >= R.hi;
            const float en = loIsMax ? R.hi : R.lo;
            if (en >= (float)x && en < (float)(x + 1))
            {
                B.ok = true;
                B.e = en;
                B.S = R;
                B.p = loIsMax ? R.hiP : R.loP;
                B.q = loIsMax ? R.hiQ : R.loQ;
            }
        }
    }
    return B;
}

__device__ __forceinline__ float leak_weight(const Boundary& B, int x, float sigma)
{
    return B.ok ? expf(-((float)(x + 1) - B.e) / sigma) : 0.f;
}

__global__ __launch_bounds__(256) void k_smooth(SmoothParams P)
{
    const uint64_t n = (uint64_t)P.W * P.H;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    {
        const int x = (int)(i % (uint64_t)P.W), y = (int)(i / (uint64_t)P.W);
        const float4 c = P.color[i];
        float4 o = c;
        if (P.sigma > 0.f)
        {
            const Boundary B = boundary_at(P, P.vis + (uint64_t)y * P.W, x, y);
            if (B.ok)
            {
                const float w = leak_weight(B, x, P.sigma);
                const float4 l = P.color[i - 1];
                const float u = 1.f - w;
                o = make_float4((u * c.x) + (w * l.x), (u * c.y) + (w * l.y), (u * c.z) + (w * l.z), (u * c.w) + (w * l.w));
            }
        }
        P.out[i] = o;
    }
}

// d(edge x at row y) / d(endpoint coordinates): e = Px + (Qx - Px) t,
// t = (y - Py) / (Qy - Py).
__device__ __forceinline__ void edge_partials(float px, float py, float qx, float qy, float y, float& dPx, float& dPy,
                                              float& dQx, float& dQy)
{
    const float inv = 1.f / (qy - py);
    const float t = (y - py) * inv;
    const float s = (qx - px) * inv;
    dPx = 1.f - t;
    dQx = t;
    dPy = -(1.f - t) * s;
    dQy = -t * s;
}

__global__ __launch_bounds__(256) void k_smooth_backward(SmoothParams P)
{
    const uint64_t n = (uint64_t)P.W * P.H;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    {
        const int x = (int)(i % (uint64_t)P.W), y = (int)(i / (uint64_t)P.W);
        const float4* visRow = P.vis + (uint64_t)y * P.W;
        const float4 g = P.grad[i];
        if (!(P.sigma > 0.f))
        {
            P.out[i] = g;
            continue;
        }
        const Boundary BL = boundary_at(P, visRow, x, y);       // between x-1 and x
        const Boundary BR = boundary_at(P, visRow, x + 1, y);   // between x and x+1
        const float wl = leak_weight(BL, x, P.sigma);
        const float wr = leak_weight(BR, x + 1, P.sigma);
        float4 gi = make_float4((1.f - wl) * g.x, (1.f - wl) * g.y, (1.f - wl) * g.z, (1.f - wl) * g.w);
        if (BR.ok)
        {
            const float4 gn = P.grad[i + 1];
            gi = make_float4(gi.x + wr * gn.x, gi.y + wr * gn.y, gi.z + wr * gn.z, gi.w + wr * gn.w);
        }
        P.out[i] = gi;
        // this pixel's left boundary: dL/de, then the edge's vertices
        if (BL.ok && P.posGrad != nullptr)
        {
            const float4 c = P.color[i], l = P.color[i - 1];
            const float dLdw = ((g.x * (l.x - c.x)) + (g.y * (l.y - c.y))) + ((g.z * (l.z - c.z)) + (g.w * (l.w - c.w)));
            const float dLde = dLdw * wl / P.sigma;
            if (dLde != 0.f)
            {
                const WinnerSpan& S = BL.S;
                float dPx, dPy, dQx, dQy;
                edge_partials(S.sx[BL.p], S.sy[BL.p], S.sx[BL.q], S.sy[BL.q], (float)y, dPx, dPy, dQx, dQy);
                if (S.vid[BL.p] < P.numVerts)
                {
                    float* d = reinterpret_cast<float*>(P.posGrad + (uint64_t)S.k * P.numVerts + S.vid[BL.p]);
                    unsafeAtomicAdd(d, dLde * dPx);
                    unsafeAtomicAdd(d + 1, dLde * dPy);
                }
                if (S.vid[BL.q] < P.numVerts)
                {
                    float* d = reinterpret_cast<float*>(P.posGrad + (uint64_t)S.k * P.numVerts + S.vid[BL.q]);
                    unsafeAtomicAdd(d, dLde * dQx);
                    unsafeAtomicAdd(d + 1, dLde * dQy);
                }
            }
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
    hipLaunchKernelGGL(k_smooth, dim3(smooth_grid(p)), dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_smooth_backward(const SmoothParams& p, hipStream_t s)
{
    (void)hipGetLastError();   // drop a stale error from an earlier call
    if (p.W <= 0 || p.H <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_smooth_backward, dim3(smooth_grid(p)), dim3(256), 0, s, p);
    return hipGetLastError();
}

}  // namespace slg
