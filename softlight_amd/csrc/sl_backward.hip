// sl_backward.hip -- fixed-visibility adjoint of the draw path.
//
// The forward pass (k_raster with a visibility attachment) leaves, for every
// pixel, the submitted-triangle number of the fragment that won the depth test
// and that fragment's perspective-correct weights b0, b1, b2 of the source
// triangle's vertices.  For an unclipped triangle the weights are bit-equal to
// the interpolation weights the forward used, so the backward rebuilds the
// fragment's varyings -- and with them the texture taps -- exactly as the
// forward computed them, then scatters
//   dL/d colour[v_k] += b_k * (dL/dC (.) texel)          (per channel)
//   dL/d texel[tap_j] += w_j * (dL/dC (.) colour)         (bilinear weight w_j)
// with float atomics.  Within a wave, lanes that hit the same vertex or texel
// are combined first (one atomic per distinct address and channel).
//
// No reference counterpart: the reference renders forward only.
#include "sl_device.h"
#include "sl_stages.h"
#include "sl_backward.h"

namespace slg {

// Last draw whose first sequence number is below seq (draws ordered by seqBase).
__device__ __forceinline__ uint32_t find_grad_draw(const GradDesc* D, uint32_t n, uint32_t seq)
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

// Adds v (c channels) to dst[4*key + ...] for every lane with active set.
// Lanes of the wave that share a key are summed first with a butterfly over
// the wave (non-members contribute zero); the lowest such lane issues the
// atomics.  Every lane of the wave must call this together.
template <int C>
__device__ __forceinline__ void wave_scatter_add(float* dst, uint64_t key, bool active, const float (&v)[C])
{
    const int lane = threadIdx.x & 63;
    unsigned long long remaining = __ballot(active);
    while (remaining)
    {
        const int leader = __ffsll((long long)remaining) - 1;
        const uint64_t k = __shfl(key, leader, 64);
        const bool mine = active && key == k;
        const unsigned long long m = __ballot(mine);
        if (__popcll(m) == 1)
        {
            if (lane == leader)
            {
#pragma unroll
                for (int c = 0; c < C; ++c) unsafeAtomicAdd(dst + 4 * k + c, v[c]);
            }
        }
        else
        {
            float s[C];
#pragma unroll
            for (int c = 0; c < C; ++c) s[c] = mine ? v[c] : 0.f;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1)
            {
#pragma unroll
                for (int c = 0; c < C; ++c) s[c] += __shfl_xor(s[c], off, 64);
            }
            if (lane == leader)
            {
#pragma unroll
                for (int c = 0; c < C; ++c) unsafeAtomicAdd(dst + 4 * k + c, s[c]);
            }
        }
        remaining &= ~m;
    }
}

// One lane per pixel; a workgroup covers a 16 x 16 block (a wave 16 x 4).
__global__ __launch_bounds__(256) void k_backward(BackwardParams P)
{
    const int bxN = (P.W + 15) >> 4;
    const int x = (int)(blockIdx.x % bxN) * 16 + (int)(threadIdx.x & 15);
    const int y = (int)(blockIdx.x / bxN) * 16 + (int)(threadIdx.x >> 4);
    const bool inb = x < P.W && y < P.H;
    const uint64_t p = (uint64_t)y * P.W + x;

    float4 vis = make_float4(0.f, 0.f, 0.f, 0.f);
    if (inb) vis = P.vis[p];
    const uint32_t seq = __float_as_uint(vis.x);
    bool live = inb && seq != 0u;
    const GradDesc* G = nullptr;
    uint64_t vid[3] = {0, 0, 0};
    if (live)
    {
        G = &P.draws[find_grad_draw(P.draws, P.numDraws, seq)];
        const DrawDesc& D = G->d;
        const uint64_t t = (uint64_t)(seq - 1u - D.seqBase);
        live = seq > D.seqBase && t < D.numTris;
        if (live)
        {
            const uint64_t inst = t / D.trisPerInstance;
            const uint64_t ti = t - inst * D.trisPerInstance;
            const uint64_t e = D.elementBegin + 3u * ti;
            vid[0] = e; vid[1] = e + 1; vid[2] = e + 2;
            if (D.indexed)
            {
                vid[0] = fetch_index(D, e);
                vid[1] = fetch_index(D, e + 1);
                vid[2] = fetch_index(D, e + 2);
            }
        }
    }

    const float b[3] = {vis.y, vis.z, vis.w};
    float4 g = make_float4(0.f, 0.f, 0.f, 0.f);
    if (live) g = P.gradImage[p];

    // varyings of the winning fragment, in the forward's operation order
    // (interpolate_tri_varyings, SL_FragmentProcessor.cpp:315-330)
    float4 col = make_float4(0.f, 0.f, 0.f, 0.f), tex = make_float4(1.f, 1.f, 1.f, 1.f);
    BilinearTaps taps{};
    bool textured = false;
    if (live)
    {
        const DrawDesc& D = G->d;
        const bool ct = D.fsProg == SL_PROGRAM_COLOR_TEXTURED;
        float4 a[3], uv[3];
#pragma unroll
        for (int k = 0; k < 3; ++k)
        {
            a[k] = read_attr(D, 1, vid[k], ct ? 0.f : 1.f);
            if (ct) a[k].w = 1.f;    // varying0 = (rgb, 1)
            uv[k] = ct ? read_attr(D, 2, vid[k], 0.f) : f4(0.f, 0.f, 0.f, 0.f);
        }
        col = f4(((a[0].x * b[0]) + (a[1].x * b[1])) + (a[2].x * b[2]), ((a[0].y * b[0]) + (a[1].y * b[1])) + (a[2].y * b[2]),
                 ((a[0].z * b[0]) + (a[1].z * b[1])) + (a[2].z * b[2]), ((a[0].w * b[0]) + (a[1].w * b[1])) + (a[2].w * b[2]));
        if (ct && D.tex != nullptr && D.texType == SL_COLOR_RGBA_FLOAT)
        {
            const float u = ((uv[0].x * b[0]) + (uv[1].x * b[1])) + (uv[2].x * b[2]);
            const float v = ((uv[0].y * b[0]) + (uv[1].y * b[1])) + (uv[2].y * b[2]);
            taps = bilinear_taps(D.texW, D.texH, u, v);
            const auto px = gptr(reinterpret_cast<const float*>(D.tex));
            float o[4];
#pragma unroll
            for (int c = 0; c < 4; ++c)
                o[c] = bilerp(taps, px[4 * taps.i0 + c], px[4 * taps.i1 + c], px[4 * taps.i2 + c], px[4 * taps.i3 + c]);
            tex = f4(o[0], o[1], o[2], o[3]);
            textured = true;
        }
    }

    // dL/d colour varying = g (.) texel; colour attribute k gets b_k of it.
    // COLOR_TEXTURED's alpha varying is the constant 1: no gradient.
    const bool colGrad = live && G->gradCol != nullptr;
    const bool ct = live && G->d.fsProg == SL_PROGRAM_COLOR_TEXTURED;
    const float gc[4] = {g.x * tex.x, g.y * tex.y, g.z * tex.z, ct ? 0.f : g.w * tex.w};
    float* gradCol = colGrad ? G->gradCol : nullptr;
    // lanes of a wave may belong to different draws: the scatter key is
    // per destination buffer, so group by (buffer, index)
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        const bool act = colGrad && vid[k] < G->gradColCount;
        const float v[4] = {b[k] * gc[0], b[k] * gc[1], b[k] * gc[2], b[k] * gc[3]};
        // one pass per distinct destination buffer in the wave
        unsigned long long pending = __ballot(act);
        while (pending)
        {
            const int leader = __ffsll((long long)pending) - 1;
            float* buf = reinterpret_cast<float*>(__shfl(reinterpret_cast<uint64_t>(gradCol), leader, 64));
            const bool same = act && gradCol == buf;
            pending &= ~__ballot(same);
            wave_scatter_add<4>(buf, vid[k], same, v);
        }
    }

    // dL/d texel = w_j * (g (.) colour varying)
    const bool texGrad = textured && G->gradTex != nullptr;
    float* gradTex = texGrad ? G->gradTex : nullptr;
    const float gt[4] = {g.x * col.x, g.y * col.y, g.z * col.z, g.w * col.w};
    const uint32_t ti[4] = {taps.i0, taps.i1, taps.i2, taps.i3};
    const float tw[4] = {taps.omdx * taps.omdy, taps.omdx * taps.dy, taps.dx * taps.omdy, taps.dx * taps.dy};
#pragma unroll
    for (int j = 0; j < 4; ++j)
    {
        const bool act = texGrad && ti[j] < G->gradTexTexels;
        const float v[4] = {tw[j] * gt[0], tw[j] * gt[1], tw[j] * gt[2], tw[j] * gt[3]};
        unsigned long long pending = __ballot(act);
        while (pending)
        {
            const int leader = __ffsll((long long)pending) - 1;
            float* buf = reinterpret_cast<float*>(__shfl(reinterpret_cast<uint64_t>(gradTex), leader, 64));
            const bool same = act && gradTex == buf;
            pending &= ~__ballot(same);
            wave_scatter_add<4>(buf, ti[j], same, v);
        }
    }
}

// grad = scale * 2 * (color - target), n RGBA32F texels
__global__ __launch_bounds__(256) void k_loss_l2_grad(const float4* __restrict__ color, const float4* __restrict__ target,
                                                      float4* __restrict__ grad, uint64_t n, float scale)
{
    const float s2 = scale * 2.f;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    {
        const float4 c = color[i], t = target[i];
        grad[i] = make_float4(s2 * (c.x - t.x), s2 * (c.y - t.y), s2 * (c.z - t.z), s2 * (c.w - t.w));
    }
}

hipError_t launch_backward(const BackwardParams& p, hipStream_t s)
{
    (void)hipGetLastError();   // drop a stale error from an earlier call
    if (p.W <= 0 || p.H <= 0 || p.numDraws == 0) return hipSuccess;
    const uint64_t blocks = (uint64_t)((p.W + 15) / 16) * (uint64_t)((p.H + 15) / 16);
    hipLaunchKernelGGL(k_backward, dim3((unsigned)blocks), dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_loss_l2_grad(const float4* color, const float4* target, float4* grad, uint64_t n, float scale, hipStream_t s)
{
    (void)hipGetLastError();   // drop a stale error from an earlier call
    if (n == 0) return hipSuccess;
    uint64_t g = (n + 255) / 256;
    if (g > 8192) g = 8192;
    hipLaunchKernelGGL(k_loss_l2_grad, dim3((unsigned)g), dim3(256), 0, s, color, target, grad, n, scale);
    return hipGetLastError();
}

}  // namespace slg
