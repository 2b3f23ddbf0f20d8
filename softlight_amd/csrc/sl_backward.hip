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
// with float atomics, pre-aggregated per 16 x 16 pixel block in LDS.
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

// ---------------------------------------------------------------------------
// Workgroup pre-aggregation: each workgroup (16 x 16 pixels) sums its vertex
// and texel contributions in LDS hash tables (open addressing, LDS atomics),
// then adds each distinct entry to HBM once per channel.  A table serves one
// destination buffer (the first one a lane claims); lanes of other draws, and
// keys that find no free slot within kProbes, add to HBM directly.
// ---------------------------------------------------------------------------
constexpr int kBwdTile = 16;
constexpr int kVSlots = 512;          // vertex table (a 16 x 16 tile touches at most 768 vertices)
constexpr int kTSlots = 1024;         // texel table (at most 1024 bilinear taps)
constexpr int kProbes = 16;
constexpr uint32_t kEmptyKey = 0xFFFFFFFFu;

struct GradTable
{
    uint32_t* keys;
    float4* vals;
    int bits;                          // log2 of the slot count
};

// Claims the table for buf (or finds it claimed for buf).
__device__ __forceinline__ bool claim(unsigned long long* owner, const float* buf)
{
    const unsigned long long mine = reinterpret_cast<unsigned long long>(buf);
    const unsigned long long prev = atomicCAS(owner, 0ull, mine);
    return prev == 0ull || prev == mine;
}

template <int C>
__device__ __forceinline__ void add_global(float* dst, uint32_t key, const float (&v)[C])
{
#pragma unroll
    for (int c = 0; c < C; ++c)
        if (v[c] != 0.f) unsafeAtomicAdd(dst + 4ull * key + c, v[c]);
}

template <int C>
__device__ __forceinline__ bool table_add(const GradTable& t, uint32_t key, const float (&v)[C])
{
    uint32_t h = (key * 2654435761u) >> (32 - t.bits);
    const uint32_t mask = (1u << t.bits) - 1u;
    for (int probe = 0; probe < kProbes; ++probe)
    {
        const uint32_t prev = atomicCAS(&t.keys[h], kEmptyKey, key);
        if (prev == kEmptyKey || prev == key)
        {
            float* d = reinterpret_cast<float*>(&t.vals[h]);
#pragma unroll
            for (int c = 0; c < C; ++c)
                if (v[c] != 0.f) atomicAdd(d + c, v[c]);
            return true;
        }
        h = (h + 1u) & mask;
    }
    return false;
}

template <int C>
__device__ __forceinline__ void scatter(const GradTable& t, bool useTable, float* buf, uint32_t key, const float (&v)[C])
{
    if (useTable && table_add<C>(t, key, v)) return;
    add_global<C>(buf, key, v);
}

// One lane per pixel; a workgroup covers a 16 x 16 block (a wave 16 x 4).
__global__ __launch_bounds__(256) void k_backward(BackwardParams P)
{
    __shared__ uint32_t sVKey[kVSlots];
    __shared__ float4 sVVal[kVSlots];
    __shared__ uint32_t sTKey[kTSlots];
    __shared__ float4 sTVal[kTSlots];
    __shared__ unsigned long long sVOwner, sTOwner;
    const int tid = threadIdx.x;
    for (int i = tid; i < kVSlots; i += 256) { sVKey[i] = kEmptyKey; sVVal[i] = make_float4(0.f, 0.f, 0.f, 0.f); }
    for (int i = tid; i < kTSlots; i += 256) { sTKey[i] = kEmptyKey; sTVal[i] = make_float4(0.f, 0.f, 0.f, 0.f); }
    if (tid == 0) { sVOwner = 0ull; sTOwner = 0ull; }
    __syncthreads();

    const int bxN = (P.W + kBwdTile - 1) / kBwdTile;
    const int x = (int)(blockIdx.x % bxN) * kBwdTile + (tid & (kBwdTile - 1));
    const int y = (int)(blockIdx.x / bxN) * kBwdTile + (tid / kBwdTile);
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
    const bool ct = live && G->d.fsProg == SL_PROGRAM_COLOR_TEXTURED;
    if (live)
    {
        const DrawDesc& D = G->d;
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
            if (D.nearest)
            {
                // one tap of weight 1 (SL_Sampler.hpp:122-148)
                taps = nearest_taps(D.texW, D.texH, u, v);
                tex = sample_nearest_rgbaf(D.tex, D.texW, D.texH, u, v);
            }
            else
            {
                taps = bilinear_taps(D.texW, D.texH, u, v);
                const auto px = gptr(reinterpret_cast<const float*>(D.tex));
                float o[4];
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    o[c] = bilerp(taps, px[4 * taps.i0 + c], px[4 * taps.i1 + c], px[4 * taps.i2 + c], px[4 * taps.i3 + c]);
                tex = f4(o[0], o[1], o[2], o[3]);
            }
            textured = true;
        }
    }

    // dL/d colour varying = g (.) texel; colour attribute k gets b_k of it.
    // COLOR_TEXTURED's alpha varying is the constant 1: no gradient there.
    if (live && G->gradCol != nullptr)
    {
        const GradTable tv{sVKey, sVVal, 9};
        const bool useTable = claim(&sVOwner, G->gradCol);
        const float gc[4] = {g.x * tex.x, g.y * tex.y, g.z * tex.z, ct ? 0.f : g.w * tex.w};
#pragma unroll
        for (int k = 0; k < 3; ++k)
        {
            if (vid[k] >= G->gradColCount) continue;
            const float v[4] = {b[k] * gc[0], b[k] * gc[1], b[k] * gc[2], b[k] * gc[3]};
            scatter<4>(tv, useTable, G->gradCol, (uint32_t)vid[k], v);
        }
    }

    // dL/d texel = w_j * (g (.) colour varying)
    if (textured && G->gradTex != nullptr)
    {
        const GradTable tt{sTKey, sTVal, 10};
        const bool useTable = claim(&sTOwner, G->gradTex);
        const float gt[4] = {g.x * col.x, g.y * col.y, g.z * col.z, g.w * col.w};
        const uint32_t ti[4] = {taps.i0, taps.i1, taps.i2, taps.i3};
        const float tw[4] = {taps.omdx * taps.omdy, taps.omdx * taps.dy, taps.dx * taps.omdy, taps.dx * taps.dy};
        const int ntaps = G->d.nearest ? 1 : 4;
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            if (j >= ntaps) break;
            if (ti[j] >= G->gradTexTexels) continue;
            const float v[4] = {tw[j] * gt[0], tw[j] * gt[1], tw[j] * gt[2], tw[j] * gt[3]};
            scatter<4>(tt, useTable, G->gradTex, ti[j], v);
        }
    }
    __syncthreads();

    // flush the tables: one HBM atomic per distinct entry and non-zero channel
    float* vbuf = reinterpret_cast<float*>(sVOwner);
    float* tbuf = reinterpret_cast<float*>(sTOwner);
    if (vbuf != nullptr)
    {
        for (int i = tid; i < kVSlots; i += 256)
        {
            const uint32_t key = sVKey[i];
            if (key == kEmptyKey) continue;
            const float4 a = sVVal[i];
            const float v[4] = {a.x, a.y, a.z, a.w};
            add_global<4>(vbuf, key, v);
        }
    }
    if (tbuf != nullptr)
    {
        for (int i = tid; i < kTSlots; i += 256)
        {
            const uint32_t key = sTKey[i];
            if (key == kEmptyKey) continue;
            const float4 a = sTVal[i];
            const float v[4] = {a.x, a.y, a.z, a.w};
            add_global<4>(tbuf, key, v);
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
    const uint64_t blocks = (uint64_t)((p.W + kBwdTile - 1) / kBwdTile) * (uint64_t)((p.H + kBwdTile - 1) / kBwdTile);
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
