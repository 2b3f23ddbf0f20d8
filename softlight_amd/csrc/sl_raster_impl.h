// sl_raster_impl.h -- k_raster, the fragment stage of the gfx950 draw path,
// and its per-tile-shape launchers.  Included by the sl_raster_*.hip
// translation units, each of which instantiates the launchers of a few
// (colour format, variant) pairs so that the instances compile in parallel.
//
// Reference routines reproduced: SL_TriRasterizer::render_triangle_simd
// generic branch (SL_TriRasterizer.cpp:806-923), render_wireframe (:116-202),
// SL_LineRasterizer (:27-80), SL_PointRasterizer (:27-98),
// SL_FragmentProcessor::flush_tri_fragments (SL_FragmentProcessor.cpp:416-531).
#pragma once

#include <algorithm>

#include <hip/hip_fp16.h>

#include "sl_device.h"
#include "sl_recurrence.h"
#include "sl_stages.h"
#include "sl_kernels.h"

namespace slg {
namespace {

// ---------------------------------------------------------------------------
// k_raster: one wave per tile
// ---------------------------------------------------------------------------
// Catch-up distances up to kStepLimit groups are stepped; longer ones use the
// binade jump (one out-of-line copy for the three barycentric components).
constexpr int kStepLimit = 256;
// Depth rejection (AUX bit 4) runs on tiles (or parts) of at least this many records.
constexpr int kHizMinRecs = 64;

__device__ __noinline__ float3 recur_advance3(float3 s, float3 d, int k)
{
    sl_recur_advance3(s.x, s.y, s.z, d.x, d.y, d.z, k);
    return s;
}

// Greatest value over the wave's 64 lanes (every lane active): pairs and
// quads with DPP quad permutes, the 16-lane row by rotating it 4 and 8 lanes,
// then the four rows' results read from lanes 0, 16, 32 and 48.  NaN inputs
// are ignored (fmaxf) unless every lane holds one.
template <int CTRL>
__device__ __forceinline__ float dpp_f32(float x)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xF, 0xF, false));
}

__device__ __forceinline__ float wave_max_f32(float v)
{
    v = fmaxf(v, dpp_f32<0xB1>(v));    // quad_perm [1,0,3,2]: lane ^ 1
    v = fmaxf(v, dpp_f32<0x4E>(v));    // quad_perm [2,3,0,1]: lane ^ 2
    v = fmaxf(v, dpp_f32<0x124>(v));   // row_ror:4
    v = fmaxf(v, dpp_f32<0x128>(v));   // row_ror:8
    const float r0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
    const float r1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 16));
    const float r2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32));
    const float r3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 48));
    return fmaxf(fmaxf(r0, r1), fmaxf(r2, r3));
}

// Orders one record's LDS updates before the next record's: a wave barrier
// plus compiler fences when the tile is one wave (LDS operations of a wave
// execute in order), a workgroup barrier otherwise.
template <int NT>
__device__ __forceinline__ void record_barrier()
{
    if (NT == 64)
    {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    else
    {
        __syncthreads();
    }
}

// Shading of one fragment from its record and its barycentric values s0..s2
// (perspective correction SL_FragmentProcessor.cpp:486-494, varyings :315-330,
// then the fragment program).  `wv` receives the visibility slot's varying.
// The destination a blend reads from a texel of n < 4 components
// (assign_alpha_pixel, SL_Framebuffer.cpp:296-311): (r, min, min, max),
// (r, g, min, max), and for three components (r, r, r, max).
__device__ __forceinline__ void blend_dst_read(int n, float one, float d[4])
{
    // selects, not conditional stores, so d stays in registers
    const float r = d[0];
    const float g = (n == 3) ? r : ((n == 1) ? 0.f : d[1]);
    const float b = (n == 4) ? d[2] : ((n == 3) ? r : 0.f);
    const float a = (n == 4) ? d[3] : one;
    d[1] = g;
    d[2] = b;
    d[3] = a;
}

// Output k (1..3) of a multiple-output program into colour attachment k at
// texel gi, written or blended as assign_pixel / assign_alpha_pixel do.
// These attachments are read and written in place: a texel is only ever
// touched by the lane that owns it.
__device__ __forceinline__ void write_extra_output(const RasterParams& P, int k, uint64_t gi, float4 o, int blendMode)
{
    const int n = P.colorXComps[k];
    if (P.colorXFloat[k])
    {
        float* px = reinterpret_cast<float*>(P.colorX[k]) + gi * (uint64_t)n;
        float d[4] = {o.x, o.y, o.z, o.w};
        if (blendMode != SL_BLEND_OFF)
        {
#pragma unroll
            for (int c = 0; c < 4; ++c) d[c] = (c < n) ? px[c] : 0.f;
            blend_dst_read(n, 1.f, d);
            blend(blendMode, o, d);
        }
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if (c < n) px[c] = d[c];
    }
    else if (P.colorXPacked[k] != 0)
    {
        const PackedFmt f = packed_format(P.colorXPacked[k]);
        uint8_t* px = P.colorX[k] + gi * (uint64_t)f.bytes;
        uint32_t w = 0;
        for (int b = 0; b < f.bytes; ++b) w |= (uint32_t)px[b] << (8 * b);
        float d[4] = {o.x, o.y, o.z, o.w};
        if (blendMode != SL_BLEND_OFF)
        {
            packed_decode(f, w, d);
            blend(blendMode, o, d);
        }
        w = packed_encode(f, d);
        for (int b = 0; b < f.bytes; ++b) px[b] = (uint8_t)(w >> (8 * b));
    }
    else
    {
        uint8_t* px = P.colorX[k] + gi * (uint64_t)n;
        float d[4] = {o.x, o.y, o.z, o.w};
        if (blendMode != SL_BLEND_OFF)
        {
#pragma unroll
            for (int c = 0; c < 4; ++c) d[c] = (c < n) ? (float)px[c] * 0.00392156862745f : 0.f;
            blend_dst_read(n, 255.f * 0.00392156862745f, d);
            blend(blendMode, o, d);
        }
        const uint32_t v = pack_rgba8(d);
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if (c < n) px[c] = (uint8_t)(v >> (8 * c));
    }
}

// A point's varyings pass unchanged (SL_PointRasterizer.cpp:66-69); a line's
// are mixed by its position s0 (interpolate_line_varyings,
// SL_FragmentProcessor.cpp:119-145).  Outputs 1..3 of a multiple-output
// program go straight to their attachments at texel gi (attachments 3, 2, 1,
// as the reference's output switch falls through); output 0 is returned.
// Only the MRT raster variants carry that code.
template <bool MRT>
__device__ __forceinline__ bool shade(const RasterParams& P, const DrawDesc& D, const float hw[3], float bc3, uint32_t state,
                                      uint32_t rid, float s0, float s1, float s2, float4& out, float4& wv,
                                      uint64_t gi, int blendMode)
{
    const float t0 = s0 * hw[0], t1 = s1 * hw[1], t2 = s2 * hw[2];
    const float sinv = rcpf(((t0 + t1) + t2) + bc3);
    const float c0 = t0 * sinv, c1 = t1 * sinv, c2 = t2 * sinv;
    const float4* V = P.vars + (uint64_t)rid * P.varStride;
    const uint32_t nvMax = P.varStride / 3;
    const bool point = (state & kStatePoint) != 0u, line = (state & kStateLine) != 0u;
    Vary4 var;
#pragma unroll
    for (int i = 0; i < SL_SHADER_MAX_VARYING_VECTORS; ++i)
    {
        if (i < D.nvary)
        {
            const float4 a = V[i];
            if (point)
            {
                var.v[i] = a;
            }
            else if (line)
            {
                var.v[i] = mix4(a, V[nvMax + i], s0);
            }
            else
            {
                const float4 b = V[nvMax + i], e = V[2 * nvMax + i];
                var.v[i] = f4(((a.x * c0) + (b.x * c1)) + (e.x * c2), ((a.y * c0) + (b.y * c1)) + (e.y * c2),
                              ((a.z * c0) + (b.z * c1)) + (e.z * c2), ((a.w * c0) + (b.w * c1)) + (e.w * c2));
            }
        }
        else
        {
            var.v[i] = f4(0.f, 0.f, 0.f, 0.f);
        }
    }
    wv = f4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int i = 0; i < SL_SHADER_MAX_VARYING_VECTORS; ++i)
        if ((uint32_t)i == D.visSlot) wv = var.v[i];
    if (!MRT || D.nout <= 1) return (D.nout > 0) && fragment_program(D, P.uniforms, var, out);
    float4 extra[3];
    const bool have = fragment_program(D, P.uniforms, var, out, extra);
    if (have)
        for (int k = (int)D.nout - 1; k >= 1; --k) write_extra_output(P, k - 1, gi, extra[k - 1], blendMode);
    return have;
}

// Attachment texels of P.colorComps components (8-bit: packed into the low
// bytes of a u32; float: the leading components of a float4), and the
// depth texel (R32F, or R16F read as float and written rounded to nearest).
__device__ __forceinline__ uint32_t load_color8(const RasterParams& P, uint64_t gi)
{
    const int n = P.colorComps;   // bytes per texel for a packed format
    if (n == 4) return reinterpret_cast<const uint32_t*>(P.color)[gi];
    const uint8_t* p = P.color + gi * (uint64_t)n;
    uint32_t v = 0;
    for (int k = 0; k < n; ++k) v |= (uint32_t)p[k] << (8 * k);
    return v;
}

__device__ __forceinline__ void store_color8(const RasterParams& P, uint64_t gi, uint32_t v)
{
    const int n = P.colorComps;
    if (n == 4) { reinterpret_cast<uint32_t*>(P.color)[gi] = v; return; }
    uint8_t* p = P.color + gi * (uint64_t)n;
    for (int k = 0; k < n; ++k) p[k] = (uint8_t)(v >> (8 * k));
}

__device__ __forceinline__ float4 load_colorF(const RasterParams& P, uint64_t gi)
{
    const int n = P.colorComps;
    if (n == 4) return reinterpret_cast<const float4*>(P.color)[gi];
    const float* p = reinterpret_cast<const float*>(P.color) + gi * (uint64_t)n;
    return make_float4(p[0], n > 1 ? p[1] : 0.f, n > 2 ? p[2] : 0.f, 0.f);
}

__device__ __forceinline__ void store_colorF(const RasterParams& P, uint64_t gi, float4 v)
{
    const int n = P.colorComps;
    if (n == 4) { reinterpret_cast<float4*>(P.color)[gi] = v; return; }
    float* p = reinterpret_cast<float*>(P.color) + gi * (uint64_t)n;
    p[0] = v.x;
    if (n > 1) p[1] = v.y;
    if (n > 2) p[2] = v.z;
}

__device__ __forceinline__ float load_depth(const RasterParams& P, uint64_t gi)
{
    return P.depthHalf ? __half2float(reinterpret_cast<const __half*>(P.depth)[gi]) : P.depth[gi];
}

__device__ __forceinline__ void store_depth(const RasterParams& P, uint64_t gi, float z)
{
    if (P.depthHalf) reinterpret_cast<__half*>(P.depth)[gi] = __float2half_rn(z);
    else P.depth[gi] = z;
}

template <int CF>
__device__ __forceinline__ void blend_into(uint32_t* col8, float4* colF, int blendMode, float4 out, int comps, int packed)
{
    if (CF == 1 && packed != 0)
    {
        // a packed texel: decoded whole, blended, clamped, encoded (SL_Framebuffer.cpp:370-560)
        const PackedFmt f = packed_format(packed);
        float o[4] = {out.x, out.y, out.z, out.w};
        if (blendMode != SL_BLEND_OFF)
        {
            packed_decode(f, *col8, o);
            blend(blendMode, out, o);
        }
        *col8 = packed_encode(f, o);
        return;
    }
    if (CF == 1)
    {
        if (blendMode == SL_BLEND_OFF)
        {
            const float o[4] = {out.x, out.y, out.z, out.w};
            *col8 = pack_rgba8(o);
        }
        else
        {
            float dcol[4];
            unpack_rgba8(*col8, dcol);
            blend_dst_read(comps, 255.f * 0.00392156862745f, dcol);   // max (255) cast to float
            blend(blendMode, out, dcol);
            *col8 = pack_rgba8(dcol);
        }
    }
    if (CF == 2)
    {
        if (blendMode == SL_BLEND_OFF)
        {
            *colF = out;
        }
        else
        {
            const float4 cur = *colF;
            float dcol[4] = {cur.x, cur.y, cur.z, cur.w};
            blend_dst_read(comps, 1.f, dcol);
            blend(blendMode, out, dcol);
            *colF = make_float4(dcol[0], dcol[1], dcol[2], dcol[3]);
        }
    }
}

constexpr uint32_t kNoPending = 0xFFFFFFFFu;

// k_raster: one workgroup per screen tile.
//
// Shading is deferred where that cannot change the result: a fragment of an
// unblended draw that passes the depth test only becomes the pixel's pending
// fragment (its record and barycentric values, in LDS); a later passing
// fragment replaces it, and it is shaded once -- when the tile ends, or before
// a blended fragment reads the pixel's colour.  The pending fragment is shaded
// with exactly the arithmetic an immediate shade would use, so colours are
// unchanged, but the dependent global loads of varyings and texels are paid
// once per pixel instead of once per record.
//
// A tile of more than one wave is split by rows: each wave owns TH/NW rows,
// stages and walks the tile's record list on its own (skipping records that
// miss its rows with one uniform test) and never touches another wave's
// pixels, so no workgroup barrier is needed between records.
//
// AUX bit 0: debug buffers (last primitive, depth-pass count); bit 1: the
// visibility attachment; bit 2: the batch holds wireframe, line or point
// records (the variants without it carry only the filled-triangle code);
// bit 3: the batch has draws with several fragment outputs (launched with
// bit 2 set); bit 4: depth rejection of records that fail every depth test
// in a wave (filled triangles, LESS / LESS_EQUAL; results are identical with
// it and without it).
//
// Heavy tiles (RasterParams::items): a workgroup takes a work item, a whole
// tile or one part of a split tile's record list.  MODE 1 walks a part from
// +inf (-inf) and leaves the depth it ends with, the part's own result m; a
// MODE 0 part starts from fold(initial depth, m of the earlier parts) --
// exactly the depth the sequential walk has there -- so it counts the
// fragments that pass exactly and ends with its last passing fragment
// pending; it leaves those per pixel, and the tile's last part to finish
// merges: the tile as it was before the batch, the parts' depths folded in
// order, the last part with a passing fragment giving the one to shade.
template <int CF, int AUX, int TW, int TH, int SEGS, int MODE>
__global__ __launch_bounds__(TH * 4 * SEGS) void k_raster(RasterParams P)
{
    constexpr bool DBG = (AUX & 1) != 0;
    constexpr bool VIS = (AUX & 2) != 0;
    constexpr bool EXTRA = (AUX & 4) != 0;
    constexpr bool MRT = (AUX & 8) != 0;
    constexpr bool HIZ = (AUX & 16) != 0;
    constexpr int NT = TH * 4 * SEGS;       // lanes: rows x segments x 4 phases
    constexpr int NW = NT / 64;             // waves, each owning TH / NW rows
    constexpr int SW = TW / SEGS;           // segment width (columns)
    static_assert(SW >= 4 && SW % 4 == 0, "segment width");
    static_assert(NT % 64 == 0 && TH % NW == 0, "whole waves of whole rows");
    __shared__ float sDepth[TH][TW];
    __shared__ uint32_t sCol8[CF == 1 ? TH : 1][TW];
    __shared__ float4 sColF[CF == 2 ? TH : 1][TW];
    __shared__ uint32_t sLast[DBG ? TH : 1][TW];
    __shared__ uint32_t sCount[DBG ? TH : 1][TW];
    __shared__ float4 sVis[VIS ? TH : 1][TW];
    __shared__ uint32_t sPend[TH][TW];                  // pending fragment's record, or kNoPending
    __shared__ float sPendS[3][TH][TW];                 // its barycentric values

    // a frame whose setup outgrew the bin capacities (a graph replay, whose
    // buffers are fixed at capture) leaves its targets untouched rather than
    // drawing without the records that did not fit; the host reports it
    if (P.ctrl != nullptr && P.ctrl->overflow != 0u) return;

    // the workgroup's work: a tile in launch order (heaviest first; the order
    // only affects scheduling), or a work item
    int tile;
    uint2 range;
    RasterItem item{};
    bool split = false;
    if (P.items != nullptr)
    {
        if (blockIdx.x >= P.itemCount[MODE == 1 ? 1 : 0]) return;
        item = P.items[blockIdx.x];
        if (item.parts == 0u) return;          // a split tile's entry in launch order: its parts do the work
        tile = (int)item.tile;
        range = make_uint2(item.begin, item.end);
        split = item.parts > 1u;
    }
    else
    {
        if (MODE == 1) return;
        tile = (P.tileOrder != nullptr) ? (int)P.tileOrder[blockIdx.x] : (int)blockIdx.x;
        range = P.tileRange[tile];
    }
    const int tx = tile % P.tilesX, ty = tile / P.tilesX;
    const int x0 = tx * TW, y0 = ty * TH;
    const int lane = threadIdx.x;
    const int wave = lane >> 6, wl = lane & 63;
    constexpr int kTilePix = TW * TH;
    const uint64_t slot0 = (uint64_t)(item.slot - item.part);   // first part slot of a split tile

    // tile init: fused clear or load of the tile as it is before the batch
    auto init_tile = [&]() {
        for (int k = lane; k < kTilePix; k += NT)
        {
            const int r = k / TW, c = k % TW;
            const int gx = x0 + c, gy = y0 + r;
            const bool inb = gx < P.W && gy < P.H;
            const uint64_t gi = (uint64_t)gy * P.W + gx;
            sDepth[r][c] = P.clearDepth ? P.clearDepthValue : (inb ? load_depth(P, gi) : 0.f);
            if (CF == 1) sCol8[r][c] = P.clearColor ? P.clearRGBA8 : (inb ? load_color8(P, gi) : 0u);
            if (CF == 2)
                sColF[r][c] = P.clearColor ? make_float4(P.clearRGBAF[0], P.clearRGBAF[1], P.clearRGBAF[2], P.clearRGBAF[3])
                                           : (inb ? load_colorF(P, gi) : make_float4(0.f, 0.f, 0.f, 0.f));
            if (DBG)
            {
                sLast[r][c] = P.clearDepth ? 0u : (inb ? P.dbgLast[gi] : 0u);
                sCount[r][c] = P.clearDepth ? 0u : (inb ? P.dbgCount[gi] : 0u);
            }
            if (VIS) sVis[r][c] = (P.clearDepth || !inb) ? make_float4(0.f, 0.f, 0.f, 0.f) : P.vis[gi];
            sPend[r][c] = kNoPending;
        }
    };
    if (MODE == 1)
    {
        // a part's own result: its fragments depth-tested from the far end
        const float far = (P.splitClass == kSplitMax) ? -__builtin_inff() : __builtin_inff();
        for (int k = lane; k < kTilePix; k += NT) sDepth[k / TW][k % TW] = far;
    }
    else if (split)
    {
        // a part starts from the depth the sequential walk has there: the
        // tile's depth before the batch with the earlier parts' results
        // folded in; its counts start at zero and the merge adds them
        for (int k = lane; k < kTilePix; k += NT)
        {
            const int r = k / TW, c = k % TW;
            const int gx = x0 + c, gy = y0 + r;
            const bool inb = gx < P.W && gy < P.H;
            // the fold pass left fold(initial, m_0 .. m_j) in part j's slot
            float d = P.clearDepth ? P.clearDepthValue : (inb ? load_depth(P, (uint64_t)gy * P.W + gx) : 0.f);
            if (P.splitClass != kSplitConst && item.part > 0u) d = P.partDepth[(slot0 + item.part - 1u) * kTilePix + k];
            sDepth[r][c] = d;
            if (DBG)
            {
                sLast[r][c] = 0u;
                sCount[r][c] = 0u;
            }
            sPend[r][c] = kNoPending;
        }
    }
    else
    {
        init_tile();
    }
    __syncthreads();

    // shades pixel (r, c)'s pending fragment into the tile (caller owns the pixel)
    auto resolve = [&](int r, int c) {
        const uint32_t prid = sPend[r][c];
        if (prid == kNoPending) return;
        sPend[r][c] = kNoPending;
        const SetupRec& Q = P.recs[prid];
        const DrawDesc& QD = P.draws[Q.draw];
        const float qhw[3] = {Q.hw[0], Q.hw[1], Q.hw[2]};
        float4 out, wv;
        const bool have = shade<MRT>(P, QD, qhw, Q.bc3, Q.state, prid, sPendS[0][r][c], sPendS[1][r][c], sPendS[2][r][c], out, wv,
                                (uint64_t)(y0 + r) * P.W + (x0 + c), SL_BLEND_OFF);
        if (have) blend_into<CF>(&sCol8[CF == 1 ? r : 0][CF == 1 ? c : 0], &sColF[CF == 2 ? r : 0][CF == 2 ? c : 0], SL_BLEND_OFF, out, P.colorComps, P.colorPacked);
        if (VIS) sVis[r][c] = make_float4(__uint_as_float(Q.seq), wv.x, wv.y, wv.z);
    };

    // lane = (row, segment, phase): the lane owns the columns x = phase (mod 4)
    // of its row's segment, whichever record it works on
    const int row = lane / (4 * SEGS);
    const int phase = lane & 3;
    const int seg = (lane >> 2) % SEGS;
    const int sx0 = x0 + seg * SW;
    const int sxLast = min(sx0 + SW, P.W) - 1;
    const int y = y0 + row;
    const float yf = (float)y;
    uint32_t tested = 0, shaded = 0;

    // Each wave stages records into LDS kBatch at a time.  The global loads
    // of the next batch (its record ids two batches ahead, its record words
    // one batch ahead) are issued before the current batch is processed, so
    // their latency hides behind the raster work.
    constexpr int kBatch = 16;
    // depth rejection (below) on tiles of at least kHizMinRecs records:
    // lighter tiles reject too little to pay for the bound's reduction
    const bool hiz = HIZ && !P.depthHalf && range.y - range.x >= (uint32_t)kHizMinRecs;
    constexpr int kRecWords = sizeof(SetupRec) / 4;
    constexpr int kLaneWords = kBatch * kRecWords / 64;
    static_assert(kBatch * kRecWords % 64 == 0 && kBatch < 32, "record staging; 32-bit record masks");
    __shared__ uint32_t sRidW[NW][kBatch];
    __shared__ SetupRec sRecW[NW][kBatch];
    uint32_t* sRid = sRidW[wave];
    SetupRec* sRec = sRecW[wave];
    // record id of entry b + wl (lanes below kBatch)
    auto fetch_ids = [&](uint32_t b) -> uint32_t {
        return (wl < kBatch && b + (uint32_t)wl < range.y) ? P.tileRecs[b + wl] : 0u;
    };
    // this lane's words of the batch starting at b (ids in lanes 0..kBatch-1)
    auto fetch_words = [&](uint32_t b, uint32_t ids, uint32_t (&v)[kLaneWords]) {
        const int nb = (int)min((uint32_t)kBatch, range.y - b);
#pragma unroll
        for (int j = 0; j < kLaneWords; ++j)
        {
            const int q = wl + 64 * j;
            const int r = q / kRecWords, w = q - r * kRecWords;
            const uint32_t id = __shfl(ids, r, 64);
            v[j] = (r < nb) ? reinterpret_cast<const uint32_t*>(P.recs + id)[w] : 0u;
        }
    };
    uint32_t idsCur = 0, idsNext = 0;
    uint32_t vCur[kLaneWords] = {};
    if (range.x < range.y)
    {
        idsCur = fetch_ids(range.x);
        fetch_words(range.x, idsCur, vCur);
        idsNext = fetch_ids(range.x + kBatch);
    }
    for (uint32_t base = range.x; base < range.y; base += kBatch)
    {
        const int nb = (int)min((uint32_t)kBatch, range.y - base);
        if (wl < kBatch) sRid[wl] = idsCur;
#pragma unroll
        for (int j = 0; j < kLaneWords; ++j) reinterpret_cast<uint32_t*>(sRec)[wl + 64 * j] = vCur[j];
        record_barrier<64>();
        uint32_t vNext[kLaneWords] = {};
        uint32_t idsAfter = 0;
        if (base + kBatch < range.y)
        {
            fetch_words(base + kBatch, idsNext, vNext);
            idsAfter = fetch_ids(base + 2 * kBatch);
        }
        // The batch's records whose rows and column box reach one of this
        // lane's pixels, as a mask in record order.  Lane r works out record
        // r's mask over the wave's 64 lanes -- the wave's rows it spans times
        // the lanes whose columns (x = phase mod 4 within their segment, left
        // of the image's edge) meet its column box -- and every lane takes
        // its own bit of each record's mask (read from lane r's registers).
        uint64_t hitMask = 0;
        // Depth rejection.  zTop: the greatest depth the wave's pixels hold at
        // the batch's start (float depth only: an R16F write rounds, which can
        // land above a clear value that is not a half).  Depth tests LESS and
        // LESS_EQUAL only lower a pixel's depth, so zTop bounds it for every
        // record of the batch before the first one that can raise it (depth
        // writes with the test off, GREATER(_EQUAL) or NOT_EQUAL).  A filled
        // triangle with a LESS(_EQUAL) test whose computed depth provably
        // exceeds zTop at every pixel of the wave it can reach fails each of
        // its depth tests: it leaves every lane's walk (rejMask), and lane r
        // counts record r's tests over the wave instead.
        float zTop = -INFINITY;
        if (HIZ && hiz)
        {
#pragma unroll
            for (int k = 0; k < SW / 4; ++k) zTop = fmaxf(zTop, sDepth[row][seg * SW + phase + 4 * k]);
            zTop = wave_max_f32(zTop);
        }
        bool rej = false, raise = false;
        uint32_t rejTested = 0;   // depth tests of this lane's record in the wave, if it is rejected
        if (wl < nb)
        {
            const SetupRec& R = sRec[wl];
            constexpr int LPR = 4 * SEGS, RPW = 64 / LPR;   // lanes per row, rows per wave
            const int wy0 = y0 + (wave * 64) / LPR;
            const int r0 = max(R.minY - wy0, 0), r1 = min(min(R.maxY, P.H) - wy0, RPW) - 1;
            uint64_t rowsel = 0;
#pragma unroll
            for (int r = 0; r < RPW; ++r)
                if (r >= r0 && r <= r1) rowsel |= 1ull << (r * LPR);
            const int jlo = max((int)(R.xbox & 0xFFFFu) - x0, 0);
            const int jhi = min(min((int)(R.xbox >> 16) - x0, TW - 1), P.W - 1 - x0);
            const uint64_t cm = (jlo <= jhi) ? ((~0ull >> (63 - jhi)) & (~0ull << jlo)) : 0ull;
            uint64_t g = cm;
#pragma unroll
            for (int k = 1; k < SW / 4; ++k) g |= cm >> (4 * k);
            uint64_t colBits = 0;
#pragma unroll
            for (int sg = 0; sg < SEGS; ++sg) colBits |= ((g >> (sg * SW)) & 0xFull) << (sg * 4);
            hitMask = colBits * rowsel;   // colBits < 2^LPR: the row copies do not overlap
            const uint32_t rst = R.state;
            const int rdt = (int)(rst & 7u);
            raise = HIZ && ((rst >> 3) & 1u) != 0u &&
                    (rdt == SL_DEPTH_TEST_OFF || rdt == SL_DEPTH_TEST_GREATER_THAN ||
                     rdt == SL_DEPTH_TEST_GREATER_EQUAL || rdt == SL_DEPTH_TEST_NOT_EQUAL);
            if (HIZ && hiz && hitMask != 0ull && (rst & (kStateWire | kStateLine | kStatePoint)) == 0u &&
                (rdt == SL_DEPTH_TEST_LESS_THAN || rdt == SL_DEPTH_TEST_LESS_EQUAL))
            {
                // The lane computes s_i = fl(x bx_i + fma(by_i, y, bz_i)), then
                // s_i += 4 bx_i per group up to its pixel (at most X / 4 + 1
                // steps from the span's start, x >= 0), and z = sum z_i s_i.
                // With A_i = X |bx_i| + Y |by_i| + |bz_i| (>= |S_i| and each
                // product term over 0 <= x <= X, 0 <= y <= Y) every rounding
                // of that chain and of the corner values below is at most
                // 2^-24 A_i, so the computed depth of any pixel in the
                // rectangle lies within sum |z_i| A_i 2^-24 (X / 4 + 13) of
                // the plane's exact value, whose minimum is at a corner.
                const float xa = (float)(x0 + jlo), xb = (float)(x0 + jhi);
                const float ya = (float)(wy0 + r0), yb = (float)(wy0 + r1);
                float zlo = INFINITY, mag = 0.f;
                float zc[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int i = 0; i < 3; ++i)
                {
                    const float bxi = R.bx[i], byi = R.by[i], bzi = R.bz[i], zi = R.z[i];
                    const float v = fmaf(byi, ya, bzi), w = fmaf(byi, yb, bzi);
                    zc[0] = fmaf(zi, fmaf(bxi, xa, v), zc[0]);
                    zc[1] = fmaf(zi, fmaf(bxi, xb, v), zc[1]);
                    zc[2] = fmaf(zi, fmaf(bxi, xa, w), zc[2]);
                    zc[3] = fmaf(zi, fmaf(bxi, xb, w), zc[3]);
                    mag = fmaf(fabsf(zi), fmaf(xb, fabsf(bxi), fmaf(yb, fabsf(byi), fabsf(bzi))), mag);
                }
                zlo = fminf(fminf(zc[0], zc[1]), fminf(zc[2], zc[3]));
                // twice the bound above, and NaN / infinite terms never reject
                const float margin = mag * (5.9604645e-8f * (2.f * (0.25f * xb + 13.f))) * 1.0625f;
                rej = (zlo - margin) > zTop;
                if (rej && MODE != 1)
                {
                    // its depth tests in the wave: per row, the span's pixels
                    // in the tile and the image (every lane's share of the
                    // row's columns, as the walk would count them)
                    const int xr = min(x0 + TW, P.W);
                    for (int r = r0; r <= r1; ++r)
                    {
                        int32_t xMin, xMax;
                        scanline_step(R, (float)(wy0 + r), xMin, xMax);
                        if ((uint32_t)xMin < (uint32_t)xMax)
                        {
                            const int a = max((int)xMin, x0), b = min((int)xMax, xr);
                            if (a < b) rejTested += (uint32_t)(b - a);
                        }
                    }
                }
            }
        }
        uint32_t rejMask = 0;
        if constexpr (HIZ)
        {
            // records before the batch's first depth-raising record
            const uint64_t rb = __ballot(raise);
            const uint64_t before = (rb != 0ull) ? ((1ull << (__ffsll((unsigned long long)rb) - 1)) - 1ull) : ~0ull;
            rejMask = (uint32_t)(__ballot(rej) & before);
            if (wl < kBatch && ((rejMask >> wl) & 1u)) tested += rejTested;
        }
        // only the records that reach some lane of the wave (and, with
        // rejection, are left after it: rejected ones are counted above)
        uint32_t cand = 0;
        uint32_t todo = (uint32_t)__ballot(hitMask != 0ull);   // lanes nb.. hold no mask
        if constexpr (HIZ) todo &= ~rejMask;
        for (; todo != 0u; todo &= todo - 1u)
        {
            const int bi = __ffs(todo) - 1;
            const uint32_t mlo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)hitMask, bi);
            const uint32_t mhi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(hitMask >> 32), bi);
            const uint64_t m = ((uint64_t)mhi << 32) | mlo;
            cand |= (uint32_t)((m >> wl) & 1ull) << bi;
        }
        // Each lane walks its own candidates in record order.  A pixel is only
        // ever touched by the lane that owns it, so lanes may be at different
        // records at the same time and every pixel still sees its fragments
        // in submission order.
        while (cand != 0u)
        {
            const int bi = __ffs(cand) - 1;
            cand &= cand - 1u;
            const SetupRec& R = sRec[bi];
            const uint32_t rid = sRid[bi];
            const uint32_t st = R.state;
            const int dtest = (int)(st & 7u);
            const bool dmask = ((st >> 3) & 1u) != 0;
            const int blendMode = (int)((st >> 4) & 7u);
            const bool hasOut = ((st >> 7) & 1u) != 0;
            // deferred: unblended fragments with outputs, and any
            // fragment whose visibility record is wanted
            const bool defer = (hasOut && blendMode == SL_BLEND_OFF) || (VIS && !hasOut);
            // depth test and fragment work of this lane's pixel x of row y;
            // s0..s2 are a triangle's barycentrics, a line's position
            // along it (s0)
            auto fragment = [&](int x, float z, float s0, float s1, float s2) {
                const int c = x - x0;
                if (MODE == 1)
                {
                    if (depth_test(dtest, z, sDepth[row][c])) sDepth[row][c] = z;
                    return;
                }
                ++tested;
                if (depth_test(dtest, z, sDepth[row][c]))
                {
                    if (defer)
                    {
                        sPend[row][c] = rid;
                        sPendS[0][row][c] = s0;
                        sPendS[1][row][c] = s1;
                        sPendS[2][row][c] = s2;
                    }
                    else if (hasOut)
                    {
                        // a blended fragment reads the colour: settle the pixel first
                        resolve(row, c);
                        const DrawDesc& D = P.draws[R.draw];
                        const float hw[3] = {R.hw[0], R.hw[1], R.hw[2]};
                        float4 out, wv;
                        const bool have = shade<MRT>(P, D, hw, R.bc3, st, rid, s0, s1, s2, out, wv,
                                                (uint64_t)y * P.W + x, blendMode);
                        if (have) blend_into<CF>(&sCol8[CF == 1 ? row : 0][CF == 1 ? c : 0],
                                                 &sColF[CF == 2 ? row : 0][CF == 2 ? c : 0], blendMode, out, P.colorComps, P.colorPacked);
                        if (VIS) sVis[row][c] = make_float4(__uint_as_float(R.seq), wv.x, wv.y, wv.z);
                    }
                    if (dmask) sDepth[row][c] = P.depthHalf ? __half2float(__float2half_rn(z)) : z;
                    if (DBG)
                    {
                        sLast[row][c] = R.seq;
                        sCount[row][c] += 1u;
                    }
                    ++shaded;
                }
            };
            if (EXTRA && (st & (kStateLine | kStatePoint)) != 0u)
            {
                // This lane's texels of a point or line: columns x, x + 4,
                // ... up to xe of row y, and for a line the walk's texel
                // (W, y - 1), which is texel 0 of row y, before (pre) or
                // after (post) them in walk order.  One loop, so the
                // fragment work is inlined once.
                const bool isLine = (st & kStateLine) != 0u;
                int x = 1, xe = 0;
                bool pre = false, post = false;
                const LineWalk w = line_walk_of(R);
                if (!isLine)
                {
                    // a point: its one texel, if this lane owns it
                    const int px = (int)R.x[0];
                    if ((int)R.y[0] == y && px >= sx0 && px <= sxLast && (px & 3) == phase) x = xe = px;
                }
                else
                {
                    // a line (SL_LineRasterizer::render_line, SL_LineRasterizer.cpp:
                    // 27-80): this row's steps of the Bresenham walk in closed form
                    const int dyAbs = w.dErr >> 1;
                    int xa = 1, xb = 0;
                    if (!w.steep)
                    {
                        const int n = (y - w.Y1) * w.yErr;
                        if (n >= 0 && n <= dyAbs)
                        {
                            xa = w.X1 + ((n == 0) ? 0 : (w.dx * (2 * n - 1)) / w.dErr + 1);
                            xb = w.X1 + ((n == dyAbs) ? w.dx : (w.dx * (2 * n + 1)) / w.dErr);
                        }
                    }
                    else
                    {
                        const int k = y - w.X1;
                        if (k >= 0 && k <= w.dx) xa = xb = w.Y1 + w.yErr * line_minor_steps(w, k);
                    }
                    if (sx0 == 0 && phase == 0 && y >= 1)
                    {
                        const int ka = w.steep ? (y - 1) - w.X1 : P.W - w.X1;
                        bool hit = false;
                        if (ka >= 0 && ka <= w.dx)
                        {
                            const int m = w.Y1 + w.yErr * line_minor_steps(w, ka);
                            hit = w.steep ? (m == P.W) : (m == y - 1);
                        }
                        const int k0 = w.steep ? y - w.X1 : -w.X1;
                        pre = hit && !(xa <= 0 && 0 <= xb && k0 < ka);
                        post = hit && !pre;
                    }
                    const int xs = max(xa, sx0);
                    x = xs + ((phase - xs) & 3);
                    xe = min(xb, sxLast);
                }
                for (;;)
                {
                    int vx, vy, px;
                    if (pre) { vx = P.W; vy = y - 1; px = 0; pre = false; }
                    else if (x <= xe) { vx = x; vy = y; px = x; x += 4; }
                    else if (post) { vx = P.W; vy = y - 1; px = 0; post = false; }
                    else break;
                    float z = R.z[0], interp = 0.f;
                    if (isLine)
                    {
                        const float fx = (float)vx - R.x[0], fy = (float)vy - R.y[0];
                        interp = sqrtf((fx * fx) + (fy * fy)) * R.bx[0];   // length(p - p0) * dist
                        z = R.z[0] + interp * (R.z[1] - R.z[0]);         // mix(z0, z1, interp)
                    }
                    fragment(px, z, interp, 0.f, 0.f);
                }
                continue;
            }
            int32_t xMin, xMax;
            scanline_step(R, yf, xMin, xMax);
            if ((uint32_t)xMin < (uint32_t)xMax)
            {
                // This lane's pixels: x, x + 4, ... below xe, then those of
                // a second run [xn, xen) (wireframe).  One loop serves both
                // kinds so the fragment work is inlined once.
                const bool wire = EXTRA && (st & kStateWire) != 0u;
                int x, xe, xn = 0, xen = 0;
                float s0, s1, s2, d0, d1, d2;
                if (wire)
                {
                    // wireframe: the run from the left end over the distance
                    // the left end moved since the row above, then the run up
                    // to the right end over the distance it moved; the second
                    // can start inside the first (a pixel visited twice, in
                    // run order).  Barycentrics are evaluated per pixel from
                    // the row constants s = fma(dy, y, c) (d = 0).
                    int32_t xMin1, xMax1;
                    scanline_step(R, yf - 1.f, xMin1, xMax1);
                    const int32_t g0 = wire_gap(xMin, xMin1), g1 = wire_gap(xMax, xMax1);
                    const int64_t firstEnd = (int64_t)xMin + g0;
                    const int hi = (firstEnd < (int64_t)xMax) ? (int)firstEnd : xMax;
                    const int xs = max(xMin, sx0);
                    x = xs + ((phase - xs) & 3);
                    xe = min(min(hi, sx0 + SW), P.W);
                    if (firstEnd < (int64_t)xMax)
                    {
                        const int32_t t = (int32_t)((uint32_t)xMax - (uint32_t)g1 - 1u);
                        const int xs2 = max(max(xMin, t) + 1, sx0);
                        xn = xs2 + ((phase - xs2) & 3);
                        xen = min(min(xMax, sx0 + SW), P.W);
                    }
                    s0 = fmaf(R.by[0], yf, R.bz[0]);
                    s1 = fmaf(R.by[1], yf, R.bz[1]);
                    s2 = fmaf(R.by[2], yf, R.bz[2]);
                    d0 = d1 = d2 = 0.f;
                }
                else
                {
                    const int xs = max(xMin, sx0);
                    xe = min(min(xMax, sx0 + SW), P.W);
                    const int xf = xMin + ((phase - xMin) & 3);   // first span pixel of this lane's column class
                    const int g0 = (xs > xf) ? ((xs - xf + 3) >> 2) : 0;
                    x = xf + 4 * g0;
                    // bc(x) = fl(fl(x*dx) + fma(dy, y, c)), then += 4*dx per group of 4 pixels
                    const float fx = (float)xf;
                    s0 = (fx * R.bx[0]) + fmaf(R.by[0], yf, R.bz[0]);
                    s1 = (fx * R.bx[1]) + fmaf(R.by[1], yf, R.bz[1]);
                    s2 = (fx * R.bx[2]) + fmaf(R.by[2], yf, R.bz[2]);
                    d0 = R.bx[0] * 4.f;
                    d1 = R.bx[1] * 4.f;
                    d2 = R.bx[2] * 4.f;
                    if (x < xe)
                    {
                        // advance to the segment's first group: plain steps when
                        // few, the exact k-step jump when many
                        if (g0 <= kStepLimit)
                        {
                            // three independent chains, four steps per trip
                            int gi = 0;
                            for (; gi + 4 <= g0; gi += 4)
                            {
#pragma unroll
                                for (int u = 0; u < 4; ++u)
                                {
                                    s0 = s0 + d0;
                                    s1 = s1 + d1;
                                    s2 = s2 + d2;
                                }
                            }
                            for (; gi < g0; ++gi)
                            {
                                s0 = s0 + d0;
                                s1 = s1 + d1;
                                s2 = s2 + d2;
                            }
                        }
                        else
                        {
                            const float3 a = recur_advance3(make_float3(s0, s1, s2), make_float3(d0, d1, d2), g0);
                            s0 = a.x;
                            s1 = a.y;
                            s2 = a.z;
                        }
                    }
                }
                const float c0 = s0, c1 = s1, c2 = s2;
                for (;;)
                {
                    for (; x < xe; x += 4)
                    {
                        if (wire)
                        {
                            const float fx = (float)x;
                            s0 = fmaf(R.bx[0], fx, c0);
                            s1 = fmaf(R.bx[1], fx, c1);
                            s2 = fmaf(R.bx[2], fx, c2);
                        }
                        fragment(x, (((R.z[0] * s0) + (R.z[1] * s1)) + (R.z[2] * s2)) + R.bc3, s0, s1, s2);
                        s0 = s0 + d0;
                        s1 = s1 + d1;
                        s2 = s2 + d2;
                    }
                    if (xn >= xen) break;
                    x = xn;
                    xe = xen;
                    xn = xen;
                }
            }
        }
        record_barrier<64>();
        idsCur = idsNext;
        idsNext = idsAfter;
#pragma unroll
        for (int j = 0; j < kLaneWords; ++j) vCur[j] = vNext[j];
    }
    __syncthreads();

    auto add_counters = [&]() {
        unsigned long long t = tested, sh = shaded;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
        {
            t += __shfl_xor(t, off, 64);
            sh += __shfl_xor(sh, off, 64);
        }
        if ((lane & 63) == 0)
        {
            FrameCtrl::CounterSlot& cs = P.ctrl->counters[blockIdx.x % kCounterSlots];
            atomicAdd(&cs.tested, t);
            atomicAdd(&cs.shaded, sh);
        }
        tested = 0;
        shaded = 0;
    };

    if (MODE == 1)
    {
        float* out = P.partDepth + (uint64_t)item.slot * kTilePix;
        for (int k = lane; k < kTilePix; k += NT) out[k] = sDepth[k / TW][k % TW];
        return;
    }
    if (split)
    {
        // this part's results, then the merge by the tile's last part
        PartPixel* out = P.partRes + (uint64_t)item.slot * kTilePix;
        for (int k = lane; k < kTilePix; k += NT)
        {
            const int r = k / TW, c = k % TW;
            PartPixel q;
            q.rid = (sPend[r][c] == kNoPending) ? kNoPart : sPend[r][c];
            q.s0 = sPendS[0][r][c];
            q.s1 = sPendS[1][r][c];
            q.s2 = sPendS[2][r][c];
            q.count = DBG ? sCount[r][c] : 0u;
            q.seq = DBG ? sLast[r][c] : 0u;
            out[k] = q;
        }
        add_counters();
        __shared__ uint32_t sFinished;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        if (lane == 0) sFinished = atomicAdd(&P.partDone[item.done], 1u);
        __syncthreads();
        if (sFinished != item.parts - 1u) return;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        init_tile();
        __syncthreads();
        for (int k = lane; k < kTilePix; k += NT)
        {
            const int r = k / TW, c = k % TW;
            // the depth after every part (the fold pass's last slot)
            if (P.splitClass != kSplitConst) sDepth[r][c] = P.partDepth[(slot0 + item.parts - 1u) * kTilePix + k];
            // the fragment to shade: the last part that left one
            for (uint32_t j = item.parts; j-- > 0u;)
            {
                const PartPixel q = P.partRes[(slot0 + j) * kTilePix + k];
                if (q.rid != kNoPart)
                {
                    sPend[r][c] = q.rid;
                    sPendS[0][r][c] = q.s0;
                    sPendS[1][r][c] = q.s1;
                    sPendS[2][r][c] = q.s2;
                    break;
                }
            }
            if (DBG)
            {
                for (uint32_t j = 0; j < item.parts; ++j)
                {
                    const PartPixel q = P.partRes[(slot0 + j) * kTilePix + k];
                    if (q.count != 0u)
                    {
                        sCount[r][c] += q.count;
                        sLast[r][c] = q.seq;
                    }
                }
            }
        }
        __syncthreads();
    }

    // shade what is still pending, then store the tile
    for (int k = lane; k < TW * TH; k += NT)
    {
        const int r = k / TW, c = k % TW;
        resolve(r, c);
        const int gx = x0 + c, gy = y0 + r;
        if (gx < P.W && gy < P.H)
        {
            const uint64_t gi = (uint64_t)gy * P.W + gx;
            store_depth(P, gi, sDepth[r][c]);
            if (CF == 1) store_color8(P, gi, sCol8[r][c]);
            if (CF == 2) store_colorF(P, gi, sColF[r][c]);
            if (DBG)
            {
                P.dbgLast[gi] = sLast[r][c];
                P.dbgCount[gi] = sCount[r][c];
            }
            if (VIS) P.vis[gi] = sVis[r][c];
        }
    }

    if (!split) add_counters();
}

}  // namespace

// Between the two passes of the heavy-tile split: for every pixel of a split
// tile, part j's slot gets fold(initial depth, m_0 .. m_j), the depth the
// sequential walk has after part j.  One workgroup per split tile (the work
// item of its part 0), one thread per pixel.
template <int TW, int TH>
__global__ __launch_bounds__(256) void k_split_fold(RasterParams P)
{
    if (P.ctrl != nullptr && P.ctrl->overflow != 0u) return;
    if (blockIdx.x >= P.itemCount[1]) return;
    const RasterItem item = P.items[blockIdx.x];
    if (item.part != 0u) return;
    constexpr int kTilePix = TW * TH;
    const int tx = (int)item.tile % P.tilesX, ty = (int)item.tile / P.tilesX;
    for (int k = threadIdx.x; k < kTilePix; k += 256)
    {
        const int gx = tx * TW + k % TW, gy = ty * TH + k / TW;
        const bool inb = gx < P.W && gy < P.H;
        float d = P.clearDepth ? P.clearDepthValue : (inb ? load_depth(P, (uint64_t)gy * P.W + gx) : 0.f);
        float* m = P.partDepth + (uint64_t)item.slot * kTilePix + k;
        for (uint32_t j = 0; j < item.parts; ++j)
        {
            const float mj = m[(uint64_t)j * kTilePix];
            if (depth_test(P.splitDtest, mj, d)) d = mj;
            m[(uint64_t)j * kTilePix] = d;
        }
    }
}

// Dynamic LDS that caps the resident workgroups of a k_raster instance at
// wgPerCU per CU (0: no cap): each workgroup then holds LDS / wgPerCU of the
// CU's LDS (its static arrays plus the rest as unused dynamic LDS), within
// the per-workgroup limit.
struct LdsLimits { size_t perCU, perBlock; };

static LdsLimits lds_limits()
{
    static const LdsLimits L = [] {
        int dev = 0, cu = 0, blk = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess) cu = 0;
        if (hipDeviceGetAttribute(&blk, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) blk = 0;
        return LdsLimits{(size_t)std::max(cu, 0), (size_t)std::max(blk, 0)};
    }();
    return L;
}

template <class Kernel>
static size_t occupancy_lds(Kernel kernel, uint32_t wgPerCU)
{
    if (wgPerCU == 0) return 0;
    hipFuncAttributes a{};
    if (hipFuncGetAttributes(&a, reinterpret_cast<const void*>(kernel)) != hipSuccess) return 0;
    const LdsLimits L = lds_limits();
    const size_t stat = a.sharedSizeBytes, want = std::min(L.perCU / wgPerCU, L.perBlock);
    return want > stat ? want - stat : 0;
}

template <int CF, int AUX, int TW, int TH, int SEGS, int MODE>
static void launch_raster_shape(const RasterParams& p, int grid, hipStream_t s)
{
    static_assert(TW * TH > 0, "tile");
    const auto kernel = k_raster<CF, AUX, TW, TH, SEGS, MODE>;
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(TH * 4 * SEGS), occupancy_lds(kernel, p.wgPerCU), s, p);
    if (MODE == 1) hipLaunchKernelGGL((k_split_fold<TW, TH>), dim3(grid), dim3(256), 0, s, p);
}

// The cases mirror kTileConfigs (checked below).
template <int CF, int AUX, int MODE>
static void launch_raster_mode(const RasterParams& p, int grid, int cfg, hipStream_t s)
{
    switch (cfg)
    {
        case 1: launch_raster_shape<CF, AUX, 32, 8, 2, MODE>(p, grid, s); break;
        case 2: launch_raster_shape<CF, AUX, 16, 16, 1, MODE>(p, grid, s); break;
        case 3: launch_raster_shape<CF, AUX, 16, 8, 2, MODE>(p, grid, s); break;
        case 4: launch_raster_shape<CF, AUX, 16, 8, 4, MODE>(p, grid, s); break;
        case 5: launch_raster_shape<CF, AUX, 16, 16, 4, MODE>(p, grid, s); break;
        case 6: launch_raster_shape<CF, AUX, 32, 16, 4, MODE>(p, grid, s); break;
        case 7: launch_raster_shape<CF, AUX, 32, 32, 2, MODE>(p, grid, s); break;
        case 8: launch_raster_shape<CF, AUX, 32, 8, 4, MODE>(p, grid, s); break;
        case 9: launch_raster_shape<CF, AUX, 32, 8, 8, MODE>(p, grid, s); break;
        case 10: launch_raster_shape<CF, AUX, 8, 8, 2, MODE>(p, grid, s); break;
        case 11:
            // the runtime maps configuration 11 to kLargeTileFallback for the
            // variants whose 64x64 tile does not fit in LDS (effective_tile_config)
            if constexpr (CF != 2 && (AUX & 2) == 0) launch_raster_shape<CF, AUX, 64, 64, 1, MODE>(p, grid, s);
            else launch_raster_shape<CF, AUX, 32, 32, 2, MODE>(p, grid, s);
            break;
        default: launch_raster_shape<CF, AUX, 32, 16, 1, MODE>(p, grid, s); break;
    }
}

template <int CF, int AUX>
void launch_raster_cfg(const RasterParams& p, int grid, int cfg, hipStream_t s)
{
    launch_raster_mode<CF, AUX, 0>(p, grid, cfg, s);
}

// The first pass over split tiles' parts (MODE 1); EXTRA: the batch holds
// wireframe, line or point records.
template <int EXTRA>
void launch_raster_depth_cfg(const RasterParams& p, int grid, int cfg, hipStream_t s)
{
    launch_raster_mode<0, EXTRA ? 4 : 0, 1>(p, grid, cfg, s);
}
static_assert(kTileConfigs[0].tileW == 32 && kTileConfigs[0].tileH == 16 && kTileConfigs[0].segs == 1 &&
              kTileConfigs[1].tileW == 32 && kTileConfigs[1].tileH == 8 && kTileConfigs[1].segs == 2 &&
              kTileConfigs[2].tileW == 16 && kTileConfigs[2].tileH == 16 && kTileConfigs[2].segs == 1 &&
              kTileConfigs[3].tileW == 16 && kTileConfigs[3].tileH == 8 && kTileConfigs[3].segs == 2 &&
              kTileConfigs[4].tileW == 16 && kTileConfigs[4].tileH == 8 && kTileConfigs[4].segs == 4 &&
              kTileConfigs[5].tileW == 16 && kTileConfigs[5].tileH == 16 && kTileConfigs[5].segs == 4 &&
              kTileConfigs[6].tileW == 32 && kTileConfigs[6].tileH == 16 && kTileConfigs[6].segs == 4 &&
              kTileConfigs[7].tileW == 32 && kTileConfigs[7].tileH == 32 && kTileConfigs[7].segs == 2 &&
              kTileConfigs[8].tileW == 32 && kTileConfigs[8].tileH == 8 && kTileConfigs[8].segs == 4 &&
              kTileConfigs[9].tileW == 32 && kTileConfigs[9].tileH == 8 && kTileConfigs[9].segs == 8 &&
              kTileConfigs[10].tileW == 8 && kTileConfigs[10].tileH == 8 && kTileConfigs[10].segs == 2 &&
              kTileConfigs[11].tileW == 64 && kTileConfigs[11].tileH == 64 && kTileConfigs[11].segs == 1 &&
              kLargeTileConfig == 11 && kLargeTileFallback == 7 && kNumTileConfigs == 12,
              "launch table and tile configurations disagree");

}  // namespace slg
