// sl_stages.h -- per-primitive and per-fragment stage functions of the gfx950
// draw path (vertex programs, clipping, triangle setup, scanline spans,
// samplers, fragment programs, output merger).
//
// Each function names the reference routine whose results it reproduces; the
// arithmetic order follows the reference's portable scalar branch, with
// rcp = IEEE 1/x and fmadd = fused, and everything else unfused.
#pragma once

#include "sl_device.h"

namespace slg {

struct Vary4 { float4 v[SL_SHADER_MAX_VARYING_VECTORS]; };

// Pointers read from the draw table point to device global memory; casting
// them to the global address space lets loads through them use global (not
// flat) instructions.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* gptr(const T* p)
{
    return (const __attribute__((address_space(1))) T*)p;
}

__device__ __forceinline__ float4 f4(float a, float b, float c, float d) { return make_float4(a, b, c, d); }

// ls::math::mat4 * vec4, column-major: ((c0*x + c1*y) + c2*z) + c3*w
__device__ __forceinline__ float4 mat_mul(const float* m, float4 p)
{
    float r[4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
    {
        const float a = m[0 * 4 + i] * p.x;
        const float b = m[1 * 4 + i] * p.y;
        const float c = m[2 * 4 + i] * p.z;
        const float d = m[3 * 4 + i] * p.w;
        r[i] = ((a + b) + c) + d;
    }
    return f4(r[0], r[1], r[2], r[3]);
}

__device__ __forceinline__ float dot4(float4 a, float4 b)
{
    return (((a.x * b.x) + (a.y * b.y)) + (a.z * b.z)) + (a.w * b.w);
}

__device__ __forceinline__ float4 mix4(float4 a, float4 b, float t)
{
    return f4(a.x + t * (b.x - a.x), a.y + t * (b.y - a.y), a.z + t * (b.z - a.z), a.w + t * (b.w - a.w));
}

__device__ __forceinline__ float comp(float4 v, int i)
{
    return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w));
}

// ---------------------------------------------------------------------------
// Vertex fetch: SL_VertexArray::offset(bind, vert) = offset + stride*vert
// (SL_VertexArray.hpp:120).  Reads past the buffer end return zeros.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float4 read_attr(const DrawDesc& d, int bind, uint64_t vertId, float defW)
{
    float o[4] = {0.f, 0.f, 0.f, defW};
    if ((uint32_t)bind >= d.numBindings || d.vbo == nullptr) return f4(o[0], o[1], o[2], o[3]);
    const BindingDesc& b = d.bind[bind];
    const int64_t off = b.offset + b.stride * (int64_t)vertId;
    const int n = b.dims + 1;
    if (b.type == VERTEX_DATA_FLOAT)
    {
        if (off >= 0 && (uint64_t)off + 4u * (uint64_t)n <= d.vboBytes)
        {
            if ((off & 3) == 0)
            {
                const auto p = gptr(reinterpret_cast<const float*>(d.vbo + off));
                for (int i = 0; i < n; ++i) o[i] = p[i];
            }
            else
            {
                const auto p = gptr(d.vbo + off);
                for (int i = 0; i < n; ++i)
                {
                    const uint32_t b = (uint32_t)p[4 * i] | ((uint32_t)p[4 * i + 1] << 8) | ((uint32_t)p[4 * i + 2] << 16) |
                                       ((uint32_t)p[4 * i + 3] << 24);
                    o[i] = __uint_as_float(b);
                }
            }
        }
    }
    else if (b.type == VERTEX_DATA_BYTE)
    {
        if (off >= 0 && (uint64_t)off + (uint64_t)n <= d.vboBytes)
        {
            const auto p = gptr(d.vbo + off);
            for (int i = 0; i < n; ++i) o[i] = (float)p[i] * 0.00392156862745f;   // color_cast<float,uint8_t>, SL_Color.hpp:666-669
        }
    }
    return f4(o[0], o[1], o[2], o[3]);
}

__device__ __forceinline__ void model_matrix(const DrawDesc& d, const sl_std_uniforms* u, uint64_t inst, float m[16])
{
    if (d.instVbo != nullptr)
    {
        const uint64_t off = inst * 64u;
        if (off + 64u <= d.instVboBytes)
        {
            const auto p = gptr(reinterpret_cast<const float*>(d.instVbo + off));
#pragma unroll
            for (int i = 0; i < 16; ++i) m[i] = p[i];
        }
        else
        {
#pragma unroll
            for (int i = 0; i < 16; ++i) m[i] = 0.f;
        }
        return;
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) m[i] = u->model[i];
}

// Index fetch (SL_TriProcessor.cpp:125-234): u8/u16/u32 indices; an index
// past the buffer reads 0.
__device__ __forceinline__ uint64_t fetch_index(const DrawDesc& d, uint64_t e)
{
    if (e >= d.iboCount) return 0;
    switch (d.iboType)
    {
        case VERTEX_DATA_BYTE: return gptr(d.ibo)[e];
        case VERTEX_DATA_SHORT: return gptr(reinterpret_cast<const uint16_t*>(d.ibo))[e];
        default: return gptr(reinterpret_cast<const uint32_t*>(d.ibo))[e];
    }
}

// The built-in vertex programs (replace SL_VertexShader::shader,
// SL_Shader.hpp:58-64), split into the clip-space position and the varyings.
// The programs are pure, so a triangle can be culled and binned from its
// positions before its varyings are computed; each part performs the same
// operations in the same order as the whole program would.  Both work on a
// triangle's three vertices at once and issue every fetch before any
// arithmetic, so the fetch latencies overlap.
//
//   PASSTHROUGH_GOURAUD  position = attribute 0 (w defaults to 1), varying 0 =
//                        attribute 1 (w defaults to 1)
//   LAMBERT_TEXTURED     position = mvp * (model * (p.xyz, 1)); varyings
//                        (uv, 0, 0) and (model * (n.xyz, 0)).xyz
//   COLOR_TEXTURED       position as above; varyings (c.xyz, 1) and (uv, 0, 0)
//
// model = the instance's mat4 when an instance buffer is bound, else the
// uniform model matrix.  need[k] == false skips vertex k.
__device__ __forceinline__ void vertex_positions3(const DrawDesc& d, const uint8_t* uniArena, const uint32_t vid[3],
                                                  uint64_t inst, const bool need[3], float4 out[3])
{
    if (d.vsProg == SL_PROGRAM_PASSTHROUGH_GOURAUD)
    {
#pragma unroll
        for (int k = 0; k < 3; ++k)
            if (need[k]) out[k] = read_attr(d, 0, vid[k], 1.f);
        return;
    }
    if (d.vsProg == SL_PROGRAM_LAMBERT_TEXTURED || d.vsProg == SL_PROGRAM_LAMBERT_MRT || d.vsProg == SL_PROGRAM_COLOR_TEXTURED)
    {
        const sl_std_uniforms* u = reinterpret_cast<const sl_std_uniforms*>(uniArena + d.uniOffset);
        float4 p[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) p[k] = need[k] ? read_attr(d, 0, vid[k], 1.f) : f4(0.f, 0.f, 0.f, 1.f);
        float m[16];
        model_matrix(d, u, inst, m);
#pragma unroll
        for (int k = 0; k < 3; ++k)
        {
            if (!need[k]) continue;
            const float4 world = mat_mul(m, f4(p[k].x, p[k].y, p[k].z, 1.f));
            out[k] = mat_mul(u->mvp, world);
        }
        return;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k)
        if (need[k]) out[k] = f4(0.f, 0.f, 0.f, 0.f);
}

__device__ __forceinline__ void vertex_varyings3(const DrawDesc& d, const uint8_t* uniArena, const uint32_t vid[3],
                                                 uint64_t inst, Vary4 out[3])
{
    // Every program's results are gathered in registers and stored once, at
    // the same places whatever the program: stores to different varyings on
    // different branches get merged into one store at a computed offset,
    // which moves the whole triangle into scratch memory.
    const float4 zero = f4(0.f, 0.f, 0.f, 0.f);
    float4 r0[3] = {zero, zero, zero}, r1[3] = {zero, zero, zero}, r2[3] = {zero, zero, zero};
    if (d.vsProg == SL_PROGRAM_PASSTHROUGH_GOURAUD)
    {
#pragma unroll
        for (int k = 0; k < 3; ++k) r0[k] = read_attr(d, 1, vid[k], 1.f);
    }
    else if (d.vsProg == SL_PROGRAM_LAMBERT_TEXTURED || d.vsProg == SL_PROGRAM_LAMBERT_MRT || d.vsProg == SL_PROGRAM_COLOR_TEXTURED)
    {
        float4 a[3], uv[3];
#pragma unroll
        for (int k = 0; k < 3; ++k)
        {
            a[k] = read_attr(d, 1, vid[k], 0.f);
            uv[k] = read_attr(d, 2, vid[k], 0.f);
        }
        if (d.vsProg != SL_PROGRAM_COLOR_TEXTURED)
        {
            const sl_std_uniforms* u = reinterpret_cast<const sl_std_uniforms*>(uniArena + d.uniOffset);
            float m[16];
            model_matrix(d, u, inst, m);
            const bool mrt = d.vsProg == SL_PROGRAM_LAMBERT_MRT;
#pragma unroll
            for (int k = 0; k < 3; ++k)
            {
                const float4 n = mat_mul(m, f4(a[k].x, a[k].y, a[k].z, 0.f));
                r0[k] = f4(uv[k].x, uv[k].y, 0.f, 0.f);
                r1[k] = f4(n.x, n.y, n.z, 0.f);
                if (mrt)
                {
                    const float4 p = read_attr(d, 0, vid[k], 1.f);
                    const float4 world = mat_mul(m, f4(p.x, p.y, p.z, 1.f));
                    r2[k] = f4(world.x, world.y, world.z, 1.f);
                }
            }
        }
        else
        {
#pragma unroll
            for (int k = 0; k < 3; ++k)
            {
                r0[k] = f4(a[k].x, a[k].y, a[k].z, 1.f);
                r1[k] = f4(uv[k].x, uv[k].y, 0.f, 0.f);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        out[k].v[0] = r0[k];
        out[k].v[1] = r1[k];
        out[k].v[2] = r2[k];
#pragma unroll
        for (int i = 3; i < SL_SHADER_MAX_VARYING_VECTORS; ++i) out[k].v[i] = zero;
    }
}

// ---------------------------------------------------------------------------
// Vertex stage helpers (SL_TriProcessor.cpp, SL_PostVertexTransform.hpp)
// ---------------------------------------------------------------------------
// face_determinant, generic branch (SL_TriProcessor.cpp:101-116)
__device__ __forceinline__ float face_determinant(float4 p0, float4 p1, float4 p2)
{
    const float w0 = rcpf(p0.w), w1 = rcpf(p1.w), w2 = rcpf(p2.w);
    const float x0 = p0.x * w0, x1 = p1.x * w1, x2 = p2.x * w2;
    const float y0 = p0.y * w0, y1 = p1.y * w1, y2 = p2.y * w2;
    const float4 xABC = f4(x0, x1, x2, 0.f);
    const float4 dy = f4(y1 - y2, y2 - y0, y0 - y1, 0.f - 0.f);
    return dot4(xABC, dy);
}

// sl_ndc_clip_status, generic branch (SL_PostVertexTransform.hpp:653-673)
__device__ __forceinline__ bool clip_inside(float4 c)
{
    const float w = c.w, nw = -c.w;
    return (c.x <= w) && (c.y <= w) && (c.z <= w) && (c.w <= w) &&
           (c.x >= nw) && (c.y >= nw) && (c.z >= nw) && (c.w >= nw);
}

__device__ __forceinline__ int clip_status(float4 c0, float4 c1, float4 c2)
{
    const int vis = (clip_inside(c0) && clip_inside(c1) && clip_inside(c2)) ? 3 : 0;
    const int part = (c0.w > 0.f || c1.w > 0.f || c2.w > 0.f) ? 1 : 0;
    return vis | part;
}

__device__ __forceinline__ float4 clip_edge(int e)
{
    switch (e)
    {
        case 0: return f4(1.f, 0.f, 0.f, 1.f);
        case 1: return f4(-1.f, 0.f, 0.f, 1.f);
        case 2: return f4(0.f, 1.f, 0.f, 1.f);
        case 3: return f4(0.f, -1.f, 0.f, 1.f);
        case 4: return f4(0.f, 0.f, 1.f, 1.f);
        default: return f4(0.f, 0.f, -1.f, 1.f);
    }
}

// Exact shortcut of clip_and_process_tris for the common case: planes are
// visited in the reference's order; a plane that keeps all three vertices
// leaves the triangle unchanged, the first plane that rejects all three ends
// it with no output.  Returns 0 = rejected, 1 = needs the full clipper.
__device__ __forceinline__ int clip_quick(const float4* p)
{
    for (int e = 0; e < 6; ++e)
    {
        const float4 E = clip_edge(e);
        const bool v0 = !signbit(dot4(p[0], E));
        const bool v1 = !signbit(dot4(p[1], E));
        const bool v2 = !signbit(dot4(p[2], E));
        if (v0 && v1 && v2) continue;
        if (!v0 && !v1 && !v2) return 0;
        return 1;
    }
    return 1;
}

// Polygon buffers of the clipper: entry k of a buffer lives at p[k * stride],
// so a thread's polygon can sit in LDS (interleaved with other threads'
// polygons, or in a wave's shared slot) instead of in scratch memory, whose
// latency every step of the clipper would otherwise pay.
template <class T>
struct PolyBuf
{
    T* p;
    int stride;
    __device__ __forceinline__ T& operator[](int k) const { return p[k * stride]; }
};

// SL_TriProcessor::clip_and_process_tris (SL_TriProcessor.cpp:370-499) on
// positions and (kVary) varyings.  Returns the polygon size (0 if it was
// clipped away or has fewer than 3 vertices); the polygon is left in
// outPos / outVar in clip space.  tpos / tvar are work buffers.  All buffers
// hold 9 entries.
template <bool kVary>
__device__ int clip_polygon(const float4* inPos, const Vary4* inVar, int nv, PolyBuf<float4> outPos, PolyBuf<Vary4> outVar,
                            PolyBuf<float4> tpos, PolyBuf<Vary4> tvar)
{
    int n = 3;
    for (int k = 0; k < 3; ++k)
    {
        outPos[k] = inPos[k];
        if constexpr (kVary) outVar[k] = inVar[k];
    }
    for (int e = 0; e < 6; ++e)
    {
        const float4 E = clip_edge(e);
        // a plane that keeps every vertex leaves the polygon as it is
        bool allIn = true;
        for (int k = 0; k < n; ++k) allIn = allIn && !signbit(dot4(outPos[k], E));
        if (allIn) continue;
        int nn = 0;
        int j = n - 1;
        float4 q0 = outPos[n - 1];
        float t0 = dot4(q0, E);
        bool vis0 = !signbit(t0);
        for (int k = 0; k < n; ++k)
        {
            const float4 q1 = outPos[k];
            const float t1 = dot4(q1, E);
            const bool vis1 = !signbit(t1);
            if (vis0 != vis1)
            {
                const float t = t0 * rcpf(t0 - t1);
                tpos[nn] = mix4(q0, q1, t);
                if constexpr (kVary)
                {
#pragma unroll
                    for (int i = 0; i < SL_SHADER_MAX_VARYING_VECTORS; ++i)
                        if (i < nv) tvar[nn].v[i] = mix4(outVar[j].v[i], outVar[k].v[i], t);
                }
                ++nn;
            }
            if (vis1)
            {
                tpos[nn] = q1;
                if constexpr (kVary) tvar[nn] = outVar[k];
                ++nn;
            }
            j = k; q0 = q1; t0 = t1; vis0 = vis1;
        }
        if (nn == 0) return 0;
        n = nn;
        for (int k = 0; k < n; ++k)
        {
            outPos[k] = tpos[k];
            if constexpr (kVary) outVar[k] = tvar[k];
        }
    }
    return n < 3 ? 0 : n;
}

// sl_perspective_divide (:239-249) + sl_ndc_to_screen_coords (:508-524)
__device__ __forceinline__ float4 divide_and_snap(float4 p, const float vp[4])
{
    const float wInv = rcpf(p.w);
    const float x = p.x * wInv, y = p.y * wInv, z = p.z * wInv;
    const float hw = vp[2] * 0.5f, hh = vp[3] * 0.5f;
    return f4(max_ref(floorf(fmaf(x + 1.f, hw, vp[0])), 0.f),
              max_ref(floorf(fmaf(y + 1.f, hh, vp[1])), 0.f), z, wInv);
}

// Bounding-box reject of SL_TriProcessor::push_bin (:256-264) plus the
// non-finite drop; returns false when the triangle produces no bin.
__device__ __forceinline__ bool bin_accept(float4 p0, float4 p1, float4 p2)
{
    if (!isfinite(p0.x) || !isfinite(p1.x) || !isfinite(p2.x) ||
        !isfinite(p0.y) || !isfinite(p1.y) || !isfinite(p2.y)) return false;
    const float minX = fminf(fminf(p0.x, p1.x), p2.x), maxX = fmaxf(fmaxf(p0.x, p1.x), p2.x);
    const float minY = fminf(fminf(p0.y, p1.y), p2.y), maxY = fmaxf(fmaxf(p0.y, p1.y), p2.y);
    return !((maxX - minX < 1.f) || (maxY - minY < 1.f));
}

// Tiles a binned triangle can cover: rows [minY, maxY), columns [minX-1, maxX-1]
// (a truncated span end may fall one pixel left of the snapped bounding box).
struct TileRect { int tx0, tx1, ty0, ty1; };

__device__ __forceinline__ TileRect tile_rect(float4 p0, float4 p1, float4 p2, int W, int H, int tilesX, int tilesY,
                                              int shx, int shy)
{
    const int minX = (int)fminf(fminf(p0.x, p1.x), p2.x);
    const int maxX = (int)fmaxf(fmaxf(p0.x, p1.x), p2.x);
    const int minY = (int)fminf(fminf(p0.y, p1.y), p2.y);
    const int maxY = (int)fmaxf(fmaxf(p0.y, p1.y), p2.y);
    TileRect r;
    const int x0 = max(minX - 1, 0), x1 = min(maxX, W - 1);
    const int y0 = max(minY, 0), y1 = min(maxY - 1, H - 1);
    if (x0 > x1 || y0 > y1) { r.tx0 = 0; r.tx1 = -1; r.ty0 = 0; r.ty1 = -1; return r; }
    r.tx0 = x0 >> shx; r.tx1 = min(x1 >> shx, tilesX - 1);
    r.ty0 = y0 >> shy; r.ty1 = min(y1 >> shy, tilesY - 1);
    return r;
}

__device__ __forceinline__ uint32_t tile_count(const TileRect& r)
{
    return (r.tx1 < r.tx0 || r.ty1 < r.ty0) ? 0u : (uint32_t)((r.tx1 - r.tx0 + 1) * (r.ty1 - r.ty0 + 1));
}

// push_bin's plane setup (SL_TriProcessor.cpp:271-314) and the scanline
// constants of SL_ScanlineBounds::init, generic branch (SL_ScanlineBounds.hpp:143-166).
__device__ __forceinline__ void setup_record(float4 p0, float4 p1, float4 p2, uint32_t draw, uint32_t seq, SetupRec& r)
{
    const float zyx = p2.x - p1.x, zyy = p2.y - p1.y;
    const float xyx = p0.x - p1.x, xyy = p0.y - p1.y;
    const float denom = rcpf((zyx * xyy) - (zyy * xyx));
    const float X0 = p0.x, X1 = p1.x, X2 = p2.x, Y0 = p0.y, Y1 = p1.y, Y2 = p2.y;
    const float ddx[3] = {(Y1 * 1.f) - (Y2 * 1.f), (Y2 * 1.f) - (Y0 * 1.f), (Y0 * 1.f) - (Y1 * 1.f)};
    const float ddy[3] = {(1.f * X2) - (1.f * X1), (1.f * X0) - (1.f * X2), (1.f * X1) - (1.f * X0)};
    const float ddz[3] = {(X1 * Y2) - (X2 * Y1), (X2 * Y0) - (X0 * Y2), (X0 * Y1) - (X1 * Y0)};
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        r.bx[k] = ddx[k] * denom;
        r.by[k] = ddy[k] * denom;
        r.bz[k] = ddz[k] * denom;
    }
    // 4th barycentric component: planes (0*denom), so bc[3] is a signed zero
    // (or NaN when denom is not finite); the depth and perspective sums add
    // it times 0.
    {
        const float b3x = 0.f * denom, b3y = 0.f * denom, b3z = 0.f * denom;
        // value at an arbitrary covered pixel: sign/NaN only depend on the planes
        const float bcY = fmaf(b3y, 1.f, b3z);
        const float bc = (1.f * b3x) + bcY;
        r.bc3 = 0.f * bc;
    }
    r.x[0] = p0.x; r.x[1] = p1.x; r.x[2] = p2.x;
    r.y[0] = p0.y; r.y[1] = p1.y; r.y[2] = p2.y;
    r.z[0] = p0.z; r.z[1] = p1.z; r.z[2] = p2.z;
    r.hw[0] = p0.w; r.hw[1] = p1.w; r.hw[2] = p2.w;
    r.minY = f2i(fminf(fminf(p0.y, p1.y), p2.y));
    r.maxY = f2i(fmaxf(fmaxf(p0.y, p1.y), p2.y));
    r.draw = draw;
    r.seq = seq;
    r.state = 0;
    // Vertices are snapped to whole non-negative pixels, so every span end
    // (a truncated point of an edge between two vertices) lies in
    // [min x - 1, max x]: the span's pixels lie in [min x - 1, max x - 1].
    {
        const int xa = f2i(fminf(fminf(p0.x, p1.x), p2.x)) - 1, xb = f2i(fmaxf(fmaxf(p0.x, p1.x), p2.x)) - 1;
        r.xbox = (uint32_t)min(max(xa, 0), 65535) | ((uint32_t)min(max(xb, 0), 65535) << 16);
    }

    // SL_ScanlineBounds::init: sort by y descending with the reference's swaps
    float ax = p0.x, ay = p0.y, bxx = p1.x, byy = p1.y, cx = p2.x, cy = p2.y, t;
    if (ay < byy) { t = ax; ax = bxx; bxx = t; t = ay; ay = byy; byy = t; }
    if (ay < cy) { t = ax; ax = cx; cx = t; t = ay; ay = cy; cy = t; }
    if (byy < cy) { t = bxx; bxx = cx; cx = t; t = byy; byy = cy; cy = t; }
    r.v0y = ay; r.v1y = byy; r.v0x = ax; r.v1x = bxx;
    r.p20y = rcpf(cy - ay);
    r.p20x = cx - ax;
    r.p21xy = (cx - bxx) * rcpf(cy - byy);
    r.p10xy = (bxx - ax) * rcpf(byy - ay);
}

// SL_ScanlineBounds::step, generic branch (:310-322); only the selected edge
// is converted (the other may be +-inf or NaN).
__device__ __forceinline__ void scanline_step(const SetupRec& r, float yf, int32_t& xMin, int32_t& xMax)
{
    const float d0 = yf - r.v0y;
    const float d1 = yf - r.v1y;
    const float alpha = d0 * r.p20y;
    const int32_t lo = f2i((r.p20x * alpha) + r.v0x);
    const float he = (d1 < 0.f) ? ((r.p21xy * d1) + r.v1x) : ((r.p10xy * d0) + r.v0x);
    const int32_t hi = f2i(he);
    xMin = (lo >= hi) ? hi : lo;
    xMax = (lo >= hi) ? lo : hi;
}

// The tiles holding the pixels a record actually covers, found by running
// the raster's own row test (scanline_step, the unsigned span check, the
// W/H clamps) over the record's rows; false when it covers no pixel.  Used
// for records of at most kExactRows rows.
constexpr int kExactRows = 8;

__device__ __forceinline__ bool exact_tile_rect(const SetupRec& r, int W, int H, int tilesX, int tilesY, int shx, int shy,
                                                TileRect& out)
{
    int xlo = 0x7FFFFFFF, xhi = -1, ylo = 0x7FFFFFFF, yhi = -1;
    const int y0 = max(r.minY, 0), y1 = min(r.maxY, H);
    for (int y = y0; y < y1; ++y)
    {
        int32_t xMin, xMax;
        scanline_step(r, (float)y, xMin, xMax);
        if (xMin >= 0 && (uint32_t)xMin < (uint32_t)xMax && xMin < W)
        {
            xlo = min(xlo, xMin);
            xhi = max(xhi, min(xMax, W) - 1);
            ylo = min(ylo, y);
            yhi = y;
        }
    }
    if (xhi < 0) return false;
    out.tx0 = xlo >> shx; out.tx1 = min(xhi >> shx, tilesX - 1);
    out.ty0 = ylo >> shy; out.ty1 = min(yhi >> shy, tilesY - 1);
    return true;
}

// ---------------------------------------------------------------------------
// Samplers (SL_Sampler.hpp:51-218): REPEAT wrap, bilinear with the x+1/y+1
// neighbours clamped to W/H (not W-1/H-1), ordered texel layout x + W*y.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float wrap_repeat(float u)
{
    return (u < 0.f ? 1.f : 0.f) + (u - truncf(u));
}

__device__ __forceinline__ uint32_t to_u16(float f) { return ((uint32_t)f2i(f)) & 0xFFFFu; }

struct BilinearTaps { uint32_t i0, i1, i2, i3; float dx, dy, omdx, omdy; };

__device__ __forceinline__ BilinearTaps bilinear_taps(uint32_t W, uint32_t H, float u, float v)
{
    BilinearTaps t;
    const float xf = wrap_repeat(u) * (float)W;
    const float yf = wrap_repeat(v) * (float)H;
    const uint32_t xi0 = to_u16(xf), yi0 = to_u16(yf);
    const uint32_t xi1 = min(xi0 + 1u, W), yi1 = min(yi0 + 1u, H);
    t.dx = xf - (float)xi0;
    t.dy = yf - (float)yi0;
    t.omdx = 1.f - t.dx;
    t.omdy = 1.f - t.dy;
    t.i0 = xi0 + W * yi0;
    t.i1 = xi0 + W * yi1;
    t.i2 = xi1 + W * yi0;
    t.i3 = xi1 + W * yi1;
    return t;
}

__device__ __forceinline__ float bilerp(const BilinearTaps& t, float p0, float p1, float p2, float p3)
{
    const float w0 = (p0 * t.omdx) * t.omdy;
    const float w1 = (p1 * t.omdx) * t.dy;
    const float w2 = (p2 * t.dx) * t.omdy;
    const float w3 = (p3 * t.dx) * t.dy;
    return ((w0 + w1) + w2) + w3;
}

// sl_sample_bilinear<RGBA8, REPEAT>: result cast back to u8 by truncation
__device__ __forceinline__ uint32_t sample_bilinear_rgba8(const uint8_t* tex, uint32_t W, uint32_t H, float u, float v)
{
    const BilinearTaps t = bilinear_taps(W, H, u, v);
    const auto px = gptr(reinterpret_cast<const uint32_t*>(tex));
    const uint32_t q0 = px[t.i0], q1 = px[t.i1], q2 = px[t.i2], q3 = px[t.i3];
    uint32_t out = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c)
    {
        const int s = 8 * c;
        const float s0 = (float)((q0 >> s) & 0xFFu) * 0.00392156862745f;
        const float s1 = (float)((q1 >> s) & 0xFFu) * 0.00392156862745f;
        const float s2 = (float)((q2 >> s) & 0xFFu) * 0.00392156862745f;
        const float s3 = (float)((q3 >> s) & 0xFFu) * 0.00392156862745f;
        const uint32_t b = ((uint32_t)f2i(bilerp(t, s0, s1, s2, s3) * 255.f)) & 0xFFu;
        out |= b << s;
    }
    return out;
}

// sl_sample_nearest<..., REPEAT> (SL_Sampler.hpp:122-148): the texel at
// (uint)(W * wrap(u)), (uint)(H * wrap(v)); wrap(u) = 1 (a tiny negative u)
// reads texel 0 of the next row, or the zeroed padding below the last row.
__device__ __forceinline__ uint32_t nearest_index(uint32_t W, uint32_t H, float u, float v)
{
    const uint32_t xi = (uint32_t)f2i((float)W * wrap_repeat(u));
    const uint32_t yi = (uint32_t)f2i((float)H * wrap_repeat(v));
    return xi + W * yi;
}

__device__ __forceinline__ uint32_t sample_nearest_rgba8(const uint8_t* tex, uint32_t W, uint32_t H, float u, float v)
{
    return gptr(reinterpret_cast<const uint32_t*>(tex))[nearest_index(W, H, u, v)];
}

__device__ __forceinline__ float4 sample_nearest_rgbaf(const uint8_t* tex, uint32_t W, uint32_t H, float u, float v)
{
    const auto px = gptr(reinterpret_cast<const float*>(tex)) + 4u * nearest_index(W, H, u, v);
    return f4(px[0], px[1], px[2], px[3]);
}

// The nearest sample as bilinear taps (one texel, weight 1) for the adjoint.
__device__ __forceinline__ BilinearTaps nearest_taps(uint32_t W, uint32_t H, float u, float v)
{
    BilinearTaps t;
    t.i0 = t.i1 = t.i2 = t.i3 = nearest_index(W, H, u, v);
    t.dx = t.dy = 0.f;
    t.omdx = t.omdy = 1.f;
    return t;
}

__device__ __forceinline__ float4 sample_bilinear_rgbaf(const uint8_t* tex, uint32_t W, uint32_t H, float u, float v)
{
    const BilinearTaps t = bilinear_taps(W, H, u, v);
    const auto px = gptr(reinterpret_cast<const float*>(tex));
    float o[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        o[k] = bilerp(t, px[4 * t.i0 + k], px[4 * t.i1 + k], px[4 * t.i2 + k], px[4 * t.i3 + k]);
    return f4(o[0], o[1], o[2], o[3]);
}

// Built-in fragment programs (replace SL_FragmentShader::shader, SL_Shader.hpp:89-98)
// math::clamp(x, 0, 1) as min(max(x, 0), 1) with the reference's max / min
// operand order (NaN stays NaN).
__device__ __forceinline__ float clamp01_ref(float x)
{
    const float m = (x < 0.f) ? 0.f : x;
    return (1.f < m) ? 1.f : m;
}

__device__ __forceinline__ float4 clamp01_ref4(float4 v)
{
    return f4(clamp01_ref(v.x), clamp01_ref(v.y), clamp01_ref(v.z), clamp01_ref(v.w));
}

// extra[0..2]: outputs 1..3 (SL_PROGRAM_LAMBERT_MRT only).
__device__ __forceinline__ bool fragment_program(const DrawDesc& d, const uint8_t* uniArena, const Vary4& var, float4& out,
                                                 float4* extra = nullptr)
{
    if (d.fsProg == SL_PROGRAM_PASSTHROUGH_GOURAUD)
    {
        out = var.v[0];
        return true;
    }
    if (d.fsProg == SL_PROGRAM_LAMBERT_TEXTURED || d.fsProg == SL_PROGRAM_LAMBERT_MRT)
    {
        const sl_std_uniforms* u = reinterpret_cast<const sl_std_uniforms*>(uniArena + d.uniOffset);
        uint32_t texel = 0;
        if (d.tex != nullptr && d.texType == SL_COLOR_RGBA_8U)
            texel = d.nearest ? sample_nearest_rgba8(d.tex, d.texW, d.texH, var.v[0].x, var.v[0].y)
                              : sample_bilinear_rgba8(d.tex, d.texW, d.texH, var.v[0].x, var.v[0].y);
        const float a0 = (float)(texel & 0xFFu) * 0.00392156862745f;
        const float a1 = (float)((texel >> 8) & 0xFFu) * 0.00392156862745f;
        const float a2 = (float)((texel >> 16) & 0xFFu) * 0.00392156862745f;
        const float a3 = (float)((texel >> 24) & 0xFFu) * 0.00392156862745f;
        const float nx = var.v[1].x, ny = var.v[1].y, nz = var.v[1].z;
        const float len2 = ((nx * nx) + (ny * ny)) + (nz * nz);
        const float inv = 1.0f / sqrtf(len2);
        const float mx = nx * inv, my = ny * inv, mz = nz * inv;
        const float ndl = -(((mx * u->light_dir[0]) + (my * u->light_dir[1])) + (mz * u->light_dir[2]));
        const float lam = (ndl > 0.f) ? ndl : 0.f;
        const float alb[3] = {a0, a1, a2};
        float o[3];
#pragma unroll
        for (int c = 0; c < 3; ++c)
        {
            const float light = u->ambient[c] + (u->diffuse[c] * lam);
            const float v = (alb[c] * light) * u->tint[c];
            o[c] = (v < 1.f) ? v : 1.f;
        }
        out = f4(o[0], o[1], o[2], a3 * u->tint[3]);
        if (extra != nullptr && d.fsProg == SL_PROGRAM_LAMBERT_MRT)
        {
            // the G-buffer outputs of tests/sl_mrt_test.cpp:105-123: position, uv, normal
            extra[0] = clamp01_ref4(var.v[2]);
            extra[1] = clamp01_ref4(var.v[0]);
            extra[2] = clamp01_ref4(f4(mx, my, mz, 0.f));
        }
        return true;
    }
    if (d.fsProg == SL_PROGRAM_COLOR_TEXTURED)
    {
        float4 t = f4(1.f, 1.f, 1.f, 1.f);
        if (d.tex != nullptr && d.texType == SL_COLOR_RGBA_FLOAT)
            t = d.nearest ? sample_nearest_rgbaf(d.tex, d.texW, d.texH, var.v[1].x, var.v[1].y)
                          : sample_bilinear_rgbaf(d.tex, d.texW, d.texH, var.v[1].x, var.v[1].y);
        out = f4(var.v[0].x * t.x, var.v[0].y * t.y, var.v[0].z * t.z, var.v[0].w * t.w);
        return true;
    }
    out = f4(0.f, 0.f, 0.f, 0.f);
    return true;
}

// ---------------------------------------------------------------------------
// Output merger (SL_Framebuffer.cpp:44-80,153-166,270-368)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float clamp01(float x) { return (x < 0.f) ? 0.f : ((x > 1.f) ? 1.f : x); }

__device__ __forceinline__ void blend(int mode, float4 s, float d[4])
{
    const float as = s.w;
    const float mod = 1.f - as;
    const float sv[4] = {s.x, s.y, s.z, s.w};
    if (mode == SL_BLEND_ALPHA)
    {
        const float dstMod = mod * d[3];
        const float dstAlpha = dstMod + as;
        const float r = rcpf(dstAlpha);
        float o[4];
#pragma unroll
        for (int c = 0; c < 3; ++c) o[c] = fmaf(sv[c], as, d[c] * dstMod) * r;
        o[3] = dstAlpha;
#pragma unroll
        for (int c = 0; c < 4; ++c) d[c] = o[c];
    }
    else if (mode == SL_BLEND_PREMULTIPLED_ALPHA)
    {
#pragma unroll
        for (int c = 0; c < 4; ++c) d[c] = fmaf(d[c], mod, sv[c]);
    }
    else if (mode == SL_BLEND_ADDITIVE)
    {
#pragma unroll
        for (int c = 0; c < 4; ++c) d[c] = fmaf(sv[c], as, d[c]);
    }
    else if (mode == SL_BLEND_SCREEN)
    {
#pragma unroll
        for (int c = 0; c < 4; ++c) d[c] = (sv[c] * as) + (d[c] * mod);
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) d[c] = clamp01(d[c]);
}

__device__ __forceinline__ uint32_t pack_rgba8(const float c[4])
{
    uint32_t o = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) o |= (((uint32_t)f2i(c[k] * 255.f)) & 0xFFu) << (8 * k);
    return o;
}

__device__ __forceinline__ void unpack_rgba8(uint32_t p, float d[4])
{
#pragma unroll
    for (int k = 0; k < 4; ++k) d[k] = (float)((p >> (8 * k)) & 0xFFu) * 0.00392156862745f;
}

__device__ __forceinline__ bool depth_test(int f, float a, float b)
{
    switch (f)
    {
        case SL_DEPTH_TEST_OFF: return true;
        case SL_DEPTH_TEST_LESS_THAN: return a < b;
        case SL_DEPTH_TEST_LESS_EQUAL: return a <= b;
        case SL_DEPTH_TEST_GREATER_THAN: return a > b;
        case SL_DEPTH_TEST_GREATER_EQUAL: return a >= b;
        case SL_DEPTH_TEST_EQUAL: return a == b;
        case SL_DEPTH_TEST_NOT_EQUAL: return a != b;
        default: return false;
    }
}

}  // namespace slg
