// sl_kernels.hip -- gfx950 kernels of the SoftLight triangle pipeline.
//
//   k_setup    counting pass: position part of the vertex programs,
//              culling, clip classification, perspective divide, viewport
//              snap and bounding-box reject for one triangle per lane; per
//              256-triangle chunk, the mask of triangles that yield records
//              and the packed record / bin-entry total (position-only
//              clipping counts the straddling triangles).
//   k_emit     emitting pass: global offsets from the chunk totals, then the
//              surviving triangles' plane setup, varyings, records and
//              (tile, record) bin entries, in submission order.  Straddling
//              triangles get their slots reserved and are finished by k_clip.
//   k_clip     6-plane Sutherland-Hodgman clipping with varyings and the
//              reversed fan, writing into the reserved slots.
//   k_bin_*    ordered binning: stable counting sort of the bin entries by tile.
//   k_raster   one wave per screen tile: walks the tile's records in
//              submission order; lane = (row, x-phase) replays the
//              reference's scanline span and its sequential barycentric
//              recurrence, depth-tests, shades and blends with the tile's
//              colour/depth held in LDS.
//
// Reference routines reproduced: SL_TriProcessor::process_verts / push_bin /
// clip_and_process_tris (SL_TriProcessor.cpp:248-640), SL_TriRasterizer::
// render_triangle_simd generic branch (SL_TriRasterizer.cpp:806-923),
// SL_FragmentProcessor::flush_tri_fragments (SL_FragmentProcessor.cpp:416-531).
#include <hip/hip_fp16.h>

#include "sl_device.h"
#include "sl_recurrence.h"
#include "sl_stages.h"
#include "sl_kernels.h"

namespace slg {

constexpr uint32_t kRecBits = 30;                        // packed counts: bin entries << kRecBits | records
constexpr int kVertexSlotBits = 9;                       // per-chunk vertex cache of k_setup
constexpr int kVertexSlots = 1 << kVertexSlotBits;
constexpr unsigned long long kNoVertex = ~0ull;

// ---------------------------------------------------------------------------
// small utilities
// ---------------------------------------------------------------------------
__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, uint64_t n)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_fill_u64(unsigned long long* __restrict__ p, unsigned long long v, uint64_t n)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}

// Per-frame reset of the control block: every word is cleared except
// `sticky`, which gathers the previous frame's overflow and error flags.
__global__ void __launch_bounds__(256) k_ctrl_reset(FrameCtrl* __restrict__ ctrl)
{
    __shared__ uint32_t prev;
    if (threadIdx.x == 0) prev = ctrl->overflow | (ctrl->error << 1);
    __syncthreads();
    uint32_t* w = reinterpret_cast<uint32_t*>(ctrl);
    constexpr uint32_t kWords = sizeof(FrameCtrl) / 4;
    constexpr uint32_t kSticky = offsetof(FrameCtrl, sticky) / 4;
    for (uint32_t i = threadIdx.x; i < kWords; i += blockDim.x)
        w[i] = (i == kSticky) ? (w[i] | prev) : 0u;
}

__device__ __forceinline__ uint32_t find_draw(const DrawDesc* __restrict__ draws, uint32_t n, uint64_t g)
{
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi)
    {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (draws[mid].triBase <= g) lo = mid; else hi = mid - 1;
    }
    return lo;
}


struct TriIn
{
    float4 pos[3];
    Vary4 var[3];
    uint32_t vid[3];          // vertex ids (element indices < 2^32, checked on the host)
    uint32_t inst;
    uint32_t draw;
    uint32_t seq;
    int status;               // 0 culled/invisible, 1 partial, 3 full
};

// The three vertex programs' varyings (and, with a visibility attachment, the
// source-vertex weights, which travel as one more varying so that clipping
// and perspective correction treat them like the others).
__device__ __forceinline__ void load_varyings(const SetupParams& P, const DrawDesc& D, TriIn& T)
{
    vertex_varyings3(D, P.uniforms, T.vid, T.inst, T.var);
    if (D.visSlot != kNoVisSlot)
    {
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int i = 0; i < SL_SHADER_MAX_VARYING_VECTORS; ++i)
                if ((uint32_t)i == D.visSlot) T.var[k].v[i] = f4(k == 0 ? 1.f : 0.f, k == 1 ? 1.f : 0.f, k == 2 ? 1.f : 0.f, 0.f);
    }
}

// Vertex ids, instance and sequence number of global triangle g.
__device__ __forceinline__ void load_triangle_ids(uint64_t g, const DrawDesc& D, uint32_t di, TriIn& T)
{
    const uint64_t t = g - D.triBase;
    const uint64_t inst = t / D.trisPerInstance;
    const uint64_t ti = t - inst * D.trisPerInstance;
    // a point repeats its vertex in all three slots, a line its second
    const uint64_t per = verts_per_prim(D.prim);
    const uint64_t e = D.elementBegin + per * ti;
    const uint64_t e1 = (per > 1) ? e + 1 : e, e2 = (per > 2) ? e + 2 : e1;
    T.vid[0] = (uint32_t)e; T.vid[1] = (uint32_t)e1; T.vid[2] = (uint32_t)e2;
    if (D.indexed)
    {
        T.vid[0] = (uint32_t)fetch_index(D, e);
        T.vid[1] = (uint32_t)fetch_index(D, e1);
        T.vid[2] = (uint32_t)fetch_index(D, e2);
    }
    T.inst = (uint32_t)inst;
    T.draw = di;
    T.seq = D.seqBase + 1u + (uint32_t)t;
}

// Cull, scissor matrix and clip status of a triangle whose clip-space
// positions are in T.pos.  Points and lines are not culled; a point is
// visible (3) when inside the clip volume (sl_ndc_clip_status, one vertex,
// SL_PostVertexTransform.hpp:536-554); a line is dropped when a vertex has
// w < 0, else gets the two-vertex status (:560-587; SL_LineProcessor.cpp:
// 256-265).
// kKinds: 1 = the draw draws triangles, 2 = points or lines, 3 = either.
template <int kKinds = 3>
__device__ __forceinline__ void classify_triangle(const DrawDesc& D, TriIn& T)
{
    if ((kKinds & 2) && (D.prim == kPrimPoint || D.prim == kPrimLine))
    {
        T.pos[0] = mat_mul(D.scissor, T.pos[0]);
        T.pos[1] = mat_mul(D.scissor, T.pos[1]);
        if (D.prim == kPrimPoint)
        {
            T.status = (T.pos[0].w > 0.f && clip_inside(T.pos[0])) ? 3 : 0;
        }
        else if (T.pos[0].w < 0.f || T.pos[1].w < 0.f)
        {
            T.status = 0;
        }
        else
        {
            T.status = ((clip_inside(T.pos[0]) && clip_inside(T.pos[1])) ? 3 : 0) |
                       ((T.pos[0].w > 0.f || T.pos[1].w > 0.f) ? 1 : 0);
        }
        return;
    }
    if (!(kKinds & 1)) return;
    if (D.cull != SL_CULL_OFF)
    {
        const float det = face_determinant(T.pos[0], T.pos[1], T.pos[2]);
        const bool culled = (D.cull == SL_CULL_FRONT_FACE) ^ (bool)signbit(det);
        if (culled) { T.status = 0; return; }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) T.pos[k] = mat_mul(D.scissor, T.pos[k]);
    T.status = clip_status(T.pos[0], T.pos[1], T.pos[2]);
}

// Clip positions of a triangle's corners from the vertex pass (a draw with
// posBase != kNoPositions): an id the pass covered is read, any other (an
// index past the vertex buffer) runs the position program here.  The pass ran
// the same program on the same inputs, so the bits equal the per-corner
// evaluation's.
__device__ __forceinline__ void positions_from_pass(const SetupParams& P, const DrawDesc& D, const uint32_t vid[3],
                                                    uint64_t inst, const bool want[3], float4 out[3])
{
    bool rest[3];
    bool any = false;
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        rest[k] = want[k] && vid[k] >= D.posVerts;
        any = any || rest[k];
        if (want[k] && !rest[k]) out[k] = P.positions[D.posBase + inst * D.posVerts + vid[k]];
    }
    if (any) vertex_positions3(D, P.uniforms, vid, inst, rest, out);
}

// Vertex positions, cull, scissor matrix and clip status of a triangle whose
// vertex ids and instance are in T.
template <int kKinds = 3>
__device__ __forceinline__ void load_positions(const SetupParams& P, const DrawDesc& D, TriIn& T)
{
    const uint32_t per = verts_per_prim(D.prim);
    const bool need[3] = {true, per > 1, per > 2};
    if (D.posBase != kNoPositions) positions_from_pass(P, D, T.vid, T.inst, need, T.pos);
    else vertex_positions3(D, P.uniforms, T.vid, T.inst, need, T.pos);
    classify_triangle<kKinds>(D, T);
}

// Fetch, vertex positions, cull, scissor matrix and clip status of one
// triangle (SL_TriProcessor::process_verts, SL_TriProcessor.cpp:574-631).
// Varyings are left to load_varyings: most triangles never need them.
template <int kKinds = 3>
__device__ __forceinline__ void load_triangle_pos(const SetupParams& P, uint64_t g, const DrawDesc& D, uint32_t di, TriIn& T)
{
    load_triangle_ids(g, D, di, T);
    load_positions<kKinds>(P, D, T);
}

template <bool PAIRS = true>
__device__ __forceinline__ void write_record(const SetupParams& P, uint32_t recId, uint32_t pairOff,
                                             float4 p0, float4 p1, float4 p2, const Vary4& a, const Vary4& b,
                                             const Vary4& c, int nv, uint32_t draw, uint32_t seq, uint32_t state,
                                             const TileRect& tr)
{
    SetupRec r;
    setup_record(p0, p1, p2, draw, seq, r);
    r.state = state;
    P.recs[recId] = r;
    float4* V = P.vars + (uint64_t)recId * P.varStride;
    const uint32_t nvMax = P.varStride / 3;
#pragma unroll
    for (int i = 0; i < SL_SHADER_MAX_VARYING_VECTORS; ++i)   // static indices keep the varyings in registers
    {
        if (i < nv)
        {
            V[0 * nvMax + i] = a.v[i];
            V[1 * nvMax + i] = b.v[i];
            V[2 * nvMax + i] = c.v[i];
        }
    }
    if (!PAIRS) return;
    uint32_t o = pairOff;
    for (int ty = tr.ty0; ty <= tr.ty1; ++ty)
        for (int tx = tr.tx0; tx <= tr.tx1; ++tx)
        {
            P.pairKeys[o] = (uint32_t)(ty * P.tilesX + tx);
            P.pairVals[o] = recId;
            ++o;
        }
}

// The (tile, record) entries of a wave's records, entry j of a record at
// pairOff + j in row-major order over its tile rectangle (as write_record
// writes them).  When some record of the wave covers more than kLanePairs
// tiles and more than kShareRatio times the wave's mean, the wave writes all
// of its records' entries together, 64 a step:
// entry k of the wave's concatenated lists goes to the lane k % 64, which
// finds its record by a binary search over the lanes' inclusive counts.
// Called by every lane of the wave; `has` marks the lanes with a record.
constexpr uint32_t kLanePairs = 8, kShareRatio = 6;

__device__ __forceinline__ void write_pairs_wave(const SetupParams& P, bool has, uint32_t recId, uint32_t pairOff,
                                                 const TileRect& tr)
{
    const int lane = threadIdx.x & 63;
    const int wi = tr.tx1 - tr.tx0 + 1, hi = tr.ty1 - tr.ty0 + 1;
    const bool any = has && wi > 0 && hi > 0;
    const uint32_t w = any ? (uint32_t)wi : 1u;
    const uint32_t n = any ? (uint32_t)wi * (uint32_t)hi : 0u;
    uint32_t mx = n, inc = n;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = max(mx, (uint32_t)__shfl_xor((int)mx, o, 64));
#pragma unroll
    for (int o = 1; o < 64; o <<= 1)
    {
        const uint32_t v = (uint32_t)__shfl_up((int)inc, o, 64);
        if (lane >= o) inc += v;
    }
    const uint32_t total = (uint32_t)__shfl((int)inc, 63, 64);
    // a lane's own loop costs its count; the shared loop costs the wave's
    // total / 64 steps of several times the work (the search and the
    // shuffles): share only when one lane's list dominates the wave
    if (mx <= kLanePairs || (uint64_t)mx * 64u <= (uint64_t)kShareRatio * total)
    {
        if (!any) return;
        uint32_t o = pairOff;
        for (int ty = tr.ty0; ty <= tr.ty1; ++ty)
            for (int tx = tr.tx0; tx <= tr.tx1; ++tx, ++o)
            {
                P.pairKeys[o] = (uint32_t)(ty * P.tilesX + tx);
                P.pairVals[o] = recId;
            }
        return;
    }
    const int tx0 = any ? tr.tx0 : 0, ty0 = any ? tr.ty0 : 0;
    for (uint32_t base = 0; base < total; base += 64)
    {
        const uint32_t k = base + (uint32_t)lane;
        // the first lane whose inclusive count exceeds k
        int L = 0;
#pragma unroll
        for (int step = 32; step > 0; step >>= 1)
        {
            const uint32_t v = (uint32_t)__shfl((int)inc, L + step - 1, 64);
            if (v <= k) L += step;
        }
        const uint32_t incL = (uint32_t)__shfl((int)inc, L, 64);
        const uint32_t nL = (uint32_t)__shfl((int)n, L, 64);
        const uint32_t wL = (uint32_t)__shfl((int)w, L, 64);
        const int tx0L = __shfl(tx0, L, 64), ty0L = __shfl(ty0, L, 64);
        const uint32_t offL = (uint32_t)__shfl((int)pairOff, L, 64);
        const uint32_t recL = (uint32_t)__shfl((int)recId, L, 64);
        if (k < total)
        {
            const uint32_t j = k - (incL - nL);
            P.pairKeys[offL + j] = (uint32_t)((ty0L + (int)(j / wL)) * P.tilesX + tx0L + (int)(j % wL));
            P.pairVals[offL + j] = recL;
        }
    }
}

// ---------------------------------------------------------------------------
// Setup: a counting pass (k_setup) and an emitting pass (k_emit).
//
// k_setup runs the position part of the vertex programs, classification and
// bin reservation for one triangle per lane and leaves, per chunk of
// kSetupBlock triangles, the mask of triangles that yield records and the
// chunk's packed (bin entries << kRecBits | records) total, which it also adds
// to its group and super-group sums.  Nothing waits on another workgroup.
// k_emit then turns the sums into each group's global offsets and writes the
// records in submission order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long wave_incl_scan(unsigned long long v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1)
    {
        const unsigned long long o = __shfl_up(v, off, 64);
        if (lane >= off) v += o;
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1)
    {
        const uint32_t o = __shfl_up(v, off, 64);
        if (lane >= off) v += o;
    }
    return v;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Last draw whose first triangle is <= g (draws are sorted by triBase and
// draws[0].triBase == 0), searched 64 candidates per round by the calling
// wave; every lane returns the result.
__device__ __forceinline__ uint32_t find_draw_wave(const DrawDesc* __restrict__ draws, uint32_t n, uint64_t g)
{
    const uint32_t lane = threadIdx.x & 63;
    uint32_t lo = 0, hi = n;   // the answer is in [lo, hi); draws[lo].triBase <= g
    while (hi - lo > 1)
    {
        const uint32_t step = (hi - lo + 63) / 64;
        const uint32_t idx = lo + lane * step;
        const bool ok = idx < hi && draws[idx].triBase <= g;
        const unsigned long long b = __ballot(ok);   // a prefix of lanes, lane 0 always set
        lo += (uint32_t)(63 - __clzll((long long)b)) * step;
        hi = min(hi, lo + step);
    }
    return lo;
}

// Records and bin entries a triangle yields, from its classified clip-space
// positions (T.pos after classify_triangle).  Both setup passes run this.
struct TriCount
{
    float4 sp[3];              // snapped screen positions (unclipped triangles; a line's two ends; a point's texel)
    TileRect tr;               // tiles to bin to (unclipped triangles, points, lines)
    TileRect tr2;              // a line's further tiles: column 0, rows its wrapped texels reach
    float t0, t1;              // a clipped line's varying mix weights at its two ends
    uint32_t nrec, npair;
    bool clipped;              // needs k_clip (nrec fan triangles)
    bool empty;                // small triangle that covers no pixel: dropped, counted as binned
    bool needClip;             // straddles a clip plane: counted by count_clipped
    bool lineClip;             // a line clipped to the viewport (varyings mixed by t0, t1)
};

// ---------------------------------------------------------------------------
// Points and lines
// ---------------------------------------------------------------------------
// Texel (x, y) addresses attachment texel x + W*y (SL_Framebuffer.cpp:33-38):
// a snapped x of W (NDC +1, or a line clipped to the viewport's right edge)
// is texel 0 of row y + 1.  False at or past W*H, where the reference writes
// outside the attachment's texels; those fragments are dropped.
__device__ __forceinline__ bool linear_texel(int x, int y, int W, int H, int& px, int& py)
{
    const int64_t idx = (int64_t)x + (int64_t)W * y;
    if (x < 0 || y < 0 || idx >= (int64_t)W * H) return false;
    py = (int)(idx / W);
    px = (int)(idx - (int64_t)py * W);
    return true;
}

// LS_EPSILON (LightSetup, absent).  It is compared against differences of
// snapped screen coordinates (whole numbers), so any value in (0, 1] gives
// the same results.
constexpr float kLineEpsilon = 1.0e-5f;

__device__ __forceinline__ bool line_clip_segment(float num, float denom, float& tE, float& tL)
{
    if (fabsf(denom) < kLineEpsilon) return num < 0.f;
    const float t = num / denom;
    if (denom > 0.f)
    {
        if (t > tL) return false;
        if (t > tE) tE = t;
    }
    else
    {
        if (t < tE) return false;
        if (t < tL) tL = t;
    }
    return true;
}

// SL_LineProcessor::clip_and_process_lines (SL_LineProcessor.cpp:69-169):
// Liang-Barsky in screen space against viewportDims read as (xMin, yMin,
// xMax, yMax).  Returns false when nothing is left; a, b become the ends,
// with depth and w mixed toward the other end by t0 / t1 when clipped.
__device__ __forceinline__ bool clip_line(float4& a, float4& b, const float vp[4], float& t0, float& t1, bool& clipped)
{
    const float dx = b.x - a.x, dy = b.y - a.y;
    const float xMin = vp[0], yMin = vp[1], xMax = vp[2], yMax = vp[3];
    clipped = false;
    t0 = t1 = 0.f;
    if (fabsf(dx) < kLineEpsilon && fabsf(dy) < kLineEpsilon && a.x >= xMin && a.x <= xMax && a.y >= yMin && a.y <= yMax)
        return true;
    const float ex = a.x - b.x, ey = a.y - b.y;
    const float dist = 1.f / sqrtf((ex * ex) + (ey * ey));
    float tE = 0.f, tL = 1.f;
    const float x1 = a.x, y1 = a.y;
    if (!(line_clip_segment(xMin - x1, dx, tE, tL) && line_clip_segment(x1 - xMax, -dx, tE, tL) &&
          line_clip_segment(yMin - y1, dy, tE, tL) && line_clip_segment(y1 - yMax, -dy, tE, tL)))
        return false;
    float4 p0 = a, p1 = b;
    if (tL < 1.f)
    {
        p1.x = x1 + tL * dx;
        p1.y = y1 + tL * dy;
    }
    if (tE > 0.f)
    {
        p0.x = x1 + tE * dx;
        p0.y = y1 + tE * dy;
    }
    const float4 d0 = f4(p0.x - a.x, p0.y - a.y, p0.z - a.z, p0.w - a.w);
    const float4 d1 = f4(p1.x - b.x, p1.y - b.y, p1.z - b.z, p1.w - b.w);
    t0 = sqrtf(dot4(d0, d0)) * dist;   // length(p0 - a) * dist
    t1 = sqrtf(dot4(d1, d1)) * dist;
    p0.z = a.z + t0 * (b.z - a.z);
    p0.w = a.w + t0 * (b.w - a.w);
    p1.z = b.z + t1 * (a.z - b.z);
    p1.w = b.w + t1 * (a.w - b.w);
    a = p0;
    b = p1;
    clipped = true;
    return true;
}

// The walk of sl_draw_line_bresenham (SL_Geometry.hpp:314-366) between the
// (uint16_t) casts of two screen ends (SL_LineRasterizer.cpp:44-48).
__device__ __forceinline__ LineWalk make_line_walk(float ax, float ay, float bx, float by)
{
    int x1 = (int16_t)(uint16_t)f2i(ax), y1 = (int16_t)(uint16_t)f2i(ay);
    int x2 = (int16_t)(uint16_t)f2i(bx), y2 = (int16_t)(uint16_t)f2i(by);
    const bool steep = abs(x1 - x2) < abs(y1 - y2);
    int t;
    if (steep)
    {
        t = x1; x1 = y1; y1 = t;
        t = x2; x2 = y2; y2 = t;
    }
    if (x1 > x2)
    {
        t = x1; x1 = x2; x2 = t;
        t = y1; y1 = y2; y2 = t;
    }
    LineWalk w;
    w.X1 = x1;
    w.Y1 = y1;
    w.dx = x2 - x1;
    w.dErr = abs(y2 - y1) * 2;
    w.yErr = (y2 > y1) ? 1 : -1;
    w.steep = steep ? 1 : 0;
    return w;
}

// A line's addressed texels: those of its walk inside the target
// (columns dx0..dx1, rows dy0..dy1, either range empty when none), and, when
// the walk reaches column W, the wrapped ones in column 0, rows wy0..wy1.
struct LineBox { int dx0, dx1, dy0, dy1, wy0, wy1; };

__device__ __forceinline__ bool line_box(const LineWalk& w, int W, int H, LineBox& b)
{
    const int majLo = w.X1, majHi = w.X1 + w.dx;
    const int minEnd = w.Y1 + w.yErr * (w.dErr / 2);
    const int minLo = min(w.Y1, minEnd), minHi = max(w.Y1, minEnd);
    const int vx0 = w.steep ? minLo : majLo, vx1 = w.steep ? minHi : majHi;
    const int vy0 = w.steep ? majLo : minLo, vy1 = w.steep ? majHi : minHi;
    b.dx0 = max(vx0, 0);
    b.dx1 = min(vx1, W - 1);
    b.dy0 = max(vy0, 0);
    b.dy1 = min(vy1, H - 1);
    b.wy0 = 1;
    b.wy1 = 0;
    if (vx0 <= W && vx1 >= W)   // the walk passes column W (columns beyond it address nothing kept)
    {
        b.wy0 = max(vy0 + 1, 1);
        b.wy1 = min(vy1 + 1, H - 1);
    }
    return (b.dx0 <= b.dx1 && b.dy0 <= b.dy1) || b.wy0 <= b.wy1;
}

__device__ __forceinline__ TileRect rect_tiles(int x0, int x1, int y0, int y1, int tilesX, int tilesY, int shx, int shy)
{
    TileRect r;
    if (x0 > x1 || y0 > y1) { r.tx0 = 0; r.tx1 = -1; r.ty0 = 0; r.ty1 = -1; return r; }
    r.tx0 = x0 >> shx; r.tx1 = min(x1 >> shx, tilesX - 1);
    r.ty0 = y0 >> shy; r.ty1 = min(y1 >> shy, tilesY - 1);
    return r;
}

__device__ __forceinline__ void count_point(const SetupParams& P, const DrawDesc& D, const TriIn& T, TriCount& C)
{
    if (T.status != 3) return;
    const float4 s = divide_and_snap(T.pos[0], D.vp);
    int px, py;
    if (!linear_texel((uint16_t)f2i(s.x), (uint16_t)f2i(s.y), P.W, P.H, px, py))
    {
        C.empty = true;   // binned by the reference, addresses no kept texel
        return;
    }
    C.sp[0] = f4((float)px, (float)py, s.z, s.w);
    C.tr = rect_tiles(px, px, py, py, P.tilesX, P.tilesY, P.tileShiftX, P.tileShiftY);
    C.nrec = 1;
    C.npair = 1;
}

__device__ __forceinline__ void count_line(const SetupParams& P, const DrawDesc& D, const TriIn& T, TriCount& C)
{
    if (T.status == 0) return;
    float4 a = divide_and_snap(T.pos[0], D.vp), b = divide_and_snap(T.pos[1], D.vp);
    if (T.status != 3 && !clip_line(a, b, D.vp, C.t0, C.t1, C.lineClip)) return;
    const LineWalk w = make_line_walk(a.x, a.y, b.x, b.y);
    LineBox box;
    if (!line_box(w, P.W, P.H, box))
    {
        C.empty = true;   // binned by the reference, addresses no kept texel
        return;
    }
    C.sp[0] = a;
    C.sp[1] = b;
    C.tr = rect_tiles(box.dx0, box.dx1, box.dy0, box.dy1, P.tilesX, P.tilesY, P.tileShiftX, P.tileShiftY);
    // wrapped texels: column 0, minus the tile rows tr already holds there
    TileRect r2 = rect_tiles(0, 0, box.wy0, box.wy1, P.tilesX, P.tilesY, P.tileShiftX, P.tileShiftY);
    if (tile_count(C.tr) > 0 && C.tr.tx0 == 0) r2.ty0 = max(r2.ty0, C.tr.ty1 + 1);
    C.tr2 = r2;
    C.nrec = 1;
    C.npair = tile_count(C.tr) + tile_count(C.tr2);
}

template <int kKinds = 3>
__device__ __forceinline__ void count_triangle(const SetupParams& P, const DrawDesc& D, uint32_t drawIdx, const TriIn& T,
                                               TriCount& C)
{
    C.nrec = 0;
    C.npair = 0;
    C.clipped = false;
    C.empty = false;
    C.needClip = false;
    C.lineClip = false;
    C.tr = TileRect{0, -1, 0, -1};
    if ((kKinds & 2) && D.prim == kPrimPoint)
    {
        count_point(P, D, T, C);
        return;
    }
    if ((kKinds & 2) && D.prim == kPrimLine)
    {
        count_line(P, D, T, C);
        return;
    }
    if (!(kKinds & 1)) return;
    if (T.status == 3)
    {
#pragma unroll
        for (int k = 0; k < 3; ++k) C.sp[k] = divide_and_snap(T.pos[k], D.vp);
        if (bin_accept(C.sp[0], C.sp[1], C.sp[2]))
        {
            const float yMinF = fminf(fminf(C.sp[0].y, C.sp[1].y), C.sp[2].y);
            const float yMaxF = fmaxf(fmaxf(C.sp[0].y, C.sp[1].y), C.sp[2].y);
            if (f2i(yMaxF) - f2i(yMinF) <= kExactRows)
            {
                // small triangle: bin it to the tiles it covers, or drop it
                // when it covers no pixel (still counted as binned)
                SetupRec r;
                setup_record(C.sp[0], C.sp[1], C.sp[2], drawIdx, T.seq, r);
                if (exact_tile_rect(r, P.W, P.H, P.tilesX, P.tilesY, P.tileShiftX, P.tileShiftY, C.tr))
                {
                    C.nrec = 1;
                    C.npair = tile_count(C.tr);
                }
                else
                {
                    C.empty = true;
                }
            }
            else
            {
                C.tr = tile_rect(C.sp[0], C.sp[1], C.sp[2], P.W, P.H, P.tilesX, P.tilesY, P.tileShiftX, P.tileShiftY);
                C.nrec = 1;
                C.npair = tile_count(C.tr);
            }
        }
    }
    else if (T.status == 1 && clip_quick(T.pos))
    {
        C.needClip = true;   // counted by count_clipped
    }
}

// The straddling-triangle clipper shared by the counting pass and k_clip: one
// 16-lane group per triangle, lane k holding polygon vertex k.  At each plane
// every lane tests the edge from vertex k - 1 (the last vertex for k = 0) and
// writes the edge's intersection and its own vertex, when kept, into the
// group's next polygon at the position the group's ballots give it, so the
// output order and every operation on a vertex are those of the serial
// polygon loop (SL_TriProcessor.cpp:370-499).  pos / var: the group's two polygon buffers in LDS
// (var only with kVary); the triangle is in pos[0][0..2] (var[0][0..2]) on
// entry, n = 3 (0 for a group without one).  Every lane of the wave calls it
// (the planes end in wave-wide syncs).  Returns the polygon size (0 when
// fewer than 3 vertices are left) and its buffer in `cur`.
constexpr int kClipGroup = 16;
constexpr int kClipVerts = 9;   // 3 + one per plane

__device__ __forceinline__ void wave_sync()
{
    // LDS operations of one wave take effect in program order: fences keep
    // the compiler from moving them across
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool kVary>
__device__ __forceinline__ int group_clip(float4 (*pos)[kClipVerts], Vary4 (*var)[kClipVerts], int n, int nv, int& cur)
{
    const int lane = (int)(threadIdx.x & 63);
    const int gi = lane / kClipGroup, li = lane % kClipGroup;
    const unsigned long long gmask = 0xFFFFull << (gi * kClipGroup);
    const unsigned long long below = gmask & ((1ull << lane) - 1ull);
    cur = 0;
    wave_sync();
    for (int e = 0; e < 6; ++e)
    {
        const float4 E = clip_edge(e);
        const bool mine = li < n;
        const int pj = li == 0 ? n - 1 : li - 1;
        float4 q0 = f4(0.f, 0.f, 0.f, 0.f), q1 = q0;
        float t0 = 0.f, t1 = 0.f;
        bool vis0 = false, vis1 = false;
        if (mine)
        {
            q1 = pos[cur][li];
            q0 = pos[cur][pj];
            t1 = dot4(q1, E);
            t0 = dot4(q0, E);
            vis1 = !signbit(t1);
            vis0 = !signbit(t0);
        }
        const bool cross = mine && vis0 != vis1, keep = mine && vis1;
        const unsigned long long mc = __ballot(cross) & gmask, mk = __ballot(keep) & gmask;
        // a plane that keeps every vertex leaves the polygon as it is
        const bool allIn = __popcll(mk) == n;
        if (n > 0 && !allIn)
        {
            const int nn = __popcll(mc) + __popcll(mk);
            int b = __popcll(mc & below) + __popcll(mk & below);
            const int nxt = cur ^ 1;
            if (cross && b < kClipVerts)
            {
                const float t = t0 * rcpf(t0 - t1);
                pos[nxt][b] = mix4(q0, q1, t);
                if constexpr (kVary)
                {
#pragma unroll
                    for (int i = 0; i < SL_SHADER_MAX_VARYING_VECTORS; ++i)
                        if (i < nv) var[nxt][b].v[i] = mix4(var[cur][pj].v[i], var[cur][li].v[i], t);
                }
            }
            if (cross) ++b;
            if (keep && b < kClipVerts)
            {
                pos[nxt][b] = q1;
                if constexpr (kVary)
                {
#pragma unroll
                    for (int i = 0; i < SL_SHADER_MAX_VARYING_VECTORS; ++i)
                        if (i < nv) var[nxt][b].v[i] = var[cur][li].v[i];
                }
            }
            n = min(nn, kClipVerts);
            cur = nxt;
        }
        wave_sync();
    }
    return n < 3 ? 0 : n;
}

// Records and bin entries of the calling wave's straddling triangles
// (C.needClip), four at a time: group g of 16 lanes clips the g-th pending
// straddler (its positions read from its lane) with group_clip, then its
// lanes set up one fan triangle each and the group's sums go back to the
// straddler's lane.  Call with every lane of the wave (inactive lanes with
// needClip false); D is the wave's draw; slots: the wave's polygon buffers.
__device__ __forceinline__ void count_clipped(const SetupParams& P, const DrawDesc& D, const TriIn& T, TriCount& C,
                                              float4 (*slots)[2][kClipVerts])
{
    const int lane = threadIdx.x & 63;
    const int gi = lane / kClipGroup, li = lane % kClipGroup;
    for (unsigned long long m = __ballot(C.needClip); m != 0ull; m = __ballot(C.needClip))
    {
        // this group's straddler: the gi-th pending lane
        unsigned long long mm = m;
        for (int g = 0; g < gi && mm != 0ull; ++g) mm &= mm - 1ull;
        const int src = mm ? __ffsll((long long)mm) - 1 : -1;
        const int from = src < 0 ? lane : src;
        float4 p[3];
#pragma unroll
        for (int k = 0; k < 3; ++k)
            p[k] = f4(__shfl(T.pos[k].x, from, 64), __shfl(T.pos[k].y, from, 64), __shfl(T.pos[k].z, from, 64),
                      __shfl(T.pos[k].w, from, 64));
        float4 (*pos)[kClipVerts] = slots[gi];
        if (src >= 0 && li == 0)
        {
#pragma unroll
            for (int k = 0; k < 3; ++k) pos[0][k] = p[k];
        }
        int cur = 0;
        const int n = group_clip<false>(pos, nullptr, src >= 0 ? 3 : 0, 0, cur);
        if (li < n) pos[cur][li] = divide_and_snap(pos[cur][li], D.vp);
        wave_sync();
        uint32_t nrec = 0, npair = 0;
        if (li <= n - 3)
        {
            const float4 a = pos[cur][0], b = pos[cur][li + 1], c = pos[cur][li + 2];
            if (bin_accept(a, b, c))
            {
                const TileRect r = tile_rect(a, b, c, P.W, P.H, P.tilesX, P.tilesY, P.tileShiftX, P.tileShiftY);
                nrec = 1;
                npair = tile_count(r);
            }
        }
#pragma unroll
        for (int o = kClipGroup / 2; o > 0; o >>= 1)
        {
            nrec += (uint32_t)__shfl_xor((int)nrec, o, kClipGroup);
            npair += (uint32_t)__shfl_xor((int)npair, o, kClipGroup);
        }
        wave_sync();   // the next round rewrites the buffers
        // each straddler lane takes its group's sums
#pragma unroll
        for (int g = 0; g < 64 / kClipGroup; ++g)
        {
            const int sg = __shfl(src, g * kClipGroup, 64);
            const uint32_t rg = (uint32_t)__shfl((int)nrec, g * kClipGroup, 64);
            const uint32_t pg = (uint32_t)__shfl((int)npair, g * kClipGroup, 64);
            if (sg == lane)
            {
                C.nrec = rg;
                C.npair = pg;
                C.clipped = rg > 0;
                C.needClip = false;
            }
        }
    }
}

// Record, varyings and bin entries of a point or line counted by
// count_point / count_line (SL_PointProcessor::push_bin, SL_PointProcessor.cpp:
// 26-58; SL_LineProcessor::push_bin, SL_LineProcessor.cpp:26-62).
__device__ __forceinline__ void write_prim_record(const SetupParams& P, const DrawDesc& D, uint32_t recId, uint32_t pairOff,
                                                  const TriIn& T, const TriCount& C)
{
    SetupRec r;
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        r.x[k] = 0.f; r.y[k] = 0.f; r.z[k] = 0.f; r.hw[k] = 1.f;
        r.bx[k] = 0.f; r.by[k] = 0.f; r.bz[k] = 0.f;
    }
    r.bc3 = 0.f;
    r.v0y = r.v1y = r.v0x = r.v1x = 0.f;
    r.p20y = r.p20x = r.p21xy = r.p10xy = 0.f;
    r.draw = T.draw;
    r.seq = T.seq;
    r.state = pack_raster_state(D);
    float4* V = P.vars + (uint64_t)recId * P.varStride;
    const uint32_t nvMax = P.varStride / 3;
    const int nv = D.nvary;
    if (D.prim == kPrimPoint)
    {
        const int px = (int)C.sp[0].x, py = (int)C.sp[0].y;
        r.x[0] = C.sp[0].x;
        r.y[0] = C.sp[0].y;
        r.z[0] = C.sp[0].z;
        r.minY = py;
        r.maxY = py + 1;
        r.xbox = (uint32_t)px | ((uint32_t)px << 16);
#pragma unroll
        for (int i = 0; i < SL_SHADER_MAX_VARYING_VECTORS; ++i)
            if (i < nv) V[i] = T.var[0].v[i];
    }
    else
    {
        const float4 a = C.sp[0], b = C.sp[1];
        r.x[0] = a.x; r.y[0] = a.y; r.z[0] = a.z;
        r.x[1] = b.x; r.y[1] = b.y; r.z[1] = b.z;
        const float ex = b.x - a.x, ey = b.y - a.y;
        r.bx[0] = 1.f / sqrtf((ex * ex) + (ey * ey));   // inversesqrt(length_squared(sc1 - sc0)), SL_LineRasterizer.cpp:39
        const LineWalk w = make_line_walk(a.x, a.y, b.x, b.y);
        r.by[0] = __int_as_float(w.X1);
        r.by[1] = __int_as_float(w.Y1);
        r.by[2] = __int_as_float(w.dx);
        r.bz[0] = __int_as_float(w.dErr);
        r.bz[1] = __int_as_float(w.yErr);
        r.bz[2] = __int_as_float(w.steep);
        LineBox box;
        line_box(w, P.W, P.H, box);
        const bool direct = box.dx0 <= box.dx1 && box.dy0 <= box.dy1, wrap = box.wy0 <= box.wy1;
        r.minY = direct ? (wrap ? min(box.dy0, box.wy0) : box.dy0) : box.wy0;
        r.maxY = (direct ? (wrap ? max(box.dy1, box.wy1) : box.dy1) : box.wy1) + 1;
        const int xl = wrap ? 0 : box.dx0, xh = direct ? box.dx1 : 0;
        r.xbox = (uint32_t)xl | ((uint32_t)xh << 16);
        // a clipped end carries its varyings mixed toward the other end
#pragma unroll
        for (int i = 0; i < SL_SHADER_MAX_VARYING_VECTORS; ++i)
        {
            if (i < nv)
            {
                const float4 va = T.var[0].v[i], vb = T.var[1].v[i];
                V[i] = C.lineClip ? mix4(va, vb, C.t0) : va;
                V[nvMax + i] = C.lineClip ? mix4(vb, va, C.t1) : vb;
            }
        }
    }
    P.recs[recId] = r;
    uint32_t o = pairOff;
    auto put = [&](const TileRect& tr) {
        for (int ty = tr.ty0; ty <= tr.ty1; ++ty)
            for (int tx = tr.tx0; tx <= tr.tx1; ++tx)
            {
                P.pairKeys[o] = (uint32_t)(ty * P.tilesX + tx);
                P.pairVals[o] = recId;
                ++o;
            }
    };
    put(C.tr);
    if (D.prim == kPrimLine) put(C.tr2);
}

// The vertex pass: one workgroup per work item of kVertexItem (instance,
// vertex) entries of the batch's qualifying draws; each thread evaluates
// kVertexPerThread entries (all their loads before any store, so the fetches
// overlap) and writes each position once.  (A persistent grid claiming items
// from device-side counters spent its time in the claims: returning atomics
// on one address from every XCD serialise at the memory side, ~0.4 us each.)
__global__ __launch_bounds__(kVertexBlock) void k_vertex(SetupParams P)
{
    const uint32_t item = blockIdx.x;
    if (item >= P.vtxBlocks) return;
    uint32_t lo = 0, hi = P.numVtxDraws;
    while (hi - lo > 1)
    {
        const uint32_t mid = (lo + hi) >> 1;
        if (P.vtxDraws[mid].firstBlock <= item) lo = mid;
        else hi = mid;
    }
    const VtxDraw vd = P.vtxDraws[lo];
    // the draw descriptor, staged once in LDS and read into scalar registers
    __shared__ DrawDesc sD;
    static_assert(sizeof(DrawDesc) % 4 == 0 && sizeof(DrawDesc) / 4 <= kVertexBlock, "draw descriptor staging");
    if (threadIdx.x < sizeof(DrawDesc) / 4)
        reinterpret_cast<uint32_t*>(&sD)[threadIdx.x] = reinterpret_cast<const uint32_t*>(P.draws + vd.draw)[threadIdx.x];
    __syncthreads();
    DrawDesc D;
    {
        uint32_t w[sizeof(DrawDesc) / 4];
#pragma unroll
        for (int i = 0; i < (int)(sizeof(DrawDesc) / 4); ++i)
            w[i] = __builtin_amdgcn_readfirstlane(reinterpret_cast<const uint32_t*>(&sD)[i]);
        __builtin_memcpy(&D, w, sizeof(DrawDesc));
    }
    const uint32_t nv = D.posVerts;
    const uint64_t total = (uint64_t)nv * (D.numTris / D.trisPerInstance);
    const uint64_t base = (uint64_t)(item - vd.firstBlock) * kVertexItem;
    // the item's first entry as (instance, vertex); entries advance by
    // kVertexBlock, wrapping into the next instance
    const uint64_t inst0 = base / nv;
    const uint32_t rem0 = (uint32_t)(base - inst0 * nv);
    // The transforming programs' matrices (mvp and the model matrix of each
    // instance the item touches) are staged once in LDS: read per thread,
    // 32 broadcast loads per entry set this kernel's time.  The products are
    // the program's own (mat_mul on the same values).
    constexpr uint32_t kMaxItemInstances = 8;
    __shared__ float sMvp[16];
    __shared__ float sModel[kMaxItemInstances][16];
    const uint32_t span = (uint32_t)((rem0 + kVertexItem - 1u) / nv + 1u);
    const bool staged = span <= kMaxItemInstances &&
                        (D.vsProg == SL_PROGRAM_LAMBERT_TEXTURED || D.vsProg == SL_PROGRAM_LAMBERT_MRT ||
                         D.vsProg == SL_PROGRAM_COLOR_TEXTURED);
    const sl_std_uniforms* u = reinterpret_cast<const sl_std_uniforms*>(P.uniforms + D.uniOffset);
    if (staged)
    {
        if (threadIdx.x < 16) sMvp[threadIdx.x] = u->mvp[threadIdx.x];
        if (threadIdx.x < span)
        {
            float m[16];
            model_matrix(D, u, inst0 + threadIdx.x, m);
#pragma unroll
            for (int i = 0; i < 16; ++i) sModel[threadIdx.x][i] = m[i];
        }
    }
    __syncthreads();
    float4 pos[kVertexPerThread];
#pragma unroll
    for (int j = 0; j < (int)kVertexPerThread; ++j)
    {
        uint64_t inst = inst0;
        uint32_t vid = rem0 + threadIdx.x + (uint32_t)j * kVertexBlock;
        while (vid >= nv)
        {
            vid -= nv;
            ++inst;
        }
        const bool in = base + threadIdx.x + (uint64_t)j * kVertexBlock < total;
        pos[j] = f4(0.f, 0.f, 0.f, 0.f);
        if (staged)
        {
            if (in)
            {
                const float4 p = read_attr(D, 0, vid, 1.f);
                const float4 world = mat_mul(sModel[inst - inst0], f4(p.x, p.y, p.z, 1.f));
                pos[j] = mat_mul(sMvp, world);
            }
        }
        else
        {
            const uint32_t v[3] = {vid, vid, vid};
            const bool one[3] = {in, false, false};
            float4 t[3];
            t[0] = f4(0.f, 0.f, 0.f, 0.f);
            vertex_positions3(D, P.uniforms, v, inst, one, t);
            pos[j] = t[0];
        }
    }
#pragma unroll
    for (int j = 0; j < (int)kVertexPerThread; ++j)
    {
        const uint64_t e = base + threadIdx.x + (uint64_t)j * kVertexBlock;
        if (e < total) P.positions[D.posBase + e] = pos[j];
    }
}

// 8 waves per SIMD: the register budget that keeps every wave slot usable
// The dispatcher sends workgroup i to XCD i % 8 (each XCD has its own L2).
// Workgroup i takes item xcd_block_order(i): XCD x's k-th workgroup takes
// item k % B of block (k / B) * 8 + x, so each XCD walks runs of B
// consecutive items while the eight XCDs stay on neighbouring blocks -- the
// vertices that consecutive setup chunks share are fetched into one L2 instead
// of all eight.  The tail past the last whole group of 8 B items keeps its
// order.  A bijection on [0, n): the results do not depend on it.
__device__ __forceinline__ uint32_t xcd_block_order(uint32_t i, uint32_t n, uint32_t B)
{
    const uint32_t full = n / (8u * B) * (8u * B);
    if (i >= full) return i;
    const uint32_t x = i % 8u, k = i / 8u;
    return ((k / B) * 8u + x) * B + (k % B);
}

constexpr uint32_t kSetupXcdBlock = 128;   // setup chunks per XCD run (measured: 64-512 within 2%, 128 best)

// Seven workgroups per CU (72 VGPRs): at eight the 64-VGPR cap spilled the
// clip-counting loop (20 B per lane, 50 MB of scratch writes on the sphere
// frame); at six nothing spills but the pass was slower (0.1067 ms at seven,
// 0.1081 at six, 0.1102 at eight).
__global__ __launch_bounds__(kSetupBlock, 7) void k_setup(SetupParams P)
{
    const uint32_t chunk = xcd_block_order(blockIdx.x, gridDim.x, kSetupXcdBlock);
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    __shared__ uint32_t sDraw;
    if (threadIdx.x < 64)
    {
        // draws start on chunk boundaries, so one draw covers the whole chunk
        const uint32_t d = find_draw_wave(P.draws, P.numDraws, (uint64_t)chunk * kSetupBlock);
        if (threadIdx.x == 0) sDraw = d;
    }
    __syncthreads();
    const uint32_t drawIdx = __builtin_amdgcn_readfirstlane(sDraw);
    const uint64_t g = (uint64_t)chunk * kSetupBlock + threadIdx.x;
    // the chunk's draw descriptor, staged once in LDS for all lanes
    __shared__ DrawDesc sD;
    static_assert(sizeof(DrawDesc) % 4 == 0 && sizeof(DrawDesc) / 4 <= kSetupBlock, "draw descriptor staging");
    if (threadIdx.x < sizeof(DrawDesc) / 4)
        reinterpret_cast<uint32_t*>(&sD)[threadIdx.x] = reinterpret_cast<const uint32_t*>(P.draws + drawIdx)[threadIdx.x];
    __syncthreads();

    TriIn T;
    T.status = 0;
    T.draw = 0;
    T.seq = 0;
    // the descriptor is uniform: copy it to scalar registers once (the fields
    // that are used stay in SGPRs instead of being re-read from LDS)
    DrawDesc DR;
    {
        uint32_t w[sizeof(DrawDesc) / 4];
#pragma unroll
        for (int i = 0; i < (int)(sizeof(DrawDesc) / 4); ++i)
            w[i] = __builtin_amdgcn_readfirstlane(reinterpret_cast<const uint32_t*>(&sD)[i]);
        __builtin_memcpy(&DR, w, sizeof(DrawDesc));
    }
    const DrawDesc& DU = DR;
    const bool live = g < P.totalTris && g - DU.triBase < DU.numTris;
    if (live) load_triangle_ids(g, DU, drawIdx, T);
    if (DU.posBase != kNoPositions)
    {
        // positions computed once per (instance, vertex) by the vertex pass
        if (live)
        {
            const bool all3[3] = {true, true, true};
            positions_from_pass(P, DU, T.vid, T.inst, all3, T.pos);
        }
    }
    else if (DU.indexed)
    {
        // Per-chunk vertex cache: a run of kSetupBlock triangles of an indexed
        // mesh references far fewer distinct vertices than 3 * kSetupBlock,
        // so each distinct (vertex, instance) position is computed once, by
        // the lane that claims its slot, and shared through LDS.  The vertex
        // programs are pure, so the positions are the ones each triangle
        // would compute; a lane that finds no slot computes its own.
        __shared__ unsigned long long sKey[kVertexSlots];
        __shared__ float4 sPos[kVertexSlots];
        for (int i = threadIdx.x; i < kVertexSlots; i += kSetupBlock) sKey[i] = kNoVertex;
        __syncthreads();
        int slot[3] = {-1, -1, -1};
        bool own[3] = {false, false, false};
        if (live)
        {
#pragma unroll
            for (int k = 0; k < 3; ++k)
            {
                const unsigned long long key = (unsigned long long)T.vid[k] | ((unsigned long long)T.inst << 32);
                uint32_t h = (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> (64 - kVertexSlotBits));
                for (int probe = 0; probe < 8; ++probe)
                {
                    const unsigned long long prev = atomicCAS(&sKey[h], kNoVertex, key);
                    if (prev == kNoVertex || prev == key)
                    {
                        slot[k] = (int)h;
                        own[k] = prev == kNoVertex;
                        break;
                    }
                    h = (h + 1u) & (kVertexSlots - 1u);
                }
            }
            // owned slots and (table crowded) unslotted vertices
            const bool need[3] = {own[0] || slot[0] < 0, own[1] || slot[1] < 0, own[2] || slot[2] < 0};
            vertex_positions3(DU, P.uniforms, T.vid, T.inst, need, T.pos);
#pragma unroll
            for (int k = 0; k < 3; ++k)
                if (own[k]) sPos[slot[k]] = T.pos[k];
        }
        __syncthreads();
        if (live)
        {
#pragma unroll
            for (int k = 0; k < 3; ++k)
                if (slot[k] >= 0) T.pos[k] = sPos[slot[k]];
        }
    }
    else if (live)
    {
        const uint32_t per = verts_per_prim(DU.prim);
        const bool need[3] = {true, per > 1, per > 2};
        vertex_positions3(DU, P.uniforms, T.vid, T.inst, need, T.pos);
    }
    TriCount C;
    C.nrec = 0;
    C.npair = 0;
    C.clipped = false;
    C.empty = false;
    C.needClip = false;
    if (live)
    {
        classify_triangle(DU, T);
        count_triangle(P, DU, drawIdx, T, C);
    }
    __shared__ float4 sClip[kSetupBlock / 64][64 / kClipGroup][2][kClipVerts];   // each wave's clipper buffers
    count_clipped(P, DU, T, C, sClip[wave]);
    // chunk summary: exclusive scan of the packed counts (each survivor's
    // offset within the chunk), survivor ranks, staging words, totals
    const unsigned long long mine = ((unsigned long long)C.npair << kRecBits) | C.nrec;
    const unsigned long long incl = wave_incl_scan(mine);
    const unsigned long long survivors = __ballot(C.nrec > 0);
    const unsigned long long empties = __ballot(C.empty);
    __shared__ unsigned long long sSum[kSetupBlock / 64];
    __shared__ uint32_t sSurv[kSetupBlock / 64];
    __shared__ uint32_t sEmpty[kSetupBlock / 64];
    if (lane == 63) sSum[wave] = incl;
    if (lane == 0)
    {
        sSurv[wave] = (uint32_t)__popcll(survivors);
        sEmpty[wave] = (uint32_t)__popcll(empties);
    }
    __syncthreads();
    unsigned long long before = incl - mine, tot = 0;
    uint32_t rank = (uint32_t)__popcll(survivors & ((1ull << lane) - 1ull)), surv = 0, emp = 0;
#pragma unroll
    for (int w = 0; w < kSetupBlock / 64; ++w)
    {
        if (w < wave)
        {
            before += sSum[w];
            rank += sSurv[w];
        }
        tot += sSum[w];
        surv += sSurv[w];
        emp += sEmpty[w];
    }
    const unsigned long long recBefore = before & ((1ull << kRecBits) - 1), pairBefore = before >> kRecBits;
    if (C.nrec > 0) P.stage[(uint64_t)chunk * kSetupBlock + rank] = (unsigned long long)threadIdx.x | (recBefore << 8) | (pairBefore << 20);
    // A straddling triangle's clip job, reserved here (one atomic per wave);
    // its offsets are within the chunk, k_clip adds the chunk's prefix.
    {
        const unsigned long long m = __ballot(C.clipped);
        if (m != 0ull)
        {
            const int leader = __ffsll((long long)m) - 1;
            uint32_t base = 0u;
            if (lane == leader) base = atomicAdd(&P.ctrl->clipJobs, (uint32_t)__popcll(m));
            base = __shfl(base, leader, 64);
            if (C.clipped)
            {
                const uint32_t job = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
                if (job < P.clipCap)
                {
                    ClipJob J;
                    J.tri = (uint32_t)g;
                    J.recOff = (uint32_t)recBefore;
                    J.pairOff = (uint32_t)pairBefore;
                    J.nrec = C.nrec;
                    J.npair = C.npair;
                    P.clipJobs[job] = J;
                }
                else
                {
                    atomicExch(&P.ctrl->overflow, 1u);
                }
            }
        }
    }
    if (threadIdx.x == 0)
    {
        ChunkInfo& I = P.chunkInfo[chunk];
        I.agg = tot;
        I.surv = surv;
        I.draw = drawIdx;
        I.empty = emp;
        if (tot != 0)
        {
            atomicAdd(&P.blockAgg[chunk / kScanBlock], tot);
            atomicAdd(&P.blockSurv[chunk / kScanBlock], surv);
        }
    }
}

// k_chunk_scan: exclusive prefixes of the chunk totals and survivor counts.
// Workgroup b owns chunks [b * kScanBlock, (b + 1) * kScanBlock): its offset is
// the sum of the earlier blocks' totals (which k_setup accumulated), the rest
// a workgroup scan.  Also records, for every block of kSetupBlock survivors,
// the chunk that holds its first survivor, and the frame totals.
__device__ __forceinline__ unsigned long long block_sum_u64(unsigned long long v, unsigned long long* sRed)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    v = wave_sum_u64(v);
    __syncthreads();
    if (lane == 0) sRed[wave] = v;
    __syncthreads();
    unsigned long long s = 0;
    for (int w = 0; w < nw; ++w) s += sRed[w];
    return s;
}

__global__ __launch_bounds__(kScanBlock) void k_chunk_scan(SetupParams P)
{
    __shared__ unsigned long long sRed[kScanBlock / 64];
    __shared__ unsigned long long sA[kScanBlock / 64];
    __shared__ uint32_t sS[kScanBlock / 64];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const uint32_t b = blockIdx.x;
    const uint32_t c = b * kScanBlock + t;
    unsigned long long pa = 0, ps = 0;
    for (uint32_t i = t; i < b; i += kScanBlock)
    {
        pa += P.blockAgg[i];
        ps += P.blockSurv[i];
    }
    const unsigned long long prefA = block_sum_u64(pa, sRed);
    const unsigned long long prefS = block_sum_u64(ps, sRed);
    unsigned long long a = 0;
    uint32_t sv = 0, em = 0;
    if (c < P.numChunks)
    {
        const ChunkInfo& I = P.chunkInfo[c];
        a = I.agg;
        sv = I.surv;
        em = I.empty;
    }
    const unsigned long long emp = block_sum_u64(em, sRed);
    const unsigned long long ia = wave_incl_scan(a);
    const uint32_t is = wave_incl_scan_u32(sv);
    if (lane == 63)
    {
        sA[wave] = ia;
        sS[wave] = is;
    }
    __syncthreads();
    unsigned long long ea = prefA + ia - a, totA = prefA;
    uint32_t es = (uint32_t)prefS + is - sv, totS = (uint32_t)prefS;
    for (int w = 0; w < kScanBlock / 64; ++w)
    {
        if (w < wave)
        {
            ea += sA[w];
            es += sS[w];
        }
        totA += sA[w];
        totS += sS[w];
    }
    if (c < P.numChunks)
    {
        ChunkInfo& I = P.chunkInfo[c];
        I.pre = ea;
        I.survPre = es;
    }
    if (t == 0)
    {
        if (emp != 0) atomicAdd(&P.ctrl->emptyRecs, (uint32_t)emp);
        if (b == gridDim.x - 1)
        {
            P.ctrl->totalRecs = (uint32_t)(totA & ((1ull << kRecBits) - 1));
            P.ctrl->totalPairs = (uint32_t)(totA >> kRecBits);
            P.ctrl->survivors = totS;
        }
    }
}

// The chunk of every survivor (in submission order), one wave per chunk:
// the emitting pass reads its lanes' chunks directly, where runs of chunks
// without survivors (culled or off-screen meshes) would otherwise have to be
// walked.
constexpr int kSurvMapChunks = 4;   // chunks per workgroup, one per wave

__global__ __launch_bounds__(64 * kSurvMapChunks) void k_surv_map(SetupParams P)
{
    const uint32_t c = blockIdx.x * kSurvMapChunks + (threadIdx.x >> 6);
    if (c >= P.numChunks) return;
    const ChunkInfo& I = P.chunkInfo[c];
    const uint32_t pre = I.survPre, n = I.surv;
    for (uint32_t k = threadIdx.x & 63; k < n; k += 64) P.survChunk[pre + k] = c;
}

// k_emit: one workgroup per block of kEmitBlock survivors (grid-stride).  A
// lane finds its survivor's chunk by searching the chunks' survivor prefixes
// (kEmitBlock chunks per window in LDS, starting from the block's first
// chunk), reads the staging word for its lane and offsets within the chunk,
// re-fetches and re-classifies the triangle (the vertex programs are pure, so
// this gives the counting pass's bits) and writes its record, varyings and
// bin entries at the chunk's prefix plus those offsets.  Straddling triangles
// become clip jobs with their slots reserved.
constexpr int kEmitBlock = 256;
static_assert(kEmitBlock == kSetupBlock, "a block of survivors spans whole staging rows");

//
// PRIMS selects the survivors a launch writes: false, those of triangle draws
// (filled or wireframe); true, those of point and line draws.  A batch with
// both kinds runs both launches; each writes its own records at the offsets
// the chunk scan gave them.
template <bool PRIMS>
__global__ __launch_bounds__(kEmitBlock, 4) void k_emit(SetupParams P)
{
    constexpr int kKinds = PRIMS ? 2 : 1;
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = t >> 6;
    const uint32_t S = P.ctrl->survivors;
    for (uint32_t k = blockIdx.x; (uint64_t)k * kEmitBlock < S; k += gridDim.x)
    {
        const uint32_t s = k * kEmitBlock + t;
        const bool act = s < S;
        // the survivor's chunk (the chunk scan wrote it per survivor)
        const uint32_t chunk = act ? P.survChunk[s] : 0u;
        bool bad = false;
        if (chunk >= P.numChunks)
        {
            atomicExch(&P.ctrl->error, 1u);
            bad = true;        // the lane goes on without a triangle, with the rest of its wave
        }
        TriIn T;
        T.status = 0;
        T.draw = 0;
        T.seq = 0;
        TriCount C;
        C.nrec = 0;
        C.npair = 0;
        C.clipped = false;
        C.needClip = false;
        uint64_t g = 0;
        uint32_t di = 0;
        unsigned long long off = 0;
        bool act2 = act && !bad;
        if (act2)
        {
            const ChunkInfo& I = P.chunkInfo[chunk];
            const uint32_t r = s - I.survPre;
            act2 = r < I.surv;
            if (!act2) atomicExch(&P.ctrl->error, 1u);
        }
        if (act2)
        {
            const ChunkInfo& I = P.chunkInfo[chunk];
            const unsigned long long st = P.stage[(uint64_t)chunk * kSetupBlock + (s - I.survPre)];
            g = (uint64_t)chunk * kSetupBlock + (st & 0xFFu);
            di = I.draw;
            off = I.pre + ((st >> 8) & 0xFFFu) + ((st >> 20) << kRecBits);
            const uint32_t pk = P.draws[di].prim;
            if ((pk == kPrimPoint || pk == kPrimLine) != PRIMS) act2 = false;   // the other launch's survivor
        }
        // When the wave's survivors share one draw (the usual case) its
        // descriptor goes to scalar registers: uniform branches, scalar loads.
        const uint32_t dFirst = __builtin_amdgcn_readfirstlane(di);
        const bool uni = __ballot(act2 && di != dFirst) == 0ull;
        DrawDesc DR;
        if (uni)
        {
            constexpr int kDescWords = (int)(sizeof(DrawDesc) / 4);
            static_assert(sizeof(DrawDesc) % 4 == 0 && kDescWords <= 192, "draw descriptor words");
            const uint32_t* src = reinterpret_cast<const uint32_t*>(P.draws + dFirst);
            uint32_t v[3];
#pragma unroll
            for (int j = 0; j < 3; ++j) v[j] = (lane + 64 * j < kDescWords) ? src[lane + 64 * j] : 0u;
            uint32_t wds[kDescWords];
#pragma unroll
            for (int j = 0; j < kDescWords; ++j) wds[j] = __builtin_amdgcn_readlane(v[j / 64], j % 64);
            __builtin_memcpy(&DR, wds, sizeof(DrawDesc));
            if (act2)
            {
                load_triangle_pos<kKinds>(P, g, DR, di, T);
                count_triangle<kKinds>(P, DR, di, T, C);
            }
        }
        else if (act2)
        {
            const DrawDesc& D = P.draws[di];
            load_triangle_pos<kKinds>(P, g, D, di, T);
            count_triangle<kKinds>(P, D, di, T, C);
        }
        // a straddling triangle is k_clip's (its job was reserved by the
        // counting pass); any other survivor yields its record here
        bool ok = act2 && !C.needClip;
        if (ok && C.nrec == 0)
        {
            atomicExch(&P.ctrl->error, 1u);
            ok = false;
        }
        const uint32_t recOff = (uint32_t)(off & ((1ull << kRecBits) - 1));
        const uint64_t pairOff = off >> kRecBits;
        if (ok && ((uint64_t)recOff + C.nrec > P.recCap || pairOff + C.npair > P.pairCap))
        {
            atomicExch(&P.ctrl->overflow, 1u);
            ok = false;
        }
        if (ok)
        {
            if (uni)
            {
                load_varyings(P, DR, T);
                if (PRIMS)
                    write_prim_record(P, DR, recOff, (uint32_t)pairOff, T, C);
                else
                    write_record<false>(P, recOff, (uint32_t)pairOff, C.sp[0], C.sp[1], C.sp[2], T.var[0], T.var[1], T.var[2],
                                        DR.nvary, T.draw, T.seq, pack_raster_state(DR), C.tr);
            }
            else
            {
                const DrawDesc& D = P.draws[di];
                load_varyings(P, D, T);
                if (PRIMS)
                    write_prim_record(P, D, recOff, (uint32_t)pairOff, T, C);
                else
                    write_record<false>(P, recOff, (uint32_t)pairOff, C.sp[0], C.sp[1], C.sp[2], T.var[0], T.var[1], T.var[2],
                                        D.nvary, T.draw, T.seq, pack_raster_state(D), C.tr);
            }
        }
        // the records' bin entries, by the whole wave (every lane reaches this)
        if (!PRIMS) write_pairs_wave(P, ok, recOff, (uint32_t)pairOff, C.tr);
    }
}

// ---------------------------------------------------------------------------
// k_clip: finish the clipped triangles into their reserved slots
// ---------------------------------------------------------------------------
// One 16-lane group per job, four jobs per one-wave workgroup.  Lane k of a
// group holds polygon vertex k: at each plane it tests its edge from vertex
// k-1 (the last vertex for k = 0) and writes the edge's intersection and its
// own vertex, when kept, to the next polygon at the position the group's
// ballots give it, so the output order and every operation on a vertex are
// those of the serial polygon loop (SL_TriProcessor.cpp:370-499).  The fan's triangles are then
// set up one per lane, their record and bin-entry offsets taken from
// suffix sums over the group (the reverse fan numbers them from the last).
constexpr int kClipBlock = 64;
constexpr int kClipGroups = kClipBlock / kClipGroup;

__global__ __launch_bounds__(kClipBlock) void k_clip(SetupParams P)
{
    __shared__ float4 sPos[kClipGroups][2][kClipVerts];
    __shared__ Vary4 sVar[kClipGroups][2][kClipVerts];
    const int lane = (int)threadIdx.x;   // the workgroup is one wave
    const int gi = lane / kClipGroup, li = lane % kClipGroup;
    const unsigned long long gmask = 0xFFFFull << (gi * kClipGroup);
    const unsigned long long below = gmask & ((1ull << lane) - 1ull);
    const unsigned long long above = gmask & ~below & ~(1ull << lane);
    const uint32_t njobs = min(P.ctrl->clipJobs, P.clipCap);
    // the loop bound is uniform over the workgroup: every lane reaches the barriers
    for (uint32_t base = blockIdx.x * kClipGroups; base < njobs; base += gridDim.x * kClipGroups)
    {
        const uint32_t j = base + (uint32_t)gi;
        bool live = j < njobs;
        ClipJob J = {};
        uint64_t recOff = 0, pairOff = 0;
        if (live)
        {
            J = P.clipJobs[j];
            // the job's output range: the chunk's prefix (chunk scan) plus the
            // offsets within the chunk the counting pass reserved
            const unsigned long long off = P.chunkInfo[J.tri / kSetupBlock].pre + J.recOff +
                                           ((unsigned long long)J.pairOff << kRecBits);
            recOff = off & ((1ull << kRecBits) - 1);
            pairOff = off >> kRecBits;
            if (recOff + J.nrec > P.recCap || pairOff + J.npair > P.pairCap)
            {
                if (li == 0) atomicExch(&P.ctrl->overflow, 1u);   // the host grows the buffers and runs setup again
                live = false;
            }
        }
        TriIn T;
        int nv = 0;
        if (live)
        {
            // the fetch and vertex programs again; a chunk's triangles are
            // all of one draw
            const uint32_t di = P.chunkInfo[J.tri / kSetupBlock].draw;
            const DrawDesc& D = P.draws[di];
            load_triangle_pos<1>(P, J.tri, D, di, T);
            load_varyings(P, D, T);
            nv = D.nvary;
            if (li == 0)   // static indices keep the triangle in registers
            {
#pragma unroll
                for (int k = 0; k < 3; ++k)
                {
                    sPos[gi][0][k] = T.pos[k];
#pragma unroll
                    for (int i = 0; i < SL_SHADER_MAX_VARYING_VECTORS; ++i)
                        if (i < nv) sVar[gi][0][k].v[i] = T.var[k].v[i];
                }
            }
        }
        int cur = 0;
        const int n = group_clip<true>(sPos[gi], sVar[gi], live ? 3 : 0, nv, cur);
        const DrawDesc* D = live ? &P.draws[T.draw] : nullptr;
        if (li < n) sPos[gi][cur][li] = divide_and_snap(sPos[gi][cur][li], D->vp);
        __syncthreads();
        // fan triangle li: vertices 0, li + 1, li + 2
        bool acc = false;
        TileRect r = TileRect{0, -1, 0, -1};
        float4 p0 = f4(0.f, 0.f, 0.f, 0.f), p1 = p0, p2 = p0;
        if (li <= n - 3)
        {
            p0 = sPos[gi][cur][0];
            p1 = sPos[gi][cur][li + 1];
            p2 = sPos[gi][cur][li + 2];
            acc = bin_accept(p0, p1, p2);
            if (acc) r = tile_rect(p0, p1, p2, P.W, P.H, P.tilesX, P.tilesY, P.tileShiftX, P.tileShiftY);
        }
        // records are numbered from the last fan triangle down, at most the
        // count the counting pass reserved
        const uint32_t rank = (uint32_t)__popcll(__ballot(acc) & above);
        acc = acc && rank < J.nrec;
        const uint32_t cnt = acc ? tile_count(r) : 0u;
        uint32_t suf = cnt;   // inclusive sum over this lane and the group's later lanes
#pragma unroll
        for (int o = 1; o < kClipGroup; o <<= 1)
        {
            const uint32_t v = (uint32_t)__shfl_down((int)suf, o, kClipGroup);
            if (li + o < kClipGroup) suf += v;
        }
        if (acc)
        {
            write_record(P, (uint32_t)recOff + rank, (uint32_t)pairOff + (suf - cnt), p0, p1, p2, sVar[gi][cur][0],
                         sVar[gi][cur][li + 1], sVar[gi][cur][li + 2], nv, T.draw, T.seq, pack_raster_state(*D), r);
        }
        __syncthreads();   // the next jobs overwrite the polygon buffers
    }
}

// ---------------------------------------------------------------------------
// Ordered binning: a stable counting sort of the (tile, record) entries, which
// k_setup/k_clip emit in record order, into per-tile lists that keep that
// order.  Entries are split into at most kMaxBinBlocks blocks; k_bin_count
// histograms each block per tile, k_bin_colscan turns every tile's column of
// block counts into block offsets and tile totals, k_bin_tilescan scans the
// totals into tile ranges, and k_bin_scatter (one wave per block) places each
// block's entries 64 at a time: a bitonic sort of (tile, lane) across the wave
// groups equal tiles in entry order, and the head of each group reserves its
// slots with one LDS atomic.  Entry counts are read from the device, so no
// host round trip is needed between setup and raster.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t bin_entries(const FrameCtrl* ctrl, uint32_t cap)
{
    return min(ctrl->totalPairs, cap);
}

// Block b of the binning kernels owns entries [b*per, (b+1)*per).
__global__ __launch_bounds__(256) void k_bin_count(const uint32_t* __restrict__ keys, const FrameCtrl* __restrict__ ctrl,
                                                   uint32_t cap, uint32_t per, uint32_t* __restrict__ counts, int numTiles)
{
    extern __shared__ uint32_t sCnt[];
    const uint32_t n = bin_entries(ctrl, cap);
    const uint32_t start = blockIdx.x * per;
    if (start >= n) return;
    const uint32_t end = min(start + per, n);
    for (int t = threadIdx.x; t < numTiles; t += blockDim.x) sCnt[t] = 0u;
    __syncthreads();
    for (uint32_t i = start + threadIdx.x; i < end; i += blockDim.x)
    {
        const uint32_t k = keys[i];
        if (k < (uint32_t)numTiles) atomicAdd(&sCnt[k], 1u);
    }
    __syncthreads();
    uint32_t* row = counts + (uint64_t)blockIdx.x * numTiles;
    for (int t = threadIdx.x; t < numTiles; t += blockDim.x) row[t] = sCnt[t];
}

// Column scan of the block x tile count matrix: 32 tiles per workgroup, each
// column split into 8 parts summed independently, then scanned in order.
constexpr int kColTiles = 32, kColParts = 8;

__global__ __launch_bounds__(kColTiles * kColParts) void k_bin_colscan(const FrameCtrl* __restrict__ ctrl, uint32_t cap, uint32_t per,
                                                                      uint32_t* __restrict__ counts, uint32_t* __restrict__ totals,
                                                                      int numTiles)
{
    __shared__ uint32_t sPart[kColParts][kColTiles];
    const int tl = threadIdx.x % kColTiles, part = threadIdx.x / kColTiles;
    const int t = blockIdx.x * kColTiles + tl;
    const uint32_t n = bin_entries(ctrl, cap);
    const uint32_t blocks = (n + per - 1) / per;
    const uint32_t rpp = (blocks + kColParts - 1) / kColParts;
    const uint32_t r0 = min(blocks, part * rpp), r1 = min(blocks, r0 + rpp);
    uint32_t sum = 0;
    if (t < numTiles)
    {
#pragma unroll 8
        for (uint32_t b = r0; b < r1; ++b) sum += counts[(uint64_t)b * numTiles + t];
    }
    sPart[part][tl] = sum;
    __syncthreads();
    uint32_t run = 0;
    for (int q = 0; q < part; ++q) run += sPart[q][tl];
    if (t < numTiles)
    {
#pragma unroll 8
        for (uint32_t b = r0; b < r1; ++b)
        {
            uint32_t* c = counts + (uint64_t)b * numTiles + t;
            const uint32_t v = *c;
            *c = run;
            run += v;
        }
        if (part == kColParts - 1) totals[t] = run;
    }
}

// Exclusive prefixes of two values over a 1024-thread workgroup, and their
// totals.
__device__ __forceinline__ void block_excl_scan2(uint32_t& a, uint32_t& b, uint32_t& totA, uint32_t& totB, uint2* sWave)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t ia = a, ib = b;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1)
    {
        const uint32_t oa = __shfl_up(ia, off, 64), ob = __shfl_up(ib, off, 64);
        if (lane >= off) { ia += oa; ib += ob; }
    }
    if (lane == 63) sWave[wave] = make_uint2(ia, ib);
    __syncthreads();
    uint32_t ba = 0, bb = 0;
    totA = totB = 0;
    for (int w = 0; w < 16; ++w)
    {
        const uint2 v = sWave[w];
        if (w < wave) { ba += v.x; bb += v.y; }
        totA += v.x;
        totB += v.y;
    }
    a = ba + ia - a;
    b = bb + ib - b;
    __syncthreads();
}

// Tile ranges (exclusive scan of the tile totals), the raster's launch order
// (heaviest tiles first: tiles bucketed by floor(log2(records)) + 1, empty
// tiles last; the order inside a bucket only affects scheduling) and, with
// a split plan (P.items), the raster's work items: a tile of more than
// splitAbove records is cut into ceil(records / partRecs) parts (at most kMaxTileParts)
// of equal length while the maxParts part slots last (in tile order); the
// parts come first, then every tile in launch order, a split tile's entry
// marked as nothing to do (parts == 0).  One workgroup; each thread owns up
// to kTilesPerThread consecutive tiles.
constexpr int kTilesPerThread = 16;   // 16384 tiles at most
struct SplitPlan
{
    RasterItem* items;         // null: no work items
    uint32_t* itemCount;
    uint32_t* partDone;
    uint32_t partRecs, splitAbove, maxParts;
};

__global__ __launch_bounds__(1024) void k_bin_tilescan(const uint32_t* __restrict__ totals, uint2* __restrict__ range,
                                                       uint32_t* __restrict__ order, int numTiles, SplitPlan SP)
{
    __shared__ uint2 sWave[16];
    __shared__ uint32_t sCnt[33];
    __shared__ uint32_t sBucket[33];
    if (threadIdx.x < 33) sCnt[threadIdx.x] = 0u;
    const int per = (numTiles + 1023) / 1024;
    const int t0 = (int)threadIdx.x * per;
    uint32_t tot[kTilesPerThread];
#pragma unroll
    for (int j = 0; j < kTilesPerThread; ++j) tot[j] = (j < per && t0 + j < numTiles) ? totals[t0 + j] : 0u;
    auto wanted = [&](uint32_t n) -> uint32_t {
        return (n > SP.splitAbove) ? min((n + SP.partRecs - 1u) / SP.partRecs, kMaxTileParts) : 1u;
    };
    uint32_t se = 0, sw = 0;
#pragma unroll
    for (int j = 0; j < kTilesPerThread; ++j)
    {
        se += tot[j];
        if (SP.items != nullptr && j < per)
        {
            const uint32_t p = wanted(tot[j]);
            sw += (p > 1u) ? p : 0u;
        }
    }
    uint32_t totE, totW;
    block_excl_scan2(se, sw, totE, totW, sWave);
    // ranges and the launch-order buckets: a thread's consecutive tiles
    // mostly share a bucket, so each run of equal buckets is counted with one
    // LDS atomic (same-address atomics serialise)
    uint32_t run = se, runCount = 0;
#pragma unroll
    for (int j = 0; j < kTilesPerThread; ++j)
    {
        if (j < per && t0 + j < numTiles)
        {
            range[t0 + j] = make_uint2(run, run + tot[j]);
            run += tot[j];
            const int b = 32 - __clz((int)tot[j]);
            ++runCount;
            const bool last = !(j + 1 < kTilesPerThread && j + 1 < per && t0 + j + 1 < numTiles) ||
                              32 - __clz((int)tot[j + 1 < kTilesPerThread ? j + 1 : j]) != b;
            if (last)
            {
                atomicAdd(&sCnt[b], runCount);
                runCount = 0;
            }
        }
    }
    // split decisions: accepted while the parts wanted so far (in tile order) fit
    uint32_t splitMask = 0, nParts = 0;
    if (SP.items != nullptr)
    {
        uint32_t rw = sw, sa = 0, sc = 0;
#pragma unroll
        for (int j = 0; j < kTilesPerThread; ++j)
        {
            if (j < per && t0 + j < numTiles)
            {
                const uint32_t p = wanted(tot[j]);
                if (p > 1u)
                {
                    if (rw + p <= SP.maxParts)
                    {
                        splitMask |= 1u << j;
                        sa += p;
                        sc += 1u;
                    }
                    rw += p;
                }
            }
        }
        uint32_t totA, totC;
        block_excl_scan2(sa, sc, totA, totC, sWave);
        nParts = totA;
        uint32_t ra = sa, rc = sc, rb = se;
#pragma unroll
        for (int j = 0; j < kTilesPerThread; ++j)
        {
            if (j < per && t0 + j < numTiles && ((splitMask >> j) & 1u))
            {
                const uint32_t n = tot[j], p = wanted(n);
                SP.partDone[rc] = 0u;
                for (uint32_t k = 0; k < p; ++k)
                {
                    RasterItem it;
                    it.tile = (uint32_t)(t0 + j);
                    it.begin = rb + (uint32_t)(((uint64_t)n * k) / p);
                    it.end = rb + (uint32_t)(((uint64_t)n * (k + 1)) / p);
                    it.slot = ra + k;
                    it.part = k;
                    it.parts = p;
                    it.done = rc;
                    it.pad = 0u;
                    SP.items[ra + k] = it;
                }
                ra += p;
                rc += 1u;
            }
            if (j < per) rb += tot[j];
        }
    }
    __syncthreads();
    if (threadIdx.x < 64)
    {
        // bucket b's first launch slot: the tiles of the buckets above it
        const int l = (int)threadIdx.x;
        const uint32_t c = (l <= 32) ? sCnt[32 - l] : 0u;
        uint32_t inc = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1)
        {
            const uint32_t v = (uint32_t)__shfl_up((int)inc, o, 64);
            if (l >= o) inc += v;
        }
        if (l <= 32) sBucket[32 - l] = inc - c;
    }
    if (threadIdx.x == 0)
    {
        if (SP.items != nullptr)
        {
            SP.itemCount[0] = nParts + (uint32_t)numTiles;
            SP.itemCount[1] = nParts;
        }
    }
    __syncthreads();
    uint32_t rb = se;
#pragma unroll
    for (int j = 0; j < kTilesPerThread; ++j)
    {
        if (j < per && t0 + j < numTiles)
        {
            const uint32_t pos = atomicAdd(&sBucket[32 - __clz((int)tot[j])], 1u);
            order[pos] = (uint32_t)(t0 + j);
            if (SP.items != nullptr)
            {
                RasterItem it;
                it.tile = (uint32_t)(t0 + j);
                it.begin = rb;
                it.end = rb + tot[j];
                it.slot = 0u;
                it.part = 0u;
                it.parts = ((splitMask >> j) & 1u) ? 0u : 1u;   // a split tile's parts come first
                it.done = 0u;
                it.pad = 0u;
                SP.items[nParts + pos] = it;
            }
            rb += tot[j];
        }
    }
}

__global__ __launch_bounds__(64) void k_bin_scatter(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                                                    const FrameCtrl* __restrict__ ctrl, uint32_t cap, uint32_t per,
                                                    const uint32_t* __restrict__ counts, const uint2* __restrict__ range,
                                                    uint32_t* __restrict__ out, int numTiles)
{
    extern __shared__ uint32_t sBase[];
    const uint32_t n = bin_entries(ctrl, cap);
    const uint32_t start = blockIdx.x * per;
    if (start >= n) return;
    const uint32_t end = min(start + per, n);
    const uint32_t* row = counts + (uint64_t)blockIdx.x * numTiles;
    // the block's base slot per tile; 32 tiles' loads in flight per lane (the
    // pass is a chain of load round trips: fewer, wider rounds)
    constexpr int kInit = 32;
    for (int t0 = 0; t0 < numTiles; t0 += 64 * kInit)
    {
        uint32_t rx[kInit], rw[kInit];
#pragma unroll
        for (int j = 0; j < kInit; ++j)
        {
            const int t = t0 + j * 64 + (int)threadIdx.x;
            rx[j] = (t < numTiles) ? range[t].x : 0u;
            rw[j] = (t < numTiles) ? row[t] : 0u;
        }
#pragma unroll
        for (int j = 0; j < kInit; ++j)
        {
            const int t = t0 + j * 64 + (int)threadIdx.x;
            if (t < numTiles) sBase[t] = rx[j] + rw[j];
        }
    }
    __syncthreads();
    const int lane = threadIdx.x;
    const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);   // lanes <= lane
    // Entries are placed 64 at a time, in order.  A step sorts kAhead chunks
    // and issues the loads of the next kAhead chunks first.
    constexpr int kAhead = 4;
    uint32_t pk[kAhead], pv[kAhead];
#pragma unroll
    for (int a = 0; a < kAhead; ++a)
    {
        const uint32_t i = start + (uint32_t)(a * 64 + lane);
        pk[a] = (i < end) ? keys[i] : 0xFFFFFFFFu;
        pv[a] = (i < end) ? vals[i] : 0u;
    }
    for (uint32_t c0 = start; c0 < end; c0 += 64 * kAhead)
    {
        // the kAhead chunks are sorted together (independent shuffle chains
        // interleave) and placed one after the other, in order
        uint32_t sv[kAhead], val[kAhead];
#pragma unroll
        for (int a = 0; a < kAhead; ++a)
        {
            const uint32_t c = c0 + (uint32_t)(a * 64);
            const uint32_t key = pk[a], rawVal = pv[a];
            const uint32_t inext = c + (uint32_t)(kAhead * 64 + lane);
            pk[a] = (inext < end) ? keys[inext] : 0xFFFFFFFFu;
            pv[a] = (inext < end) ? vals[inext] : 0u;
            const bool valid = key < (uint32_t)numTiles;
            val[a] = valid ? rawVal : 0u;
            // bitonic sort of (tile << 6 | lane) across the wave: equal tiles
            // end up adjacent and in entry order; invalid entries sort last
            sv[a] = ((valid ? key : (uint32_t)numTiles) << 6) | (uint32_t)lane;
        }
#pragma unroll
        for (int k = 2; k <= 64; k <<= 1)
        {
#pragma unroll
            for (int j = k >> 1; j > 0; j >>= 1)
            {
                const bool up = (lane & k) == 0;
                const bool lower = (lane & j) == 0;
#pragma unroll
                for (int a = 0; a < kAhead; ++a)
                {
                    const uint32_t o = __shfl_xor(sv[a], j, 64);
                    sv[a] = (lower == up) ? min(sv[a], o) : max(sv[a], o);
                }
            }
        }
#pragma unroll
        for (int a = 0; a < kAhead; ++a)
        {
            if (c0 + (uint32_t)(a * 64) >= end) break;
            const uint32_t skey = sv[a] >> 6, src = sv[a] & 63u;
            const uint32_t prev = __shfl_up(skey, 1, 64);
            const bool head = lane == 0 || prev != skey;
            const unsigned long long heads = __ballot(head);
            const int headLane = 63 - __clzll((long long)(heads & upto));
            // the head of each run of equal tiles reserves the run's slots; runs
            // have distinct tiles, so the LDS atomics never collide, and chunks
            // are taken in order by this one wave
            uint32_t slot = 0;
            if (head && skey < (uint32_t)numTiles)
            {
                const unsigned long long after = (lane == 63) ? 0ull : (heads >> (lane + 1));
                const uint32_t cnt = after ? (uint32_t)__ffsll((long long)after) : (uint32_t)(64 - lane);
                slot = atomicAdd(&sBase[skey], cnt);
            }
            const uint32_t b = __shfl(slot, headLane, 64);
            const uint32_t v = __shfl(val[a], (int)src, 64);
            if (skey < (uint32_t)numTiles) out[b + (uint32_t)(lane - headLane)] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// k_blit: nearest-neighbour blit, one destination pixel per lane
// (SL_BlitProcessor::blit_nearest, SL_BlitProcessor.cpp:620-665; colour
// operators :289-416)
// ---------------------------------------------------------------------------
template <int SRC, int DST>
__global__ __launch_bounds__(256) void k_blit(BlitParams P)
{
    const uint32_t w = P.x1 - P.x0;
    const uint64_t n = (uint64_t)w * (P.y1 - P.y0);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    {
        const uint32_t x = P.x0 + (uint32_t)(i % w), y = P.y0 + (uint32_t)(i / w);
        const unsigned long long yf = ((unsigned long long)y * P.foutH) >> 16;
        const unsigned long long srcY = (unsigned long long)P.srcY1 - ((unsigned long long)P.srcY0 + yf) - 1ull;
        const unsigned long long srcX = ((unsigned long long)x * P.foutW) >> 16;
        if (srcX >= P.srcW || srcY >= P.srcH) continue;      // the reference reads outside the source here
        const uint64_t si = srcX + (uint64_t)P.srcW * srcY;
        const uint64_t di = (uint64_t)x + (uint64_t)P.dstW * y;
        if (SRC == SL_COLOR_RGBA_8U && DST == SL_COLOR_RGBA_8U)
        {
            reinterpret_cast<uint32_t*>(P.dst)[di] = reinterpret_cast<const uint32_t*>(P.src)[si];
        }
        else if (SRC == SL_COLOR_RGBA_FLOAT && DST == SL_COLOR_RGBA_FLOAT)
        {
            reinterpret_cast<float4*>(P.dst)[di] = reinterpret_cast<const float4*>(P.src)[si];
        }
        else
        {
            float c[4];
            if (SRC == SL_COLOR_RGBA_8U)
            {
                unpack_rgba8(reinterpret_cast<const uint32_t*>(P.src)[si], c);   // color_cast<float, uint8_t>
            }
            else if (SRC == SL_COLOR_RGBA_FLOAT)
            {
                const float4 v = reinterpret_cast<const float4*>(P.src)[si];
                c[0] = v.x; c[1] = v.y; c[2] = v.z; c[3] = v.w;
            }
            else
            {
                // R -> RGBA: (min, min, r, max) of the source type
                c[0] = 0.f; c[1] = 0.f; c[2] = reinterpret_cast<const float*>(P.src)[si]; c[3] = 1.f;
            }
            if (DST == SL_COLOR_RGBA_8U) reinterpret_cast<uint32_t*>(P.dst)[di] = pack_rgba8(c);
            else reinterpret_cast<float4*>(P.dst)[di] = make_float4(c[0], c[1], c[2], c[3]);
        }
    }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
static inline int grid_for(uint64_t n, int block, int cap = 4096)
{
    uint64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > (uint64_t)cap) g = cap;
    return (int)g;
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s)
{
    (void)hipGetLastError();   // drop a stale error from an earlier call
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256)), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t launch_ctrl_reset(FrameCtrl* ctrl, hipStream_t s)
{
    (void)hipGetLastError();   // drop a stale error from an earlier call
    hipLaunchKernelGGL(k_ctrl_reset, dim3(1), dim3(256), 0, s, ctrl);
    return hipGetLastError();
}

hipError_t launch_vertex(const SetupParams& p, hipStream_t s)
{
    (void)hipGetLastError();   // drop a stale error from an earlier call
    if (p.vtxBlocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_vertex, dim3(p.vtxBlocks), dim3(kVertexBlock), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_fill_u64(unsigned long long* p, unsigned long long v, uint64_t n, hipStream_t s)
{
    (void)hipGetLastError();   // drop a stale error from an earlier call
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, 256)), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t launch_setup(const SetupParams& p, uint64_t chunks, hipStream_t s)
{
    (void)hipGetLastError();   // drop a stale error from an earlier call
    hipLaunchKernelGGL(k_setup, dim3((unsigned)chunks), dim3(kSetupBlock), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_chunk_scan(const SetupParams& p, hipStream_t s)
{
    (void)hipGetLastError();   // drop a stale error from an earlier call
    const uint32_t blocks = (p.numChunks + kScanBlock - 1) / kScanBlock;
    hipLaunchKernelGGL(k_chunk_scan, dim3(blocks), dim3(kScanBlock), 0, s, p);
    hipLaunchKernelGGL(k_surv_map, dim3((p.numChunks + kSurvMapChunks - 1) / kSurvMapChunks), dim3(64 * kSurvMapChunks), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_emit(const SetupParams& p, bool tris, bool prims, hipStream_t s)
{
    (void)hipGetLastError();   // drop a stale error from an earlier call
    const uint32_t grid = p.numChunks < 2048u ? p.numChunks : 2048u;   // grid-stride over blocks of survivors
    if (tris) hipLaunchKernelGGL(k_emit<false>, dim3(grid), dim3(kEmitBlock), 0, s, p);
    if (prims) hipLaunchKernelGGL(k_emit<true>, dim3(grid), dim3(kEmitBlock), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_clip(const SetupParams& p, uint32_t jobs, hipStream_t s)
{
    (void)hipGetLastError();   // drop a stale error from an earlier call
    if (jobs == 0) return hipSuccess;
    hipLaunchKernelGGL(k_clip, dim3(grid_for(jobs, kClipGroups, 1024)), dim3(kClipBlock), 0, s, p);
    return hipGetLastError();
}

template <int SRC, int DST>
static void launch_blit_t(const BlitParams& p, uint64_t n, hipStream_t s)
{
    hipLaunchKernelGGL((k_blit<SRC, DST>), dim3(grid_for(n, 256, 16384)), dim3(256), 0, s, p);
}

hipError_t launch_blit(const BlitParams& p, int srcType, int dstType, hipStream_t s)
{
    (void)hipGetLastError();   // drop a stale error from an earlier call
    const uint64_t n = (uint64_t)(p.x1 - p.x0) * (p.y1 - p.y0);
    if (n == 0) return hipSuccess;
    const bool d8 = dstType == SL_COLOR_RGBA_8U;
    if (srcType == SL_COLOR_RGBA_8U) d8 ? launch_blit_t<SL_COLOR_RGBA_8U, SL_COLOR_RGBA_8U>(p, n, s) : launch_blit_t<SL_COLOR_RGBA_8U, SL_COLOR_RGBA_FLOAT>(p, n, s);
    else if (srcType == SL_COLOR_RGBA_FLOAT) d8 ? launch_blit_t<SL_COLOR_RGBA_FLOAT, SL_COLOR_RGBA_8U>(p, n, s) : launch_blit_t<SL_COLOR_RGBA_FLOAT, SL_COLOR_RGBA_FLOAT>(p, n, s);
    else d8 ? launch_blit_t<SL_COLOR_R_FLOAT, SL_COLOR_RGBA_8U>(p, n, s) : launch_blit_t<SL_COLOR_R_FLOAT, SL_COLOR_RGBA_FLOAT>(p, n, s);
    return hipGetLastError();
}

BinPlan bin_plan(uint32_t entries)
{
    BinPlan b;
    uint64_t per = ((uint64_t)entries + kMaxBinBlocks - 1) / kMaxBinBlocks;
    per = (per + 63) / 64 * 64;
    if (per < kMinBinBlock) per = kMinBinBlock;
    b.per = (uint32_t)per;
    b.blocks = (uint32_t)(((uint64_t)entries + per - 1) / per);
    if (b.blocks == 0) b.blocks = 1;
    return b;
}

hipError_t launch_binning(const uint32_t* keys, const uint32_t* vals, const FrameCtrl* ctrl, uint32_t cap,
                          BinPlan plan, uint32_t* counts, uint32_t* totals, uint2* range, uint32_t* out,
                          uint32_t* order, int numTiles, hipStream_t s, RasterItem* items, uint32_t* itemCount,
                          uint32_t* partDone, uint32_t partRecs, uint32_t splitAbove, uint32_t maxParts)
{
    SplitPlan sp{items, itemCount, partDone, partRecs, splitAbove, maxParts};
    (void)hipGetLastError();   // drop a stale error from an earlier call
    const size_t lds = sizeof(uint32_t) * (size_t)numTiles;
    hipLaunchKernelGGL(k_bin_count, dim3(plan.blocks), dim3(256), lds, s, keys, ctrl, cap, plan.per, counts, numTiles);
    hipLaunchKernelGGL(k_bin_colscan, dim3((numTiles + kColTiles - 1) / kColTiles), dim3(kColTiles * kColParts), 0, s,
                       ctrl, cap, plan.per, counts, totals, numTiles);
    if (numTiles > 1024 * kTilesPerThread) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_bin_tilescan, dim3(1), dim3(1024), 0, s, totals, range, order, numTiles, sp);
    hipLaunchKernelGGL(k_bin_scatter, dim3(plan.blocks), dim3(64), lds, s, keys, vals, ctrl, cap, plan.per, counts, range, out,
                       numTiles);
    return hipGetLastError();
}

hipError_t launch_raster_depth(const RasterParams& p, int grid, int cfg, hipStream_t s)
{
    (void)hipGetLastError();   // drop a stale error from an earlier call
    if (p.extra) launch_raster_depth_cfg<1>(p, grid, cfg, s);
    else launch_raster_depth_cfg<0>(p, grid, cfg, s);
    return hipGetLastError();
}

hipError_t launch_raster(const RasterParams& p, int numTiles, int cfg, bool debug, hipStream_t s, bool reject)
{
    (void)hipGetLastError();   // drop a stale error from an earlier call
    const bool floatColor = p.colorType == SL_COLOR_R_FLOAT || p.colorType == SL_COLOR_RG_FLOAT ||
                            p.colorType == SL_COLOR_RGB_FLOAT || p.colorType == SL_COLOR_RGBA_FLOAT;
    const int cf = (p.color == nullptr) ? 0 : (floatColor ? 2 : 1);
    if (p.vis != nullptr)
    {
        // the visibility attachment is only built for RGBA32F colour targets
        // and filled triangles
        if (cf != 2 || p.extra || p.mrt) return hipErrorInvalidValue;
        if (debug) launch_raster_cfg<2, 3>(p, numTiles, cfg, s);
        else launch_raster_cfg<2, 2>(p, numTiles, cfg, s);
    }
    else if (p.mrt)
    {
        if (cf == 0) return hipErrorInvalidValue;   // outputs need colour attachment 0
        if (debug)
        {
            if (cf == 1) launch_raster_cfg<1, 13>(p, numTiles, cfg, s);
            else launch_raster_cfg<2, 13>(p, numTiles, cfg, s);
        }
        else
        {
            if (cf == 1) launch_raster_cfg<1, 12>(p, numTiles, cfg, s);
            else launch_raster_cfg<2, 12>(p, numTiles, cfg, s);
        }
    }
    else if (p.extra)
    {
        if (debug)
        {
            if (cf == 0) launch_raster_cfg<0, 5>(p, numTiles, cfg, s);
            else if (cf == 1) launch_raster_cfg<1, 5>(p, numTiles, cfg, s);
            else launch_raster_cfg<2, 5>(p, numTiles, cfg, s);
        }
        else
        {
            if (cf == 0) launch_raster_cfg<0, 4>(p, numTiles, cfg, s);
            else if (cf == 1) launch_raster_cfg<1, 4>(p, numTiles, cfg, s);
            else launch_raster_cfg<2, 4>(p, numTiles, cfg, s);
        }
    }
    else if (debug && reject)
    {
        if (cf == 0) launch_raster_cfg<0, 17>(p, numTiles, cfg, s);
        else if (cf == 1) launch_raster_cfg<1, 17>(p, numTiles, cfg, s);
        else launch_raster_cfg<2, 17>(p, numTiles, cfg, s);
    }
    else if (debug)
    {
        if (cf == 0) launch_raster_cfg<0, 1>(p, numTiles, cfg, s);
        else if (cf == 1) launch_raster_cfg<1, 1>(p, numTiles, cfg, s);
        else launch_raster_cfg<2, 1>(p, numTiles, cfg, s);
    }
    else if (reject)
    {
        if (cf == 0) launch_raster_cfg<0, 16>(p, numTiles, cfg, s);
        else if (cf == 1) launch_raster_cfg<1, 16>(p, numTiles, cfg, s);
        else launch_raster_cfg<2, 16>(p, numTiles, cfg, s);
    }
    else
    {
        if (cf == 0) launch_raster_cfg<0, 0>(p, numTiles, cfg, s);
        else if (cf == 1) launch_raster_cfg<1, 0>(p, numTiles, cfg, s);
        else launch_raster_cfg<2, 0>(p, numTiles, cfg, s);
    }
    return hipGetLastError();
}

}  // namespace slg
