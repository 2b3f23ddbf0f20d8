// sl_device.h -- data layouts and pinned arithmetic shared by the gfx950
// kernels of the SoftLight draw path and their host launch code.
//
// All kernels are built with -ffp-contract=off: every multiply and add below
// rounds on its own unless fmaf() is written out, which is what makes the
// coverage, depth and order bit-identical to the reference's scalar path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sl_gpu.h"

#define SL_HD __host__ __device__ __forceinline__

namespace slg {

// ---------------------------------------------------------------------------
// Fixed geometry of the pipeline
// ---------------------------------------------------------------------------
constexpr int kSetupBlock = 256;       // triangles per setup workgroup (one per lane)
constexpr int kMaxFan = 7;             // a 9-gon from 6-plane clipping fans into 7 triangles

// Tile raster configurations.  A workgroup owns one tileW x tileH screen
// tile; lane = (row, column segment, x-phase): each lane walks the pixels
// x = xMin + phase + 4k of its row that fall in its segment of tileW/segs
// columns (the reference steps its barycentrics in groups of 4 pixels).
struct TileConfig { int tileW, tileH, segs; };
constexpr int kNumTileConfigs = 12;
constexpr TileConfig kTileConfigs[kNumTileConfigs] = {
    {32, 16, 1},   // 0: 1 wave
    {32, 8, 2},    // 1: 1 wave
    {16, 16, 1},   // 2: 1 wave
    {16, 8, 2},    // 3: 1 wave
    {16, 8, 4},    // 4: 2 waves x 4 rows
    {16, 16, 4},   // 5: 4 waves x 4 rows
    {32, 16, 4},   // 6: 4 waves x 4 rows
    {32, 32, 2},   // 7: 4 waves x 8 rows
    {32, 8, 4},    // 8: 2 waves x 4 rows
    {32, 8, 8},    // 9: 4 waves x 2 rows
    {8, 8, 2},     // 10: 1 wave
    {64, 64, 1},   // 11: 4 waves x 16 rows
};
constexpr int kDefaultTileConfig = 0;

// Configuration 11's 64x64 tile keeps 4096 pixels' depth, colour and pending
// fragments in LDS: 112 KB with an 8-bit colour target, more than the 160 KB a
// workgroup may hold with an RGBA32F target or the visibility attachment, so
// a batch with either is rasterised on configuration 7 (32x32) instead.
constexpr int kLargeTileConfig = 11, kLargeTileFallback = 7;
SL_HD constexpr int effective_tile_config(int cfg, bool floatColor, bool vis)
{
    return (cfg == kLargeTileConfig && (floatColor || vis)) ? kLargeTileFallback : cfg;
}

SL_HD int ilog2i(int v) { int r = 0; while ((1 << r) < v) ++r; return r; }

// ---------------------------------------------------------------------------
// Pinned scalar arithmetic (mirrors the reference's generic branches)
// ---------------------------------------------------------------------------
// float -> int32 with truncation; NaN and out-of-range give INT32_MIN, the
// value x86 cvttss2si produces for the reference's (int32_t) casts.
SL_HD int32_t f2i(float f)
{
    return (f >= -2147483648.0f && f < 2147483648.0f) ? (int32_t)f : INT32_MIN;
}

SL_HD float rcpf(float x) { return 1.0f / x; }

// ls::math::max as std::max: (a < b) ? b : a (NaN in a propagates).
SL_HD float max_ref(float a, float b) { return (a < b) ? b : a; }

// ---------------------------------------------------------------------------
// Draw table entry: everything a kernel needs about one draw call, with
// handles already resolved to device pointers on the host.
// ---------------------------------------------------------------------------
struct BindingDesc
{
    int64_t offset;
    int64_t stride;
    int32_t dims;      // SL_Dimension (components - 1)
    int32_t type;      // SL_DataType
};

// DrawDesc::visSlot when the target has no visibility attachment.
constexpr uint32_t kNoVisSlot = 0xFFu;

// Primitive kinds.  A draw call's processor and rasteriser follow the render
// mode of its first mesh (sl_processor_type_for_draw_mode,
// SL_ShaderProcessor.cpp:12-37; rasterizer.mMode = mRenderMode,
// SL_VertexProcessor.cpp:154, set from meshes[0].mode,
// SL_ProcessorPool.cpp:286,339); whether a mesh reads its index buffer is the
// processor's own test of that mesh's mode (SL_TriProcessor.cpp:682,697).
enum : uint32_t { kPrimFill = 0, kPrimWire = 1, kPrimLine = 2, kPrimPoint = 3, kPrimBad = 0xFFu };

SL_HD uint32_t prim_of_mode(int mode)
{
    switch (mode)
    {
        case RENDER_MODE_TRIANGLES: case RENDER_MODE_INDEXED_TRIANGLES: return kPrimFill;
        case RENDER_MODE_TRI_WIRE: case RENDER_MODE_INDEXED_TRI_WIRE: return kPrimWire;
        case RENDER_MODE_LINES: case RENDER_MODE_INDEXED_LINES: return kPrimLine;
        case RENDER_MODE_POINTS: case RENDER_MODE_INDEXED_POINTS: return kPrimPoint;
        default: return kPrimBad;
    }
}

// SL_TriProcessor.cpp:682,697; SL_LineProcessor.cpp:204; SL_PointProcessor.cpp:79
SL_HD bool mesh_indexed(uint32_t prim, int mode)
{
    if (prim == kPrimFill || prim == kPrimWire) return mode == RENDER_MODE_INDEXED_TRIANGLES || mode == RENDER_MODE_INDEXED_TRI_WIRE;
    if (prim == kPrimLine) return mode == RENDER_MODE_INDEXED_LINES;
    return mode == RENDER_MODE_INDEXED_POINTS;
}

SL_HD uint32_t verts_per_prim(uint32_t prim) { return prim == kPrimPoint ? 1u : prim == kPrimLine ? 2u : 3u; }

// Packed colour formats (SL_ColorCompressed.hpp): bit-field widths of r, g,
// b, a (r lowest, as the structs' bit-fields lie) and texel bytes; all zero
// for other formats.
struct PackedFmt { int bits[4]; int bytes; };

SL_HD PackedFmt packed_format(int type)
{
    switch (type)
    {
        case SL_COLOR_RGB_332: return PackedFmt{{3, 3, 2, 0}, 1};
        case SL_COLOR_RGB_565: return PackedFmt{{5, 6, 5, 0}, 2};
        case SL_COLOR_RGBA_5551: return PackedFmt{{5, 5, 5, 1}, 2};
        case SL_COLOR_RGBA_4444: return PackedFmt{{4, 4, 4, 4}, 2};
        case SL_COLOR_RGBA_1010102: return PackedFmt{{10, 10, 10, 2}, 4};
        default: return PackedFmt{{0, 0, 0, 0}, 0};
    }
}

// rgba_cast<packed, float> (e.g. SL_ColorCompressed.hpp:300-310): each
// component times its maximum, truncated to an int (low bits kept, as for
// 8-bit colour), stored in its bit-field.
SL_HD uint32_t packed_encode(const PackedFmt& f, const float c[4])
{
    uint32_t w = 0;
    int shift = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
    {
        if (f.bits[k] == 0) continue;
        const uint32_t mx = (1u << f.bits[k]) - 1u;
        w |= ((uint32_t)f2i(c[k] * (float)mx) & mx) << shift;
        shift += f.bits[k];
    }
    return w;
}

// rgba_cast<float, packed> (e.g. :354-363): field / maximum; alpha 1 when
// the format has none.
SL_HD void packed_decode(const PackedFmt& f, uint32_t w, float c[4])
{
    int shift = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
    {
        if (f.bits[k] == 0)
        {
            c[k] = (k == 3) ? 1.f : 0.f;
            continue;
        }
        const uint32_t mx = (1u << f.bits[k]) - 1u;
        c[k] = (float)((w >> shift) & mx) / (float)mx;
        shift += f.bits[k];
    }
}

// Line draws walk 16-bit Bresenham coordinates (sl_lowp_t); targets up to
// this size keep every walk's error term inside int16.
constexpr int kMaxLineTarget = 8192;

struct DrawDesc
{
    uint64_t triBase;          // first global triangle of this draw
    uint64_t numTris;          // trisPerInstance * numInstances
    uint64_t trisPerInstance;
    uint64_t elementBegin;
    const uint8_t* vbo;
    uint64_t vboBytes;
    const uint8_t* ibo;
    uint32_t iboType;          // SL_DataType of indices
    uint32_t iboCount;
    const uint8_t* instVbo;    // per-instance mat4 array (or null)
    uint64_t instVboBytes;
    uint64_t posBase;          // first entry of the draw's vertex-pass positions, or kNoPositions
    uint32_t posVerts;         // vertices per instance in them (ids below it; others run the program)
    uint32_t pad32;
    const uint8_t* tex;        // sampled texture texels (or null)
    uint32_t texType;
    uint32_t texW, texH;
    uint32_t uniOffset;        // byte offset of the uniform snapshot in the arena
    uint32_t seqBase;          // debug sequence number of the draw's first triangle
    uint32_t numBindings;
    uint32_t vsProg, fsProg;
    uint32_t visSlot;          // varying slot carrying the vertex weights, or kNoVisSlot
    uint32_t prim;             // primitive kind (kPrimFill, kPrimWire, ...)
    uint8_t  indexed, cull, dtest, dmask, blend, nvary, nout, hasUniforms;
    uint8_t  nearest;          // sample with sl_sample_nearest (SL_UNIFORM_SAMPLE_NEAREST)
    uint8_t  pad8[3];
    BindingDesc bind[SL_VAO_MAX_BINDINGS];
    float scissor[16];         // SL_ViewportState::scissor_matrix at draw time
    float vp[4];               // SL_ViewportState::viewport_rect at draw time
};

// Screen-space triangle after setup: SL_FragmentBin (SL_ShaderUtil.hpp:627-647)
// without the varyings, plus the per-triangle scanline constants of
// SL_ScanlineBounds::init (SL_ScanlineBounds.hpp:137-166).
struct SetupRec
{
    float x[3], y[3];          // snapped screen coordinates
    float z[3];                // NDC depth
    float hw[3];               // 1/w
    float bx[3], by[3], bz[3]; // barycentric planes (ddx, ddy, ddz) * denom
    float bc3;                 // 4th barycentric component times 0 (+-0, or NaN when denom is not finite)
    float v0y, v1y, v0x, v1x;  // scanline constants
    float p20y, p20x, p21xy, p10xy;
    int32_t minY, maxY;
    uint32_t draw;
    uint32_t seq;
    uint32_t state;            // the draw's raster state, see pack_raster_state
    uint32_t xbox;             // columns its spans can reach: min x - 1 | (max x - 1) << 16, clamped to 0..65535
};
static_assert(sizeof(SetupRec) == 144, "SetupRec layout");

// Point and line records reuse SetupRec's storage:
//   point  x[0], y[0]  the texel it addresses (column, row); z[0] its depth
//   line   x[0..1], y[0..1], z[0..1]  the two (clipped) screen ends and depths;
//          bx[0]  1 / |end1 - end0|;  by[0] by[1] by[2] bz[0] bz[1] bz[2]
//          (bit patterns of int32) the Bresenham walk: major-axis start,
//          minor-axis start, major length dx, 2|dminor| (dErr), minor step
//          (+-1), steep flag (major axis is y).
// minY / maxY and xbox bound the texels they address, as for triangles.
struct LineWalk { int32_t X1, Y1, dx, dErr, yErr, steep; };

__device__ __forceinline__ LineWalk line_walk_of(const SetupRec& r)
{
    LineWalk w;
    w.X1 = __float_as_int(r.by[0]);
    w.Y1 = __float_as_int(r.by[1]);
    w.dx = __float_as_int(r.by[2]);
    w.dErr = __float_as_int(r.bz[0]);
    w.yErr = __float_as_int(r.bz[1]);
    w.steep = __float_as_int(r.bz[2]);
    return w;
}

// Minor-axis offset (in steps of yErr) of step k of the walk: the number of
// times `err += dErr; if (err > dx) { y += yErr; err -= 2 dx; }` has stepped
// before pixel k (SL_Geometry.hpp:338-365), in closed form.
SL_HD int32_t line_minor_steps(const LineWalk& w, int32_t k)
{
    const int32_t num = k * w.dErr - w.dx;
    return (num > 0) ? (num + 2 * w.dx - 1) / (2 * w.dx) : 0;
}

// Depth function, depth mask, blend mode, "has colour outputs" and the
// primitive kind of a draw, packed into each of its records so the raster
// loop reads no draw table.
constexpr uint32_t kStateWire = 1u << 8;
constexpr uint32_t kStateLine = 1u << 9;
constexpr uint32_t kStatePoint = 1u << 10;

__host__ __device__ inline uint32_t pack_raster_state(const DrawDesc& d)
{
    return (uint32_t)(d.dtest & 7u) | ((uint32_t)(d.dmask & 1u) << 3) | ((uint32_t)(d.blend & 7u) << 4) |
           ((d.nout > 0 ? 1u : 0u) << 7) | (d.prim == kPrimWire ? kStateWire : 0u) |
           (d.prim == kPrimLine ? kStateLine : 0u) | (d.prim == kPrimPoint ? kStatePoint : 0u);
}

// d0 / d1 of SL_TriRasterizer::render_wireframe (SL_TriRasterizer.cpp:155-156):
// max(|a - b|, 1) on int32.  The row above a triangle's first row can have a
// span end of INT32_MIN (an extrapolated horizontal edge); the int
// subtraction then wraps and |INT32_MIN| stays INT32_MIN, giving 1.
SL_HD int32_t wire_gap(int32_t a, int32_t b)
{
    const int32_t dd = (int32_t)((uint32_t)a - (uint32_t)b);
    const int32_t ad = (dd < 0) ? (int32_t)(0u - (uint32_t)dd) : dd;
    return (ad < 1) ? 1 : ad;
}

struct ClipJob
{
    uint32_t tri;              // global triangle index
    uint32_t recOff;           // first output record, within the triangle's setup chunk
    uint32_t pairOff;          // first output (tile, record) pair, within the chunk
    uint32_t nrec;             // records reserved by k_setup
    uint32_t npair;            // pairs reserved by k_setup
};

// Setup runs as a counting pass, a chunk scan and an emitting pass.  Per
// chunk of kSetupBlock triangles the counting pass leaves this summary (and
// one staging word per surviving triangle); the chunk scan fills in the
// prefixes.
struct ChunkInfo
{
    unsigned long long agg;    // the chunk's packed (bin entries << 30 | records)
    unsigned long long pre;    // packed totals of all earlier chunks (chunk scan)
    uint32_t surv;             // triangles of the chunk that yield records (survivors)
    uint32_t survPre;          // survivors of all earlier chunks (chunk scan)
    uint32_t draw;             // the chunk's draw (draws start on chunk boundaries)
    uint32_t empty;            // records dropped because they cover no pixel
};
static_assert(sizeof(ChunkInfo) == 32, "ChunkInfo layout");

// Staging word of a survivor (slot = chunk * kSetupBlock + its rank among the
// chunk's survivors): lane | records before it in the chunk << 8 | bin entries
// before it in the chunk << 20.
constexpr int kScanBlock = 1024;          // chunks per chunk-scan workgroup

// Device-side frame control block.
struct FrameCtrl
{
    uint32_t survivors;        // triangles that yield records (chunk scan)
    uint32_t clipJobs;         // number of clip jobs appended
    uint32_t totalRecs;
    uint32_t totalPairs;
    uint32_t overflow;         // capacity exceeded (records, pairs or clip jobs)
    uint32_t error;            // the emitting pass found a survivor that yields no record
    uint32_t emptyRecs;        // records that passed push_bin's tests but cover no pixel (not binned)
    // overflow | error << 1 of every earlier frame since the host last reset
    // it: the per-frame reset (k_ctrl_reset) folds them in here before it
    // clears them, so a replayed graph frame that overflowed is still
    // reported after later replays succeed
    uint32_t sticky;
    // Fragment counters, spread over kCounterSlots cache lines (a workgroup
    // adds to slot blockIdx % kCounterSlots): tens of thousands of atomics on
    // one address serialise, on 64 lines they do not.  Readers sum the slots.
    struct CounterSlot
    {
        unsigned long long tested;   // fragments depth-tested
        unsigned long long shaded;   // fragments that passed and were shaded
        unsigned long long pad[6];
    } counters[64];
    SL_HD uint64_t tested() const { uint64_t t = 0; for (const CounterSlot& c : counters) t += c.tested; return t; }
    SL_HD uint64_t shaded() const { uint64_t t = 0; for (const CounterSlot& c : counters) t += c.shaded; return t; }
};
constexpr int kCounterSlots = 64;
static_assert(sizeof(FrameCtrl::CounterSlot) == 64, "one cache line per counter slot");

// The vertex pass (k_vertex): an indexed triangle draw whose vertices are
// shared by several triangles has each (instance, vertex) clip position
// computed once, into a per-batch array, before setup; setup, emitting and
// clipping then read positions instead of running the position program per
// triangle corner.  Work items are blocks of kVertexItem entries, one
// workgroup each, the draw of a block found in a table of (first block, draw)
// pairs.
constexpr uint64_t kNoPositions = ~0ull;
constexpr uint32_t kVertexBlock = 256;                  // threads of a vertex-pass workgroup
constexpr uint32_t kVertexPerThread = 2;                // entries per thread and work item (1, 4, 8 measured slower on the sphere frame)
constexpr uint32_t kVertexItem = kVertexBlock * kVertexPerThread;   // entries per work item
constexpr uint64_t kMaxPositionEntries = 1ull << 27;   // 2 GiB of positions per batch at most
constexpr uint64_t kMinPassEntries = 1ull << 16;       // smaller draws keep the setup pass's vertex cache
struct VtxDraw { uint32_t firstBlock, draw; };

struct SetupParams
{
    const DrawDesc* draws;
    uint32_t numDraws;
    uint64_t totalTris;
    const uint8_t* uniforms;
    int32_t W, H;
    int32_t tilesX, tilesY;
    int32_t tileShiftX, tileShiftY;   // log2 of the tile size
    SetupRec* recs;
    float4* vars;              // 3 * maxVary float4 per record
    uint32_t varStride;        // float4 per record
    uint32_t recCap;
    uint32_t* pairKeys;
    uint32_t* pairVals;
    uint32_t pairCap;
    ClipJob* clipJobs;
    uint32_t clipCap;
    ChunkInfo* chunkInfo;         // per chunk
    unsigned long long* stage;    // per chunk, kSetupBlock staging words
    unsigned long long* blockAgg; // packed counts per kScanBlock chunks (summed by atomics)
    uint32_t* blockSurv;          // survivors per kScanBlock chunks (summed by atomics)
    uint32_t* survChunk;          // per survivor (in submission order): its chunk, written by the chunk scan
    uint32_t numChunks;
    FrameCtrl* ctrl;
    float4* positions;            // vertex-pass positions (draws with posBase != kNoPositions)
    const VtxDraw* vtxDraws;      // by first block
    uint32_t numVtxDraws;
    uint32_t vtxBlocks;           // vertex-pass work items (kVertexItem entries each)
};

// Heavy tiles.  A tile whose record list is long is split into parts of
// about the same length, each raster workgroup taking one part, when the
// batch's draws allow an exact merge (splitClass): every draw unblended on a
// 32-bit float depth target, with one depth function, and either depth
// writes on with a min-type (LESS / LEQUAL: kSplitMin) or max-type (GREATER /
// GEQUAL: kSplitMax) function, or depth writes off (kSplitConst).  The depth
// a pixel holds after a run of fragments is then fold(d, m): the run's own
// result m (its fragments depth-tested from +inf / -inf) replaces d when it
// passes the test against d, so part k can start from the exact depth the
// sequential walk would have (a first pass, mode 1, computes every part's m),
// count exactly the fragments that pass, and leave its last passing
// fragment; the tile's last part to finish merges the parts in order.
// Between the passes a fold pass turns every part's m into the running
// fold(initial, m_0 .. m_j), so each part reads its starting depth once.
enum : uint32_t { kSplitNone = 0, kSplitMin = 1, kSplitMax = 2, kSplitConst = 3 };
constexpr uint32_t kDefaultSplitRecords = 256;  // records per part of a split tile
constexpr uint32_t kSplitAboveParts = 4;        // tiles of more than kSplitAboveParts parts' records split
constexpr uint32_t kMaxTileParts = 1024;        // parts of one split tile at most

// One raster workgroup's work: a whole tile (parts == 1) or part `part` of
// a split tile, records [begin, end) of its list; part results live in slot
// `slot` (the tile's parts in slots slot - part .. slot - part + parts - 1);
// `done` counts the split tile's finished parts.
struct RasterItem
{
    uint32_t tile;
    uint32_t begin, end;
    uint32_t slot;
    uint32_t part, parts;
    uint32_t done;
    uint32_t pad;
};

// A part's result for one pixel: its last passing fragment (record or
// kNoPart, barycentric values) and, with the parity attachments, its pass
// count and the last passing fragment's sequence number.
struct PartPixel
{
    uint32_t rid;
    float s0, s1, s2;
    uint32_t count;
    uint32_t seq;
};
constexpr uint32_t kNoPart = 0xFFFFFFFFu;

struct RasterParams
{
    const DrawDesc* draws;
    const uint8_t* uniforms;
    const SetupRec* recs;
    const float4* vars;
    uint32_t varStride;
    const uint32_t* tileRecs;     // record ids sorted by tile
    const uint2* tileRange;       // [start, end) into tileRecs per tile
    const uint32_t* tileOrder;    // launch order of the tiles (heaviest first), or null
    int32_t W, H, tilesX, tilesY;
    uint8_t* color;               // attachment 0 texels (8-bit or float, colorComps components) or null
    int32_t colorType;
    int32_t colorComps;           // components per colour texel (1..4; bytes per texel when packed)
    int32_t colorPacked;          // packed format of attachment 0 (SL_COLOR_RGB_332 ...), or 0
    uint8_t* colorX[3];           // colour attachments 1..3 (outputs 1..3 of a multiple-output program), or null
    int32_t colorXFloat[3];       // their texels are float (else 8-bit or packed)
    int32_t colorXComps[3];
    int32_t colorXPacked[3];
    int32_t depthHalf;            // the depth attachment is R16F (else R32F)
    float* depth;
    uint32_t* dbgLast;
    uint32_t* dbgCount;
    float4* vis;                  // visibility attachment {seq bits, b0, b1, b2} or null
    int32_t clearColor, clearDepth;   // fused clear of this batch
    uint32_t clearRGBA8;
    float clearRGBAF[4];
    float clearDepthValue;
    int32_t extra;                // some draw of the batch draws wireframe triangles, lines or points
    int32_t mrt;                  // some draw of the batch has several fragment outputs
    FrameCtrl* ctrl;
    // heavy-tile split (items == nullptr: one workgroup per tile, tileOrder)
    const RasterItem* items;      // whole tiles and parts of split tiles, parts first
    const uint32_t* itemCount;    // [0] items, [1] parts of split tiles (items [0, [1]))
    float* partDepth;             // per part slot, per tile pixel: the part's own depth result m
    PartPixel* partRes;           // per part slot, per tile pixel
    uint32_t* partDone;           // per split tile: parts finished
    uint32_t splitClass;          // kSplit*
    int32_t splitDtest;           // the batch's depth function
    uint32_t wgPerCU;             // launch: cap on resident raster workgroups per CU (0: none)
};

// ---------------------------------------------------------------------------
// Blit (SL_BlitProcessor::blit_nearest)
// ---------------------------------------------------------------------------
struct BlitParams
{
    const uint8_t* src;
    uint8_t* dst;
    uint32_t srcW, srcH, dstW, dstH;
    uint32_t srcX0, srcY0, srcX1, srcY1;
    uint32_t x0, x1, y0, y1;           // destination pixels written: [x0, x1) x [y0, y1)
    unsigned long long foutW, foutH;   // 16.16 source steps per destination pixel
};

// ---------------------------------------------------------------------------
// Backward pass (fixed-visibility adjoint)
// ---------------------------------------------------------------------------
// One draw of the saved forward batch with its gradient destinations.
struct GradDesc
{
    DrawDesc d;
    float* gradCol;            // float4 per vertex id (dL/d colour attribute), or null
    uint64_t gradColCount;     // float4 entries in gradCol
    float* gradTex;            // RGBA32F gradient of the sampled texture, or null
    uint64_t gradTexTexels;    // texels in the gradient allocation (padding included)
};

struct SmoothParams
{
    const GradDesc* draws;     // the framebuffer's saved draws, ordered by seqBase (uniOffset into uniforms)
    uint32_t numDraws;
    const uint8_t* uniforms;   // the saved draws' uniform snapshots
    const float4* color;       // hard colour image (colour attachment 0, RGBA32F)
    const float4* vis;         // visibility attachment
    const float4* grad;        // backward: dL/d smoothed image
    float4* out;               // forward: smoothed image; backward: dL/d hard colour image
    float2* posGrad;           // backward: dL/d screen position per (saved draw, vertex), or null
    const float2* posOffset;   // screen-space offsets per (saved draw, vertex), or null
    uint32_t numVerts;         // vertices per saved draw in posGrad / posOffset
    int32_t W, H;
    float sigma;
    // per pixel, written by the boundary pass: .x the row-pass weight of the
    // boundary between x - 1 and x, .y the column-pass weight of the boundary
    // between rows y and y + 1 with its sign bit set when row y receives; NaN
    // where there is no boundary
    float2* bounds;
};

struct BackwardParams
{
    const GradDesc* draws;     // ordered by seqBase
    uint32_t numDraws;
    const float4* vis;         // visibility attachment
    const float4* gradImage;   // dL/d colour attachment 0
    int32_t W, H;
};

}  // namespace slg
