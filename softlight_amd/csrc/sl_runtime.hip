// sl_runtime.hip -- host runtime of the gfx950 draw path: the C ABI of
// include/sl_gpu.h, device-resident resources, and the batched frame
// (setup -> clip -> tile sort -> raster) on the context's HIP stream.
//
// The reference's SL_Context owns resources in vectors indexed by size_t
// handles (SL_Context.hpp:66-82) and runs each draw synchronously through its
// thread pool (SL_ProcessorPool.cpp:280-327).  Here draws and clears are
// recorded and executed as one batch per framebuffer at the next
// synchronisation point; uniform blocks are snapshotted at draw time, so the
// result equals drawing each call in order.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "sl_device.h"
#include "sl_kernels.h"
#include "sl_backward.h"
#include "sl_smooth.h"

using namespace slg;

// The per-frame buffers of sl_draw_static (setup records, varyings, bin
// entries, clip jobs, counters, tile ranges, draw table, uniform arena) live
// in this array; each GPU holds its own copy until the process ends.  Host
// code uses the per-device address from hipGetSymbolAddress, never the name.
static __device__ __align__(256) unsigned char scratch[SL_STATIC_SCRATCH_BYTES];

namespace {

struct DevBuf
{
    void* p = nullptr;
    size_t bytes = 0;
};

struct Vbo { bool valid = false; DevBuf buf; size_t bytes = 0; size_t grad = SL_INVALID_ID; };
struct Ibo { bool valid = false; DevBuf buf; int type = VERTEX_DATA_INT; uint32_t count = 0; uint32_t bpe = 4; };
struct Binding { int dims = VERTEX_DIMENSION_4; int type = VERTEX_DATA_FLOAT; ptrdiff_t offset = 0; ptrdiff_t stride = 0; };
struct Vao { size_t vbo = SL_INVALID_ID; size_t ibo = SL_INVALID_ID; size_t nb = 0; Binding b[SL_VAO_MAX_BINDINGS]; };
struct Tex
{
    bool valid = false;
    DevBuf buf;
    int type = SL_COLOR_RGBA_8U;
    uint16_t w = 0, h = 0, d = 0;
    uint32_t bpt = 0;
    size_t grad = SL_INVALID_ID;
    // as a visibility attachment: the framebuffer whose draws it last
    // recorded, and that recording's generation
    size_t visFbo = SL_INVALID_ID;
    uint64_t visGen = 0;
};
// A draw kept for the backward pass, with the resources its gradients go to
// and its uniform snapshot (the smoothed-coverage passes re-run its vertex
// positions).
struct SavedDraw { DrawDesc d; size_t vbo; size_t tex; std::vector<uint8_t> uni; };
struct Fbo
{
    std::vector<size_t> colors;
    size_t depth = SL_INVALID_ID;
    size_t dbgLast = SL_INVALID_ID;
    size_t dbgCount = SL_INVALID_ID;
    size_t vis = SL_INVALID_ID;
    // draws since the depth attachment was last cleared (visibility attached)
    std::vector<SavedDraw> saved;
    bool savedFromClear = false;     // they start at a depth clear (of savedClearDepth) ...
    double savedClearDepth = 1.0;
    uint64_t visGen = 0;             // ... and recorded this generation of the visibility attachment
    bool savedStale = false;         // a resource they read was re-initialised
    bool savedBlend = false;         // one of them blends: no fixed-visibility adjoint
    std::vector<GradDesc> gradHost;  // the backward draw table last uploaded ...
    DevBuf gradDev;                  // ... and its device copy
    // smoothed coverage: screen-position gradients and offsets (float2 per
    // saved draw and vertex), and the draw table / uniform arena last uploaded
    size_t posGrad = SL_INVALID_ID, posOffset = SL_INVALID_ID;
    uint32_t posVerts = 0;
    std::vector<GradDesc> smoothHost;
    std::vector<uint8_t> smoothUniHost;
    DevBuf smoothDev, smoothUniDev;
};
struct Ubo { uint8_t bytes[SL_MAX_UNIFORM_BUFFER_SIZE]; };
struct Shader { uint8_t cull, dtest, dmask, blend, nvary, nout; uint32_t vsProg, fsProg; size_t ubo; };

// Device buffers of one frame.
struct FrameRes
{
    DevBuf draws, uni, recs, vars, keys, vals, sorted, clip, status, ranges, ctrl, binCounts, binTotals, order;
    uint32_t recCap = 0, pairCap = 0, clipCap = 0;
    // heavy-tile split: work items, their counts, per part slot the first
    // pass's depths and the part results, per split tile a finished-parts
    // counter (absent: no split)
    DevBuf items, itemCount, partDepth, partRes, partDone;
    uint32_t maxParts = 0;
    DevBuf positions, vtxDraws;   // the vertex pass's output and work table
};

struct FrameGraph
{
    bool valid = false;
    FrameRes res;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    FrameCtrl* hCtrl = nullptr;
    uint64_t tris = 0;
    uint64_t pending = 0;                          // replays since the last collection
    size_t fbo = SL_INVALID_ID;                    // target of the captured batch
    bool clearDepth = false;
    double depth = 1.0;                            // its depth clear value
    std::vector<SavedDraw> draws;                  // its draws, for the backward pass
    bool timed = false;                            // event nodes around the setup and raster kernels
};

struct Batch
{
    size_t fbo = SL_INVALID_ID;
    bool clearColor = false, clearDepth = false;
    double rgba[4] = {0, 0, 0, 0};
    double depth = 1.0;
    std::vector<DrawDesc> draws;
    std::vector<size_t> drawVbo, drawTex;          // resource ids behind each draw (backward pass)
    std::vector<uint8_t> uni;
    uint64_t totalTris = 0;
    uint32_t maxVary = 1;
    std::vector<VtxDraw> vtxDraws;                 // draws with vertex-pass positions
    uint64_t posEntries = 0;                       // their (instance, vertex) entries, block-aligned per draw

    bool empty() const { return !clearColor && !clearDepth && draws.empty(); }
    void reset() { *this = Batch{}; }
};

uint32_t bytes_per_color(int type)
{
    switch (type)
    {
        case SL_COLOR_R_8U: return 1;
        case SL_COLOR_R_32U: return 4;
        case SL_COLOR_R_FLOAT: return 4;
        case SL_COLOR_RGBA_8U: return 4;
        case SL_COLOR_RGBA_FLOAT: return 16;
        case SL_COLOR_RG_8U: return 2;
        case SL_COLOR_RGB_8U: return 3;
        case SL_COLOR_R_HALF: return 2;
        case SL_COLOR_RG_FLOAT: return 8;
        case SL_COLOR_RGB_FLOAT: return 12;
        case SL_COLOR_RGB_332: return 1;
        case SL_COLOR_RGB_565: case SL_COLOR_RGBA_5551: case SL_COLOR_RGBA_4444: return 2;
        case SL_COLOR_RGBA_1010102: return 4;
        default: return 0;
    }
}

// Colour attachments drawn to: 8-bit unsigned and float, 1..4 components.
int u8_components(int type)
{
    return type == SL_COLOR_R_8U ? 1 : type == SL_COLOR_RG_8U ? 2 : type == SL_COLOR_RGB_8U ? 3 : type == SL_COLOR_RGBA_8U ? 4 : 0;
}

int float_components(int type)
{
    return type == SL_COLOR_R_FLOAT ? 1 : type == SL_COLOR_RG_FLOAT ? 2 : type == SL_COLOR_RGB_FLOAT ? 3 : type == SL_COLOR_RGBA_FLOAT ? 4 : 0;
}

bool depth_format(int type) { return type == SL_COLOR_R_FLOAT || type == SL_COLOR_R_HALF; }

// A colour attachment draws can write: 8-bit, float or packed.
bool color_format(int type) { return u8_components(type) || float_components(type) || packed_format(type).bytes; }

// Texel units the raster walks per attachment texel: components, or bytes of a packed texel.
int color_units(int type)
{
    if (u8_components(type)) return u8_components(type);
    if (float_components(type)) return float_components(type);
    return packed_format(type).bytes;
}

// The clear colour as a texel of a packed format: rgba_cast<packed, double>
// (SL_Color.cpp:180-184), the products in double, truncated, low bits kept.
uint32_t packed_clear_word(int type, const double rgba[4])
{
    const PackedFmt f = packed_format(type);
    uint32_t w = 0;
    int shift = 0;
    for (int k = 0; k < 4; ++k)
    {
        if (f.bits[k] == 0) continue;
        const uint32_t mx = (1u << f.bits[k]) - 1u;
        const double v = rgba[k] * (double)mx;
        const int64_t iv = (v >= -9.2e18 && v < 9.2e18) ? (int64_t)v : INT64_MIN;
        w |= ((uint32_t)iv & mx) << shift;
        shift += f.bits[k];
    }
    return w;
}

// float -> IEEE binary16, round to nearest even (the depth clear value of a
// half depth attachment, (half)(float)depth, SL_Context.cpp:946-948), and back.
uint16_t host_float_to_half(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    const uint32_t sign = (u >> 16) & 0x8000u, ex = (u >> 23) & 0xFFu;
    uint32_t man = u & 0x7FFFFFu;
    if (ex == 0xFFu) return (uint16_t)(sign | 0x7C00u | (man ? 0x200u | (man >> 13) : 0u));
    const int e = (int)ex - 112;
    if (e >= 31) return (uint16_t)(sign | 0x7C00u);
    if (e <= 0)
    {
        if (e < -10) return (uint16_t)sign;
        man |= 0x800000u;
        const int sh = 14 - e;
        const uint32_t q = man >> sh, rem = man & ((1u << sh) - 1u), hf = 1u << (sh - 1);
        return (uint16_t)(sign | (q + ((rem > hf || (rem == hf && (q & 1u))) ? 1u : 0u)));
    }
    const uint32_t q = man >> 13, rem = man & 0x1FFFu;
    uint32_t r = ((uint32_t)e << 10) | q;
    if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) ++r;
    return (uint16_t)(sign | r);
}

float host_half_to_float(uint16_t h)
{
    const float m = (float)(h & 0x3FFu);
    const int ex = (h >> 10) & 0x1F;
    float v;
    if (ex == 0) v = std::ldexp(m, -24);
    else if (ex == 31) v = (h & 0x3FFu) ? std::numeric_limits<float>::quiet_NaN() : std::numeric_limits<float>::infinity();
    else v = std::ldexp(1024.f + m, ex - 25);
    return (h & 0x8000u) ? -v : v;
}

uint32_t bytes_per_index(int type)
{
    switch (type)
    {
        case VERTEX_DATA_BYTE: return 1;
        case VERTEX_DATA_SHORT: return 2;
        case VERTEX_DATA_INT: return 4;
        default: return 0;
    }
}

}  // namespace

struct sl_context
{
    hipStream_t stream = nullptr;
    std::vector<Vbo> vbos;
    std::vector<Ibo> ibos;
    std::vector<Vao> vaos;
    std::vector<Tex> texs;
    std::vector<Fbo> fbos;
    std::vector<Ubo> ubos;
    std::vector<Shader> shaders;
    int32_t viewport[4] = {0, 0, 65535, 65535};   // SL_ViewportState.hpp:80-81
    int32_t scissor[4] = {0, 0, 65535, 65535};
    unsigned numThreads = 1;
    int tileConfig = kDefaultTileConfig;
    uint32_t splitRecs = kDefaultSplitRecords;   // heavy-tile split: records per part (0: off)
    bool splitAll = false;                       // split triangle-only batches too (sl_set_raster_split)
    uint32_t wgPerCU = 0;                        // raster occupancy cap (sl_set_raster_occupancy; 0: none)
    bool depthReject = false;                    // sl_set_raster_depth_reject
    uint32_t triSeq = 0;
    sl_frame_stats stats{};
    std::string err;

    Batch batch;

    FrameRes res;                                  // eager-frame scratch (grow-only)
    FrameGraph graph;                              // captured frame
    uint64_t lastTotals[3] = {0, 0, 0};            // records, bin entries, clip jobs of the last eager frame
    FrameCtrl* hCtrl = nullptr;                    // pinned readback of the control block
    hipEvent_t ev[11] = {};                        // eager frames: 0-5 stages, 6-7 graph replays, 8-9 counting pass, 10 binning
    hipEvent_t bwdEv[2] = {};                      // first / last backward since the last collection
    hipEvent_t gEv[4] = {};                        // graph replays: setup start/end, raster start/end
    uint64_t bwdPending = 0;
    DevBuf smoothBounds;                           // the smoothing passes' per-pixel boundary weights (float2)
    uint64_t visCounter = 0;                       // generations of visibility recordings
};

namespace {

int fail(sl_context* c, int code, const char* msg)
{
    if (c) c->err = msg;
    return code;
}

int check(sl_context* c, hipError_t e, const char* what)
{
    if (e == hipSuccess) return SL_OK;
    if (c)
    {
        c->err = std::string(what) + ": " + hipGetErrorString(e);
    }
    return SL_ERR_DEVICE;
}

// Transfers and zero fills of resources, ordered on the context's stream (a
// non-blocking stream does not wait for the null stream, so a plain hipMemset
// could still be running when the next kernel of the stream starts) and
// complete when they return.
int upload(sl_context* c, void* dst, const void* src, size_t n, const char* what)
{
    hipError_t e = hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return check(c, e, what);
}

int download(sl_context* c, void* dst, const void* src, size_t n, const char* what)
{
    hipError_t e = hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return check(c, e, what);
}

int zero_fill(sl_context* c, void* dst, size_t n, const char* what)
{
    hipError_t e = hipMemsetAsync(dst, 0, n, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return check(c, e, what);
}

int ensure(sl_context* c, DevBuf& b, size_t bytes)
{
    if (bytes <= b.bytes) return SL_OK;
    if (b.p)
    {
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return check(c, e, "hipStreamSynchronize");
        (void)hipFree(b.p);
        b.p = nullptr;
        b.bytes = 0;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return fail(c, SL_ERR_OUT_OF_MEMORY, "hipMalloc failed");
    b.p = p;
    b.bytes = bytes;
    return SL_OK;
}

void release(DevBuf& b)
{
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

// SL_ViewportState::scissor_matrix / viewport_rect (SL_ViewportState.cpp:17-111)
void viewport_setup(const sl_context* c, float fboW, float fboH, float mat[16], float vp[4])
{
    const float sc[4] = {(float)c->scissor[0], (float)c->scissor[1], (float)c->scissor[2], (float)c->scissor[3]};
    float v[4] = {(float)c->viewport[0], (float)c->viewport[1], (float)c->viewport[2], (float)c->viewport[3]};
    v[0] = std::min(std::max(v[0] + sc[0], 0.f), fboW);
    v[1] = std::min(std::max(v[1] + sc[1], 0.f), fboH);
    v[2] = std::min(std::min(v[2], (sc[0] + sc[2]) - v[0]), fboW - v[0]);
    v[3] = std::min(std::min(v[3], (sc[1] + sc[3]) - v[1]), fboH - v[1]);
    std::memcpy(vp, v, sizeof(v));
    const float inv0 = 1.0f / fboW, inv1 = 1.0f / fboH;
    float x = sc[0] * inv0, y = sc[1] * inv1, w = sc[2] * inv0, h = sc[3] * inv1;
    if (x < 0.f) { w += x; x = 0.f; }
    if (y < 0.f) { h += y; y = 0.f; }
    w = std::min(1.f - x, w);
    h = std::min(1.f - y, h);
    const float nm00 = 1.0f / w, nm11 = 1.0f / h;
    const float nm30 = nm00 - 1.f, nm31 = nm11 - 1.f;
    const float m30 = std::fma(x, (-2.f * nm00), nm30);
    const float m31 = std::fma(y, (-2.f * nm11), nm31);
    const float m[16] = {nm00, 0.f, 0.f, 0.f, 0.f, nm11, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f, m30, m31, 0.f, 1.f};
    std::memcpy(mat, m, sizeof(m));
}

Tex* tex_at(sl_context* c, size_t id)
{
    return (id < c->texs.size() && c->texs[id].valid) ? &c->texs[id] : nullptr;
}

int flush(sl_context* c, bool staticScratch = false);

// ---------------------------------------------------------------------------
// batch execution
// ---------------------------------------------------------------------------
// Everything derived from a batch and its target framebuffer.
struct Plan
{
    int W, H, cfg, tilesX, tilesY, numTiles;
    TileConfig tc;
    Tex* dt;
    Tex* cx[3] = {nullptr, nullptr, nullptr};   // colour attachments 1..3
    Tex* ct;
    Tex* dbgL;
    Tex* dbgC;
    Tex* vis;
    bool debug;
    uint64_t chunks;
    uint32_t varStride;
};

int make_plan(sl_context* c, const Batch& B, Plan& p)
{
    if (B.fbo >= c->fbos.size()) return fail(c, SL_ERR_INVALID_HANDLE, "batch framebuffer");
    Fbo& F = c->fbos[B.fbo];
    p.dt = tex_at(c, F.depth);
    if (!p.dt || !depth_format(p.dt->type)) return fail(c, SL_ERR_STATE, "framebuffer needs an R_FLOAT or R_HALF depth attachment");
    p.ct = F.colors.empty() ? nullptr : tex_at(c, F.colors[0]);
    if (p.ct && !color_format(p.ct->type)) return fail(c, SL_ERR_UNSUPPORTED, "colour attachment format");
    for (int k = 0; k < 3; ++k)
    {
        // attachments 1..3, written by multiple-output programs (validated per draw)
        Tex* xt = (F.colors.size() > (size_t)k + 1) ? tex_at(c, F.colors[k + 1]) : nullptr;
        p.cx[k] = (xt && color_format(xt->type) && xt->w == p.dt->w && xt->h == p.dt->h) ? xt : nullptr;
    }
    p.dbgL = tex_at(c, F.dbgLast);
    p.dbgC = tex_at(c, F.dbgCount);
    p.debug = p.dbgL && p.dbgC;
    p.vis = tex_at(c, F.vis);
    p.W = p.dt->w;
    p.H = p.dt->h;
    if (p.ct && (p.ct->w != p.W || p.ct->h != p.H)) return fail(c, SL_ERR_STATE, "colour and depth attachment sizes differ");
    if (p.vis)
    {
        if (p.vis->type != SL_COLOR_RGBA_FLOAT || p.vis->w != p.W || p.vis->h != p.H) return fail(c, SL_ERR_STATE, "visibility attachment must be RGBA_FLOAT of the framebuffer's size");
        if (!p.ct || p.ct->type != SL_COLOR_RGBA_FLOAT) return fail(c, SL_ERR_UNSUPPORTED, "visibility needs an RGBA_FLOAT colour attachment 0");
        if (p.dt->type != SL_COLOR_R_FLOAT) return fail(c, SL_ERR_UNSUPPORTED, "visibility needs an R_FLOAT depth attachment");
    }
    p.cfg = (c->tileConfig >= 0 && c->tileConfig < kNumTileConfigs) ? c->tileConfig : kDefaultTileConfig;
    {
        const int t = p.ct ? p.ct->type : -1;
        const bool floatColor = t == SL_COLOR_R_FLOAT || t == SL_COLOR_RG_FLOAT || t == SL_COLOR_RGB_FLOAT ||
                                t == SL_COLOR_RGBA_FLOAT;
        p.cfg = effective_tile_config(p.cfg, floatColor, p.vis != nullptr);
    }
    p.tc = kTileConfigs[p.cfg];
    p.tilesX = (p.W + p.tc.tileW - 1) / p.tc.tileW;
    p.tilesY = (p.H + p.tc.tileH - 1) / p.tc.tileH;
    p.numTiles = p.tilesX * p.tilesY;
    if (p.numTiles > 16384) return fail(c, SL_ERR_UNSUPPORTED, "more than 16384 raster tiles: choose a larger tile configuration");
    if (B.totalTris >= (1ull << 32)) return fail(c, SL_ERR_CAPACITY, "more than 2^32 triangles in one batch");
    p.chunks = (B.totalTris + kSetupBlock - 1) / kSetupBlock;
    p.varStride = 3u * std::max<uint32_t>(B.maxVary, 1u);
    return SL_OK;
}

// Setup's per-chunk buffer: ChunkInfo per chunk, kSetupBlock staging words
// per chunk, the block sums (packed totals, then survivors; zeroed before
// every frame) and the first chunk of every block of kSetupBlock survivors.
uint64_t chunk_blocks(uint64_t chunks) { return (chunks + kScanBlock - 1) / kScanBlock; }
size_t chunk_buffer_bytes(uint64_t chunks)
{
    return (sizeof(ChunkInfo) + (sizeof(unsigned long long) + sizeof(uint32_t)) * kSetupBlock) * chunks +
           (sizeof(unsigned long long) + sizeof(uint32_t)) * chunk_blocks(chunks);
}

// Part slots of the heavy-tile split: the per-slot buffers (first-pass
// depths and part results, 28 B per tile pixel) kept within 64 MB.
uint32_t split_parts(const Plan& p)
{
    const size_t pix = (size_t)p.tc.tileW * p.tc.tileH;
    return (uint32_t)std::min<size_t>(16384, (64u << 20) / (pix * (sizeof(float) + sizeof(PartPixel))));
}

// The heavy-tile split applies to a batch whose draws merge exactly
// (RasterParams::splitClass): all unblended, one depth function, a float
// depth target, and depth writes on with a min- or max-type function or off.
uint32_t split_class(const Batch& B, const Plan& p, bool all)
{
    if (B.draws.empty() || p.dt->type != SL_COLOR_R_FLOAT) return kSplitNone;
    // by default only batches with point or line draws, whose records pile
    // onto few tiles (a distant mesh's vertices); triangle batches when asked
    bool prims = false;
    for (const DrawDesc& d : B.draws) prims = prims || d.prim == kPrimPoint || d.prim == kPrimLine;
    if (!all && !prims) return kSplitNone;
    const DrawDesc& d0 = B.draws[0];
    for (const DrawDesc& d : B.draws)
        if (d.blend != SL_BLEND_OFF || d.dtest != d0.dtest || d.dmask != d0.dmask) return kSplitNone;
    if (!d0.dmask) return kSplitConst;
    if (d0.dtest == SL_DEPTH_TEST_LESS_THAN || d0.dtest == SL_DEPTH_TEST_LESS_EQUAL) return kSplitMin;
    if (d0.dtest == SL_DEPTH_TEST_GREATER_THAN || d0.dtest == SL_DEPTH_TEST_GREATER_EQUAL) return kSplitMax;
    return kSplitNone;
}

// Allocates a frame's buffers at the resource set's current capacities.
int ensure_frame(sl_context* c, FrameRes& R, const Plan& p, const Batch& B)
{
    int rc;
    if ((rc = ensure(c, R.ctrl, sizeof(FrameCtrl))) != SL_OK) return rc;
    if ((rc = ensure(c, R.ranges, sizeof(uint2) * (size_t)p.numTiles)) != SL_OK) return rc;
    if ((rc = ensure(c, R.order, sizeof(uint32_t) * (size_t)p.numTiles)) != SL_OK) return rc;
    if (B.totalTris == 0) return SL_OK;
    if (R.recCap == 0) { R.recCap = 1u << 18; R.pairCap = 1u << 20; R.clipCap = 1u << 16; }
    if ((rc = ensure(c, R.draws, sizeof(DrawDesc) * B.draws.size())) != SL_OK) return rc;
    if ((rc = ensure(c, R.uni, std::max<size_t>(B.uni.size(), 16))) != SL_OK) return rc;
    if ((rc = ensure(c, R.status, chunk_buffer_bytes(p.chunks))) != SL_OK) return rc;
    if ((rc = ensure(c, R.recs, sizeof(SetupRec) * (size_t)R.recCap)) != SL_OK) return rc;
    if ((rc = ensure(c, R.vars, sizeof(float4) * (size_t)p.varStride * R.recCap)) != SL_OK) return rc;
    if ((rc = ensure(c, R.keys, sizeof(uint32_t) * (size_t)R.pairCap)) != SL_OK) return rc;
    if ((rc = ensure(c, R.vals, sizeof(uint32_t) * (size_t)R.pairCap)) != SL_OK) return rc;
    if ((rc = ensure(c, R.sorted, sizeof(uint32_t) * (size_t)R.pairCap)) != SL_OK) return rc;
    if ((rc = ensure(c, R.clip, sizeof(ClipJob) * (size_t)R.clipCap)) != SL_OK) return rc;
    if ((rc = ensure(c, R.binCounts, sizeof(uint32_t) * (size_t)kMaxBinBlocks * p.numTiles)) != SL_OK) return rc;
    if ((rc = ensure(c, R.binTotals, sizeof(uint32_t) * (size_t)p.numTiles)) != SL_OK) return rc;
    if (B.posEntries > 0)
    {
        if ((rc = ensure(c, R.positions, sizeof(float4) * (size_t)B.posEntries)) != SL_OK) return rc;
        if ((rc = ensure(c, R.vtxDraws, sizeof(VtxDraw) * B.vtxDraws.size())) != SL_OK) return rc;
    }
    if (c->splitRecs > 0)
    {
        const uint32_t parts = split_parts(p);
        const size_t pix = (size_t)p.tc.tileW * p.tc.tileH;
        if ((rc = ensure(c, R.items, sizeof(RasterItem) * ((size_t)p.numTiles + parts))) != SL_OK) return rc;
        if ((rc = ensure(c, R.itemCount, 2 * sizeof(uint32_t))) != SL_OK) return rc;
        if ((rc = ensure(c, R.partDepth, sizeof(float) * pix * parts)) != SL_OK) return rc;
        if ((rc = ensure(c, R.partRes, sizeof(PartPixel) * pix * parts)) != SL_OK) return rc;
        if ((rc = ensure(c, R.partDone, sizeof(uint32_t) * parts)) != SL_OK) return rc;
        R.maxParts = parts;
    }
    return SL_OK;
}

SetupParams setup_params(const Plan& p, const Batch& B, FrameRes& R)
{
    SetupParams sp{};
    sp.draws = reinterpret_cast<const DrawDesc*>(R.draws.p);
    sp.numDraws = (uint32_t)B.draws.size();
    sp.totalTris = B.totalTris;
    sp.uniforms = reinterpret_cast<const uint8_t*>(R.uni.p);
    sp.W = p.W; sp.H = p.H; sp.tilesX = p.tilesX; sp.tilesY = p.tilesY;
    sp.tileShiftX = ilog2i(p.tc.tileW);
    sp.tileShiftY = ilog2i(p.tc.tileH);
    sp.recs = reinterpret_cast<SetupRec*>(R.recs.p);
    sp.vars = reinterpret_cast<float4*>(R.vars.p);
    sp.varStride = p.varStride;
    sp.recCap = R.recCap;
    sp.pairKeys = reinterpret_cast<uint32_t*>(R.keys.p);
    sp.pairVals = reinterpret_cast<uint32_t*>(R.vals.p);
    sp.pairCap = R.pairCap;
    sp.clipJobs = reinterpret_cast<ClipJob*>(R.clip.p);
    sp.clipCap = R.clipCap;
    sp.chunkInfo = reinterpret_cast<ChunkInfo*>(R.status.p);
    sp.stage = reinterpret_cast<unsigned long long*>(sp.chunkInfo + p.chunks);
    sp.blockAgg = sp.stage + (size_t)kSetupBlock * p.chunks;
    sp.blockSurv = reinterpret_cast<uint32_t*>(sp.blockAgg + chunk_blocks(p.chunks));
    sp.survChunk = sp.blockSurv + chunk_blocks(p.chunks);
    sp.numChunks = (uint32_t)p.chunks;
    sp.ctrl = reinterpret_cast<FrameCtrl*>(R.ctrl.p);
    if (B.posEntries > 0)
    {
        sp.positions = reinterpret_cast<float4*>(R.positions.p);
        sp.vtxDraws = reinterpret_cast<const VtxDraw*>(R.vtxDraws.p);
        sp.numVtxDraws = (uint32_t)B.vtxDraws.size();
        sp.vtxBlocks = (uint32_t)(B.posEntries / kVertexItem);
    }
    return sp;
}

RasterParams raster_params(const Plan& p, const Batch& B, FrameRes& R, bool haveEntries, bool splitAll)
{
    RasterParams rp{};
    rp.draws = reinterpret_cast<const DrawDesc*>(R.draws.p);
    rp.uniforms = reinterpret_cast<const uint8_t*>(R.uni.p);
    rp.recs = reinterpret_cast<const SetupRec*>(R.recs.p);
    rp.vars = reinterpret_cast<const float4*>(R.vars.p);
    rp.varStride = p.varStride;
    rp.tileRecs = haveEntries ? reinterpret_cast<const uint32_t*>(R.sorted.p) : nullptr;
    rp.tileRange = reinterpret_cast<const uint2*>(R.ranges.p);
    rp.tileOrder = haveEntries ? reinterpret_cast<const uint32_t*>(R.order.p) : nullptr;
    rp.W = p.W; rp.H = p.H; rp.tilesX = p.tilesX; rp.tilesY = p.tilesY;
    rp.color = p.ct ? reinterpret_cast<uint8_t*>(p.ct->buf.p) : nullptr;
    rp.colorType = p.ct ? p.ct->type : SL_COLOR_RGBA_8U;
    rp.colorComps = p.ct ? color_units(p.ct->type) : 4;
    rp.colorPacked = (p.ct && packed_format(p.ct->type).bytes) ? p.ct->type : 0;
    rp.depthHalf = (p.dt->type == SL_COLOR_R_HALF) ? 1 : 0;
    for (int k = 0; k < 3; ++k)
    {
        const Tex* xt = p.cx[k];
        rp.colorX[k] = xt ? reinterpret_cast<uint8_t*>(xt->buf.p) : nullptr;
        rp.colorXFloat[k] = (xt && float_components(xt->type)) ? 1 : 0;
        rp.colorXComps[k] = xt ? color_units(xt->type) : 4;
        rp.colorXPacked[k] = (xt && packed_format(xt->type).bytes) ? xt->type : 0;
    }
    rp.depth = reinterpret_cast<float*>(p.dt->buf.p);
    rp.dbgLast = p.debug ? reinterpret_cast<uint32_t*>(p.dbgL->buf.p) : nullptr;
    rp.dbgCount = p.debug ? reinterpret_cast<uint32_t*>(p.dbgC->buf.p) : nullptr;
    rp.vis = p.vis ? reinterpret_cast<float4*>(p.vis->buf.p) : nullptr;
    rp.clearColor = B.clearColor ? 1 : 0;
    rp.clearDepth = B.clearDepth ? 1 : 0;
    uint32_t packed = 0;
    for (int k = 0; k < 4; ++k)
    {
        packed |= ((uint32_t)(uint8_t)(int64_t)(B.rgba[k] * 255.0)) << (8 * k);   // color_cast<uint8_t,double>
        rp.clearRGBAF[k] = (float)B.rgba[k];
    }
    rp.clearRGBA8 = rp.colorPacked ? packed_clear_word(rp.colorPacked, B.rgba) : packed;
    rp.clearDepthValue = rp.depthHalf ? host_half_to_float(host_float_to_half((float)B.depth)) : (float)B.depth;
    rp.extra = 0;
    rp.mrt = 0;
    for (const DrawDesc& d : B.draws)
    {
        rp.extra |= (d.prim != kPrimFill) ? 1 : 0;
        rp.mrt |= (d.nout > 1) ? 1 : 0;
    }
    rp.ctrl = reinterpret_cast<FrameCtrl*>(R.ctrl.p);
    rp.items = nullptr;
    rp.itemCount = nullptr;
    rp.partDepth = nullptr;
    rp.partRes = nullptr;
    rp.partDone = nullptr;
    rp.splitClass = kSplitNone;
    rp.splitDtest = 0;
    if (haveEntries && R.items.p != nullptr && R.maxParts > 0)
    {
        const uint32_t cls = split_class(B, p, splitAll);
        if (cls != kSplitNone)
        {
            rp.items = reinterpret_cast<const RasterItem*>(R.items.p);
            rp.itemCount = reinterpret_cast<const uint32_t*>(R.itemCount.p);
            rp.partDepth = reinterpret_cast<float*>(R.partDepth.p);
            rp.partRes = reinterpret_cast<PartPixel*>(R.partRes.p);
            rp.partDone = reinterpret_cast<uint32_t*>(R.partDone.p);
            rp.splitClass = cls;
            rp.splitDtest = B.draws[0].dtest;
        }
    }
    return rp;
}

// Control-block reset, block-sum reset and the setup passes: counting (the
// events bracket it), chunk scan, emitting, clipping.
int enqueue_front(sl_context* c, FrameRes& R, const Plan& p, const Batch& B, hipStream_t s,
                  hipEvent_t evSetup0 = nullptr, hipEvent_t evSetup1 = nullptr, hipEvent_t evClip0 = nullptr,
                  hipEvent_t evClip1 = nullptr)
{
    int rc;
    FrameCtrl* ctrl = reinterpret_cast<FrameCtrl*>(R.ctrl.p);
    if ((rc = check(c, launch_ctrl_reset(ctrl, s), "reset ctrl")) != SL_OK) return rc;
    if ((rc = check(c, launch_fill_u32(reinterpret_cast<uint32_t*>(R.ranges.p), 0u, 2ull * p.numTiles, s), "fill ranges")) != SL_OK) return rc;
    if (B.totalTris == 0) return SL_OK;
    const SetupParams sp = setup_params(p, B, R);
    if ((rc = check(c, launch_fill_u32(reinterpret_cast<uint32_t*>(sp.blockAgg), 0u, 3 * chunk_blocks(p.chunks), s), "fill block sums")) != SL_OK) return rc;
    if ((rc = check(c, launch_vertex(sp, s), "k_vertex")) != SL_OK) return rc;
    if (evSetup0 && (rc = check(c, hipEventRecord(evSetup0, s), "event")) != SL_OK) return rc;
    if ((rc = check(c, launch_setup(sp, p.chunks, s), "k_setup")) != SL_OK) return rc;
    if (evSetup1 && (rc = check(c, hipEventRecord(evSetup1, s), "event")) != SL_OK) return rc;
    if ((rc = check(c, launch_chunk_scan(sp, s), "k_chunk_scan")) != SL_OK) return rc;
    bool tris = false, prims = false;
    for (const DrawDesc& d : B.draws)
    {
        const bool pl = d.prim == kPrimPoint || d.prim == kPrimLine;
        prims = prims || pl;
        tris = tris || !pl;
    }
    if ((rc = check(c, launch_emit(sp, tris, prims, s), "k_emit")) != SL_OK) return rc;
    // the straddling triangles, whose jobs the counting pass reserved (a
    // grid for the job capacity: the count stays on the device)
    if (tris)
    {
        if (evClip0 && (rc = check(c, hipEventRecord(evClip0, s), "event")) != SL_OK) return rc;
        if ((rc = check(c, launch_clip(sp, R.clipCap, s), "k_clip")) != SL_OK) return rc;
        if (evClip1 && (rc = check(c, hipEventRecord(evClip1, s), "event")) != SL_OK) return rc;
    }
    return SL_OK;
}

// Binning and raster.  entries sizes the grids (a graph uses the capacity;
// the kernels read the actual counts from the control block).
int enqueue_back(sl_context* c, FrameRes& R, const Plan& p, const Batch& B, uint32_t entries, hipStream_t s,
                 hipEvent_t evBin0, hipEvent_t evBin1)
{
    int rc;
    RasterParams rp = raster_params(p, B, R, B.totalTris > 0 && entries > 0, c->splitAll);
    rp.wgPerCU = c->wgPerCU;
    if (evBin0) (void)hipEventRecord(evBin0, s);
    if (B.totalTris > 0)
    {
        const SetupParams sp = setup_params(p, B, R);
        if (entries > 0)
        {
            // with the heavy-tile split on, the tile scan also writes the
            // raster's work items
            const uint32_t partRecs = std::max<uint32_t>(c->splitRecs, 1u);
            if ((rc = check(c, launch_binning(sp.pairKeys, sp.pairVals, sp.ctrl, R.pairCap, bin_plan(entries),
                                              reinterpret_cast<uint32_t*>(R.binCounts.p), reinterpret_cast<uint32_t*>(R.binTotals.p),
                                              reinterpret_cast<uint2*>(R.ranges.p), reinterpret_cast<uint32_t*>(R.sorted.p),
                                              reinterpret_cast<uint32_t*>(R.order.p), p.numTiles, s,
                                              const_cast<RasterItem*>(rp.items), const_cast<uint32_t*>(rp.itemCount), rp.partDone,
                                              partRecs, kSplitAboveParts * partRecs, R.maxParts), "binning")) != SL_OK) return rc;
        }
    }
    if (evBin1) (void)hipEventRecord(evBin1, s);
    if (rp.items != nullptr)
    {
        // heavy tiles: the first pass over split tiles' parts, then one
        // workgroup per work item (capacity-sized grids: the counts stay on
        // the device)
        if (rp.splitClass != kSplitConst && (rc = check(c, launch_raster_depth(rp, (int)R.maxParts, p.cfg, s), "k_raster depth pass")) != SL_OK)
            return rc;
        return check(c, launch_raster(rp, p.numTiles + (int)R.maxParts, p.cfg, p.debug, s, c->depthReject), "k_raster");
    }
    return check(c, launch_raster(rp, p.numTiles, p.cfg, p.debug, s, c->depthReject), "k_raster");
}

int upload_batch(sl_context* c, FrameRes& R, const Batch& B, hipStream_t s)
{
    int rc;
    if (B.draws.empty()) return SL_OK;
    if ((rc = check(c, hipMemcpyAsync(R.draws.p, B.draws.data(), sizeof(DrawDesc) * B.draws.size(), hipMemcpyHostToDevice, s), "upload draws")) != SL_OK) return rc;
    if (!B.uni.empty() && (rc = check(c, hipMemcpyAsync(R.uni.p, B.uni.data(), B.uni.size(), hipMemcpyHostToDevice, s), "upload uniforms")) != SL_OK) return rc;
    if (!B.vtxDraws.empty() &&
        (rc = check(c, hipMemcpyAsync(R.vtxDraws.p, B.vtxDraws.data(), sizeof(VtxDraw) * B.vtxDraws.size(), hipMemcpyHostToDevice, s),
                    "upload vertex-pass table")) != SL_OK)
        return rc;
    return SL_OK;
}

// Draw i's uniform block as the batch snapshotted it (empty without one).
std::vector<uint8_t> uniform_snapshot(const Batch& B, size_t i)
{
    const DrawDesc& d = B.draws[i];
    if (!d.hasUniforms || (size_t)d.uniOffset + SL_MAX_UNIFORM_BUFFER_SIZE > B.uni.size()) return {};
    return std::vector<uint8_t>(B.uni.begin() + d.uniOffset, B.uni.begin() + d.uniOffset + SL_MAX_UNIFORM_BUFFER_SIZE);
}

// Keeps the draws of a batch that rendered with a visibility attachment: the
// backward pass maps each pixel's sequence number back to its draw.
// The framebuffer's draws now fill its visibility attachment: a later
// backward pass of another framebuffer that shares the attachment rebuilds
// its own first (rebuild_visibility).
void mark_visibility(sl_context* c, size_t fbo)
{
    Fbo& F = c->fbos[fbo];
    Tex* V = tex_at(c, F.vis);
    if (!V) return;
    V->visFbo = fbo;
    V->visGen = ++c->visCounter;
    F.visGen = V->visGen;
}

void save_draws(sl_context* c, const Batch& B, const Plan& p)
{
    Fbo& F = c->fbos[B.fbo];
    if (!p.vis)
    {
        F.saved.clear();
        F.savedStale = false;
        F.savedBlend = false;
        F.savedFromClear = false;
        return;
    }
    if (B.clearDepth)
    {
        F.saved.clear();
        F.savedStale = false;
        F.savedBlend = false;
        F.savedFromClear = true;
        F.savedClearDepth = B.depth;
    }
    else if (F.saved.empty())
    {
        F.savedFromClear = false;
    }
    for (size_t i = 0; i < B.draws.size(); ++i)
    {
        F.saved.push_back(SavedDraw{B.draws[i], B.drawVbo[i], B.drawTex[i], uniform_snapshot(B, i)});
        if (B.draws[i].blend != SL_BLEND_OFF) F.savedBlend = true;
    }
    mark_visibility(c, B.fbo);
}

// Resource storage moved: saved draws hold stale device pointers.
void stale_saved_draws(sl_context* c)
{
    for (Fbo& F : c->fbos)
        if (!F.saved.empty()) F.savedStale = true;
}

// ---------------------------------------------------------------------------
// Static scratch (sl_draw_static): the frame's buffers carved out of
// `scratch` instead of allocated.
// ---------------------------------------------------------------------------
constexpr int kMaxScratchDevices = 64;
std::mutex gScratchMutex;
void* gScratchPtr[kMaxScratchDevices] = {};

// The current device's copy of `scratch`, looked up once per device.
int static_scratch(sl_context* c, uint8_t** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return check(c, e, "hipGetDevice");
    if (dev < 0 || dev >= kMaxScratchDevices) return fail(c, SL_ERR_UNSUPPORTED, "device index beyond the static scratch table");
    std::lock_guard<std::mutex> lock(gScratchMutex);
    if (gScratchPtr[dev] == nullptr)
    {
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return check(c, e, "hipGetSymbolAddress(scratch)");
        if (ptr == nullptr) return fail(c, SL_ERR_DEVICE, "hipGetSymbolAddress(scratch) returned null");
        gScratchPtr[dev] = ptr;
    }
    *out = reinterpret_cast<uint8_t*>(gScratchPtr[dev]);
    return SL_OK;
}

size_t align256(size_t n) { return (n + 255) & ~size_t(255); }

// Bin count matrix rows for up to `pairs` entries (bin_plan never uses more).
uint64_t bin_rows(uint64_t pairs) { return std::min<uint64_t>(kMaxBinBlocks, std::max<uint64_t>(1, (pairs + kMinBinBlock - 1) / kMinBinBlock)); }

// Bytes of the fixed part: control block, tile ranges and order, draw table,
// uniform arena, setup chunk buffer, tile totals.
size_t static_fixed_bytes(const Plan& p, const Batch& B)
{
    size_t n = align256(sizeof(FrameCtrl)) + align256(sizeof(uint2) * (size_t)p.numTiles) + align256(sizeof(uint32_t) * (size_t)p.numTiles);
    if (B.totalTris == 0) return n;
    n += align256(sizeof(DrawDesc) * B.draws.size()) + align256(std::max<size_t>(B.uni.size(), 16)) + align256(chunk_buffer_bytes(p.chunks)) +
         align256(sizeof(uint32_t) * (size_t)p.numTiles);
    return n;
}

// Bytes of the capacity-dependent part.
size_t static_var_bytes(const Plan& p, uint64_t recs, uint64_t pairs, uint64_t clips)
{
    return align256(sizeof(SetupRec) * recs) + align256(sizeof(float4) * (size_t)p.varStride * recs) + 3 * align256(sizeof(uint32_t) * pairs) +
           align256(sizeof(ClipJob) * clips) + align256(sizeof(uint32_t) * bin_rows(pairs) * (size_t)p.numTiles);
}

// Points every buffer of R into the array at the given capacities (the fixed
// part at the same offsets every time).  SL_ERR_CAPACITY when they do not fit.
int bind_static(sl_context* c, uint8_t* base, FrameRes& R, const Plan& p, const Batch& B, uint64_t recs, uint64_t pairs, uint64_t clips)
{
    const size_t fixed = static_fixed_bytes(p, B);
    if (fixed > SL_STATIC_SCRATCH_BYTES || (B.totalTris > 0 && static_var_bytes(p, recs, pairs, clips) > SL_STATIC_SCRATCH_BYTES - fixed))
        return fail(c, SL_ERR_CAPACITY, "frame needs more than the SL_STATIC_SCRATCH_BYTES scratch array of sl_draw_static; use sl_draw");
    size_t off = 0;
    auto take = [&](DevBuf& b, size_t bytes) { b.p = base + off; b.bytes = bytes; off += align256(bytes); };
    take(R.ctrl, sizeof(FrameCtrl));
    take(R.ranges, sizeof(uint2) * (size_t)p.numTiles);
    take(R.order, sizeof(uint32_t) * (size_t)p.numTiles);
    if (B.totalTris == 0) return SL_OK;
    take(R.draws, sizeof(DrawDesc) * B.draws.size());
    take(R.uni, std::max<size_t>(B.uni.size(), 16));
    take(R.status, chunk_buffer_bytes(p.chunks));
    take(R.binTotals, sizeof(uint32_t) * (size_t)p.numTiles);
    take(R.recs, sizeof(SetupRec) * recs);
    take(R.vars, sizeof(float4) * (size_t)p.varStride * recs);
    take(R.keys, sizeof(uint32_t) * pairs);
    take(R.vals, sizeof(uint32_t) * pairs);
    take(R.sorted, sizeof(uint32_t) * pairs);
    take(R.clip, sizeof(ClipJob) * clips);
    take(R.binCounts, sizeof(uint32_t) * bin_rows(pairs) * (size_t)p.numTiles);
    R.recCap = (uint32_t)recs;
    R.pairCap = (uint32_t)pairs;
    R.clipCap = (uint32_t)clips;
    return SL_OK;
}

// Capacities filling the free part of the array: records : entries : clip
// jobs as the grown buffers start (1 : 4 : 1/4), or `need` (exact totals of an
// overflowed run) with the rest spread over them in proportion.
void static_capacities(const Plan& p, const Batch& B, const uint64_t need[3], uint64_t cap[3])
{
    const size_t freeBytes = SL_STATIC_SCRATCH_BYTES - std::min<size_t>(SL_STATIC_SCRATCH_BYTES, static_fixed_bytes(p, B));
    const uint64_t base[3] = {need[0] ? need[0] : 4, need[1] ? need[1] : 16, need[2] ? need[2] : 1};
    // largest scale k (in 1/64 steps) at which base * k still fits
    uint64_t lo = 64, hi = 64;
    while (hi < (1ull << 40) && static_var_bytes(p, base[0] * hi / 64, base[1] * hi / 64, base[2] * hi / 64 + 1) <= freeBytes) { lo = hi; hi *= 2; }
    while (hi - lo > 1)
    {
        const uint64_t mid = (lo + hi) / 2;
        if (static_var_bytes(p, base[0] * mid / 64, base[1] * mid / 64, base[2] * mid / 64 + 1) <= freeBytes) lo = mid; else hi = mid;
    }
    for (int k = 0; k < 3; ++k) cap[k] = std::min<uint64_t>(base[k] * lo / 64 + (k == 2 ? 1 : 0), (1ull << 30));
}

// Runs the pending batch: into the context's grow-only buffers, or (static)
// into `scratch` without allocating.
int execute_batch(sl_context* c, bool staticScratch = false)
{
    Batch& B = c->batch;
    if (B.empty()) return SL_OK;
    Plan p;
    int rc = make_plan(c, B, p);
    if (rc != SL_OK) return rc;
    hipStream_t s = c->stream;
    FrameRes local;
    FrameRes& R = staticScratch ? local : c->res;
    uint8_t* base = nullptr;
    if (staticScratch && (rc = static_scratch(c, &base)) != SL_OK) return rc;
    if (staticScratch)
    {
        // the fixed array holds no vertex-pass positions: every corner runs
        // the position program (the results are the same bits)
        for (DrawDesc& d : B.draws) d.posBase = kNoPositions;
        B.vtxDraws.clear();
        B.posEntries = 0;
    }
    uint64_t need[3] = {0, 0, 0};
    FrameCtrl* ctrl = nullptr;
    float msVertex = 0.f, msBin = 0.f, msRaster = 0.f;
    uint32_t totalRecs = 0, totalPairs = 0, jobs = 0;

    for (int attempt = 0; attempt < 4; ++attempt)
    {
        if (staticScratch)
        {
            uint64_t cap[3];
            static_capacities(p, B, need, cap);
            if (attempt > 0 && (cap[0] < need[0] || cap[1] < need[1] || cap[2] < need[2]))
                return fail(c, SL_ERR_CAPACITY, "frame needs more than the SL_STATIC_SCRATCH_BYTES scratch array of sl_draw_static; use sl_draw");
            if ((rc = bind_static(c, base, R, p, B, cap[0], cap[1], cap[2])) != SL_OK) return rc;
        }
        else if ((rc = ensure_frame(c, R, p, B)) != SL_OK)
        {
            return rc;
        }
        ctrl = reinterpret_cast<FrameCtrl*>(R.ctrl.p);
        if (attempt == 0 && (rc = upload_batch(c, R, B, s)) != SL_OK) return rc;
        if (B.totalTris == 0)
        {
            if ((rc = enqueue_front(c, R, p, B, s)) != SL_OK) return rc;
            break;
        }
        (void)hipEventRecord(c->ev[0], s);
        if ((rc = enqueue_front(c, R, p, B, s, c->ev[8], c->ev[9], c->ev[2], c->ev[3])) != SL_OK) return rc;
        (void)hipEventRecord(c->ev[1], s);
        c->stats.setup_launches += 1;
        if ((rc = check(c, hipMemcpyAsync(c->hCtrl, ctrl, sizeof(FrameCtrl), hipMemcpyDeviceToHost, s), "read ctrl")) != SL_OK) return rc;
        if ((rc = check(c, hipStreamSynchronize(s), "k_setup sync")) != SL_OK) return rc;
        if (c->hCtrl->error) return fail(c, SL_ERR_DEVICE, "setup emitting pass disagrees with the counting pass");
        if (!c->hCtrl->overflow) break;
        if (staticScratch)
        {
            // lay the array out for the exact totals and run setup again
            need[0] = c->hCtrl->totalRecs;
            need[1] = c->hCtrl->totalPairs;
            need[2] = c->hCtrl->clipJobs;
            if (attempt == 3) return fail(c, SL_ERR_CAPACITY, "static scratch layout did not converge");
            continue;
        }
        // grow to the exact requirement (+25%) and run setup again; the
        // framebuffer has not been touched yet.
        auto grow = [](uint64_t need, uint32_t cur) -> uint64_t { return std::max<uint64_t>(cur, need + need / 4 + 1024); };
        const uint64_t nr = grow(c->hCtrl->totalRecs, R.recCap), np = grow(c->hCtrl->totalPairs, R.pairCap),
                       nc = grow(c->hCtrl->clipJobs, R.clipCap);
        if (nr >= (1ull << 30) || np >= (1ull << 31) || nc >= (1ull << 31)) return fail(c, SL_ERR_CAPACITY, "frame exceeds bin capacity");
        R.recCap = (uint32_t)nr; R.pairCap = (uint32_t)np; R.clipCap = (uint32_t)nc;
        if (attempt == 3) return fail(c, SL_ERR_CAPACITY, "bin capacity growth did not converge");
    }
    if (B.totalTris > 0)
    {
        float t = 0.f;
        (void)hipEventElapsedTime(&t, c->ev[0], c->ev[1]);
        msVertex += t;
        (void)hipEventElapsedTime(&t, c->ev[8], c->ev[9]);
        c->stats.ms_setup_kernel += t;
        totalRecs = c->hCtrl->totalRecs;
        totalPairs = c->hCtrl->totalPairs;
        jobs = c->hCtrl->clipJobs;
    }
    if ((rc = enqueue_back(c, R, p, B, totalPairs, s, c->ev[10], c->ev[4])) != SL_OK) return rc;
    (void)hipEventRecord(c->ev[5], s);
    if ((rc = check(c, hipMemcpyAsync(c->hCtrl, ctrl, sizeof(FrameCtrl), hipMemcpyDeviceToHost, s), "read ctrl")) != SL_OK) return rc;
    if ((rc = check(c, hipStreamSynchronize(s), "frame sync")) != SL_OK) return rc;

    float t = 0.f;
    if (B.totalTris > 0)
    {
        // k_clip's own time, inside the setup span
        (void)hipEventElapsedTime(&t, c->ev[2], c->ev[3]);
        c->stats.ms_clip_kernel += t;
    }
    (void)hipEventElapsedTime(&t, c->ev[10], c->ev[4]);
    msBin += t;
    (void)hipEventElapsedTime(&t, c->ev[4], c->ev[5]);
    msRaster += t;
    c->stats.triangles_binned += (uint64_t)totalRecs + c->hCtrl->emptyRecs;
    c->stats.tile_pairs += totalPairs;
    c->stats.fragments_tested += c->hCtrl->tested();
    c->stats.fragments_shaded += c->hCtrl->shaded();
    c->stats.frames += 1;
    c->stats.raster_launches += 1;
    c->stats.ms_vertex += msVertex;
    c->stats.ms_bin += msBin;
    c->stats.ms_raster += msRaster;
    c->stats.ms_total += msVertex + msBin + msRaster;
    c->lastTotals[0] = totalRecs;
    c->lastTotals[1] = totalPairs;
    c->lastTotals[2] = jobs;
    save_draws(c, B, p);
    return SL_OK;
}

// ---------------------------------------------------------------------------
// HIP-graph frame: the pending batch, captured with private buffers sized from
// an eager run of the same batch, replayed without host round trips.
// ---------------------------------------------------------------------------
void graph_release(sl_context* c)
{
    FrameGraph& G = c->graph;
    if (G.exec) (void)hipGraphExecDestroy(G.exec);
    if (G.graph) (void)hipGraphDestroy(G.graph);
    G.exec = nullptr;
    G.graph = nullptr;
    for (DevBuf* b : {&G.res.draws, &G.res.uni, &G.res.recs, &G.res.vars, &G.res.keys, &G.res.vals, &G.res.sorted,
                      &G.res.clip, &G.res.status, &G.res.ranges, &G.res.ctrl, &G.res.binCounts, &G.res.binTotals, &G.res.order,
                      &G.res.items, &G.res.itemCount, &G.res.partDepth, &G.res.partRes, &G.res.partDone,
                      &G.res.positions, &G.res.vtxDraws}) release(*b);
    G.res = FrameRes{};
    if (G.hCtrl) (void)hipHostFree(G.hCtrl);
    G.hCtrl = nullptr;
    G.valid = false;
    G.pending = 0;
    G.draws.clear();
    G.timed = false;
}

int graph_capture(sl_context* c)
{
    Batch B = c->batch;                      // keep a copy: executing the batch resets it
    if (B.empty()) return fail(c, SL_ERR_STATE, "no pending draws to capture");
    int rc = flush(c);                       // eager run: sizes buffers, produces the frame
    if (rc != SL_OK) return rc;
    graph_release(c);
    FrameGraph& G = c->graph;
    Plan p;
    if ((rc = make_plan(c, B, p)) != SL_OK) return rc;
    hipStream_t s = c->stream;
    auto grow = [](uint64_t need) -> uint32_t { return (uint32_t)std::min<uint64_t>(need + need / 4 + 1024, (1ull << 30)); };
    G.res.recCap = grow(c->lastTotals[0]);
    G.res.pairCap = grow(c->lastTotals[1]);
    G.res.clipCap = grow(c->lastTotals[2]);
    if ((rc = ensure_frame(c, G.res, p, B)) != SL_OK) { graph_release(c); return rc; }
    if ((rc = upload_batch(c, G.res, B, s)) != SL_OK) { graph_release(c); return rc; }
    // the replays' reset keeps FrameCtrl::sticky: start it (and the rest) at zero
    if ((rc = check(c, launch_fill_u32(reinterpret_cast<uint32_t*>(G.res.ctrl.p), 0u, sizeof(FrameCtrl) / 4, s), "fill ctrl")) != SL_OK)
    {
        graph_release(c);
        return rc;
    }
    if (hipHostMalloc(reinterpret_cast<void**>(&G.hCtrl), sizeof(FrameCtrl), hipHostMallocDefault) != hipSuccess)
    {
        graph_release(c);
        return fail(c, SL_ERR_OUT_OF_MEMORY, "pinned control block");
    }
    if ((rc = check(c, hipStreamSynchronize(s), "sync")) != SL_OK) { graph_release(c); return rc; }
    // first with event nodes around the setup and raster kernels (per-kernel
    // device times of the replays), then without them if that capture fails
    for (int timed = 1; timed >= 0; --timed)
    {
        if ((rc = check(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture")) != SL_OK) { graph_release(c); return rc; }
        int rc1 = timed ? enqueue_front(c, G.res, p, B, s, c->gEv[0], c->gEv[1]) : enqueue_front(c, G.res, p, B, s);
        int rc2 = (rc1 == SL_OK) ? enqueue_back(c, G.res, p, B, G.res.pairCap, s, nullptr,
                                                timed ? c->gEv[2] : nullptr) : rc1;
        if (rc2 == SL_OK && timed) rc2 = check(c, hipEventRecord(c->gEv[3], s), "event");
        int rc3 = (rc2 == SL_OK) ? check(c, hipMemcpyAsync(G.hCtrl, G.res.ctrl.p, sizeof(FrameCtrl), hipMemcpyDeviceToHost, s), "read ctrl") : rc2;
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &graph);
        if (rc3 == SL_OK && e == hipSuccess && graph)
        {
            hipGraphExec_t exec = nullptr;
            if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess)
            {
                G.graph = graph;
                G.exec = exec;
                G.timed = timed != 0;
                break;
            }
        }
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        if (!timed)
        {
            graph_release(c);
            return fail(c, SL_ERR_DEVICE, rc3 != SL_OK ? "graph capture" : "graph end capture / instantiate");
        }
    }
    // submitted triangles (B.totalTris also counts the padding that starts
    // every draw on a setup-chunk boundary)
    G.tris = 0;
    for (const DrawDesc& d : B.draws) G.tris += d.numTris;
    G.fbo = B.fbo;
    G.clearDepth = B.clearDepth;
    G.depth = B.depth;
    G.draws.clear();
    for (size_t i = 0; i < B.draws.size(); ++i) G.draws.push_back(SavedDraw{B.draws[i], B.drawVbo[i], B.drawTex[i], uniform_snapshot(B, i)});
    G.valid = true;
    return SL_OK;
}

int graph_collect(sl_context* c)
{
    FrameGraph& G = c->graph;
    if (!G.valid || G.pending == 0) return SL_OK;
    int rc = check(c, hipStreamSynchronize(c->stream), "graph sync");
    if (rc != SL_OK) return rc;
    float t = 0.f;
    (void)hipEventElapsedTime(&t, c->ev[6], c->ev[7]);
    c->stats.ms_total += t;
    c->stats.frames += G.pending;
    c->stats.triangles_submitted += G.tris * G.pending;
    c->stats.triangles_binned += ((uint64_t)G.hCtrl->totalRecs + G.hCtrl->emptyRecs) * G.pending;
    c->stats.tile_pairs += (uint64_t)G.hCtrl->totalPairs * G.pending;
    c->stats.fragments_tested += G.hCtrl->tested() * G.pending;
    c->stats.fragments_shaded += G.hCtrl->shaded() * G.pending;
    if (G.timed)
    {
        // every replay runs the same work: the last replay's kernel times
        float ts = 0.f, tr = 0.f;
        if (hipEventElapsedTime(&ts, c->gEv[0], c->gEv[1]) == hipSuccess && hipEventElapsedTime(&tr, c->gEv[2], c->gEv[3]) == hipSuccess)
        {
            c->stats.ms_setup_kernel += (double)ts * G.pending;
            c->stats.setup_launches += G.pending;
            c->stats.ms_raster += (double)tr * G.pending;
            c->stats.raster_launches += G.pending;
        }
    }
    G.pending = 0;
    // the last replay's flags, and those of every earlier replay since the
    // last collection (folded into `sticky` by each replay's reset)
    const bool error = G.hCtrl->error || (G.hCtrl->sticky & 2u);
    const bool overflow = G.hCtrl->overflow || (G.hCtrl->sticky & 1u);
    if (G.hCtrl->sticky)
    {
        FrameCtrl* dctrl = reinterpret_cast<FrameCtrl*>(G.res.ctrl.p);
        rc = check(c, launch_fill_u32(&dctrl->sticky, 0u, 1, c->stream), "reset sticky flags");
        if (rc == SL_OK) rc = check(c, hipStreamSynchronize(c->stream), "sync");
        G.hCtrl->sticky = 0;
        if (rc != SL_OK) return rc;
    }
    if (error) return fail(c, SL_ERR_DEVICE, "graph frame: setup emitting pass disagrees with the counting pass");
    if (overflow) return fail(c, SL_ERR_CAPACITY, "a graph frame exceeded its captured capacity (its targets were left untouched); capture again");
    return SL_OK;
}

// Adds the device time of the backward passes enqueued since the last
// collection to the counters (waits for the last one).
void collect_backward(sl_context* c)
{
    if (c->bwdPending == 0) return;
    float t = 0.f;
    if (hipEventSynchronize(c->bwdEv[1]) == hipSuccess && hipEventElapsedTime(&t, c->bwdEv[0], c->bwdEv[1]) == hipSuccess)
        c->stats.ms_backward += t;
    c->stats.backward_launches += c->bwdPending;
    c->bwdPending = 0;
}

int flush(sl_context* c, bool staticScratch)
{
    if (c->batch.empty()) return SL_OK;
    const int rc = execute_batch(c, staticScratch);
    c->batch.reset();
    return rc;
}

// ---------------------------------------------------------------------------
// draw recording
// ---------------------------------------------------------------------------
// `procMode` is the render mode that picks the primitive kind: the mesh's
// own for sl_draw / sl_draw_instanced, the first mesh's for sl_draw_multiple.
int record_draw(sl_context* c, const sl_mesh& m, uint64_t numInstances, size_t shaderId, size_t fboId, int procMode)
{
    if (shaderId >= c->shaders.size()) return fail(c, SL_ERR_INVALID_HANDLE, "shader handle");
    if (fboId >= c->fbos.size()) return fail(c, SL_ERR_INVALID_HANDLE, "framebuffer handle");
    if (m.vaoId >= c->vaos.size()) return fail(c, SL_ERR_INVALID_HANDLE, "vao handle");
    const uint32_t prim = prim_of_mode(procMode);
    if (prim == kPrimBad || prim_of_mode(m.mode) == kPrimBad) return fail(c, SL_ERR_UNSUPPORTED, "render mode");
    const Shader& sh = c->shaders[shaderId];
    if (sh.nout > 1 && sh.fsProg != SL_PROGRAM_LAMBERT_MRT) return fail(c, SL_ERR_UNSUPPORTED, "several fragment outputs need SL_PROGRAM_LAMBERT_MRT");
    Fbo& F = c->fbos[fboId];
    const bool vis = tex_at(c, F.vis) != nullptr;
    if (vis && sh.nvary >= SL_SHADER_MAX_VARYING_VECTORS) return fail(c, SL_ERR_UNSUPPORTED, "visibility needs a free varying slot (numVaryings <= 3)");
    if (vis && prim != kPrimFill) return fail(c, SL_ERR_UNSUPPORTED, "visibility attachments cover filled triangles only");
    if (vis && sh.nout > 1) return fail(c, SL_ERR_UNSUPPORTED, "visibility attachments need a single fragment output");
    Tex* dt = tex_at(c, F.depth);
    if (!dt || !depth_format(dt->type)) return fail(c, SL_ERR_STATE, "framebuffer needs an R_FLOAT or R_HALF depth attachment");
    if (prim == kPrimLine && (dt->w > kMaxLineTarget || dt->h > kMaxLineTarget))
        return fail(c, SL_ERR_UNSUPPORTED, "line draws need a target of at most 8192 x 8192");
    for (uint32_t k = 0; k < sh.nout; ++k)
    {
        Tex* ct = (F.colors.size() > k) ? tex_at(c, F.colors[k]) : nullptr;
        if (!ct) return fail(c, SL_ERR_STATE, "framebuffer lacks a colour attachment for a fragment output");
        if (!color_format(ct->type)) return fail(c, SL_ERR_UNSUPPORTED, "colour attachment format");
        if (ct->w != dt->w || ct->h != dt->h) return fail(c, SL_ERR_STATE, "colour and depth attachment sizes differ");
    }
    const Vao& vao = c->vaos[m.vaoId];
    const Vbo* vbo = (vao.vbo < c->vbos.size() && c->vbos[vao.vbo].valid) ? &c->vbos[vao.vbo] : nullptr;
    const Ibo* ibo = (vao.ibo < c->ibos.size() && c->ibos[vao.ibo].valid) ? &c->ibos[vao.ibo] : nullptr;
    const bool indexed = mesh_indexed(prim, m.mode);
    if (indexed && !ibo) return fail(c, SL_ERR_STATE, "indexed draw without an index buffer");
    const bool needUni = sh.vsProg != SL_PROGRAM_PASSTHROUGH_GOURAUD || sh.fsProg != SL_PROGRAM_PASSTHROUGH_GOURAUD;
    if (needUni && sh.ubo >= c->ubos.size()) return fail(c, SL_ERR_STATE, "program needs a uniform buffer");

    if (c->batch.fbo != fboId && !c->batch.empty())
    {
        const int rc = flush(c);
        if (rc != SL_OK) return rc;
    }
    Batch& B = c->batch;
    B.fbo = fboId;

    if ((!indexed && m.elementEnd > 0xFFFFFFFFull) || numInstances > 0xFFFFFFFFull)
        return fail(c, SL_ERR_UNSUPPORTED, "vertex ids and instance counts are limited to 2^32");
    const uint64_t count = (m.elementEnd > m.elementBegin) ? (m.elementEnd - m.elementBegin) : 0;
    const uint64_t tpi = count / verts_per_prim(prim);   // whole primitives only
    const uint64_t numTris = tpi * numInstances;
    c->stats.triangles_submitted += numTris;
    if (numTris == 0) return SL_OK;

    DrawDesc d;
    std::memset(&d, 0, sizeof(d));
    // each draw starts on a setup-chunk boundary, so a setup workgroup only
    // ever sees one draw
    B.totalTris = (B.totalTris + kSetupBlock - 1) / kSetupBlock * kSetupBlock;
    d.triBase = B.totalTris;
    d.numTris = numTris;
    d.trisPerInstance = tpi;
    d.elementBegin = m.elementBegin;
    d.vbo = vbo ? reinterpret_cast<const uint8_t*>(vbo->buf.p) : nullptr;
    d.vboBytes = vbo ? vbo->bytes : 0;
    d.ibo = ibo ? reinterpret_cast<const uint8_t*>(ibo->buf.p) : nullptr;
    d.iboType = ibo ? (uint32_t)ibo->type : 0u;
    d.iboCount = ibo ? ibo->count : 0u;
    d.indexed = indexed ? 1 : 0;
    d.numBindings = (uint32_t)vao.nb;
    for (size_t i = 0; i < vao.nb; ++i)
    {
        d.bind[i].offset = vao.b[i].offset;
        d.bind[i].stride = vao.b[i].stride;
        d.bind[i].dims = vao.b[i].dims;
        d.bind[i].type = vao.b[i].type;
    }
    d.vsProg = sh.vsProg;
    d.fsProg = sh.fsProg;
    d.posBase = kNoPositions;
    {
        // the vertex pass pays where corners share vertices over a large
        // draw: indexed triangles of a transforming program whose vertex
        // buffer holds at most two vertices per triangle (a closed mesh has
        // about a half), with at least kMinPassEntries (instance, vertex)
        // entries (for small meshes the setup pass's per-chunk vertex cache
        // is cheaper than an extra pass)
        const bool transforms = sh.vsProg == SL_PROGRAM_LAMBERT_TEXTURED || sh.vsProg == SL_PROGRAM_LAMBERT_MRT ||
                                sh.vsProg == SL_PROGRAM_COLOR_TEXTURED;
        const int64_t stride = vao.nb > 0 ? (int64_t)vao.b[0].stride : 0;
        const uint64_t verts = (vbo && stride > 0) ? (vbo->bytes + (uint64_t)stride - 1) / (uint64_t)stride : 0;
        const uint64_t entries = verts * numInstances;
        const uint64_t blocks = (entries + kVertexItem - 1) / kVertexItem;
        if (transforms && indexed && (prim == kPrimFill || prim == kPrimWire) && verts > 0 && verts < (1ull << 31) &&
            entries <= 2 * numTris && entries >= kMinPassEntries && B.posEntries + blocks * kVertexItem <= kMaxPositionEntries)
        {
            d.posBase = B.posEntries;
            d.posVerts = (uint32_t)verts;
            B.vtxDraws.push_back(VtxDraw{(uint32_t)(B.posEntries / kVertexItem), (uint32_t)B.draws.size()});
            B.posEntries += blocks * kVertexItem;
        }
    }
    d.cull = sh.cull;
    d.dtest = sh.dtest;
    d.dmask = sh.dmask;
    d.blend = sh.blend;
    d.nvary = sh.nvary;
    d.nout = sh.nout;
    d.visSlot = kNoVisSlot;
    d.prim = prim;
    if (vis)
    {
        d.visSlot = sh.nvary;
        d.nvary = (uint8_t)(sh.nvary + 1);
    }
    d.seqBase = c->triSeq;
    c->triSeq += (uint32_t)numTris;
    viewport_setup(c, (float)dt->w, (float)dt->h, d.scissor, d.vp);
    if (sh.ubo < c->ubos.size())
    {
        const size_t off = B.uni.size();
        B.uni.resize(off + SL_MAX_UNIFORM_BUFFER_SIZE);
        std::memcpy(B.uni.data() + off, c->ubos[sh.ubo].bytes, SL_MAX_UNIFORM_BUFFER_SIZE);
        d.uniOffset = (uint32_t)off;
        d.hasUniforms = 1;
        sl_std_uniforms u;
        std::memcpy(&u, c->ubos[sh.ubo].bytes, sizeof(u));
        d.nearest = (u.flags & SL_UNIFORM_SAMPLE_NEAREST) ? 1 : 0;
        if (u.texture != SL_NO_RESOURCE && u.texture < c->texs.size() && c->texs[u.texture].valid)
        {
            const Tex& t = c->texs[u.texture];
            d.tex = reinterpret_cast<const uint8_t*>(t.buf.p);
            d.texType = (uint32_t)t.type;
            d.texW = t.w;
            d.texH = t.h;
        }
        if (u.instance_vbo != SL_NO_RESOURCE && u.instance_vbo < c->vbos.size() && c->vbos[u.instance_vbo].valid)
        {
            d.instVbo = reinterpret_cast<const uint8_t*>(c->vbos[u.instance_vbo].buf.p);
            d.instVboBytes = c->vbos[u.instance_vbo].bytes;
        }
    }
    size_t texId = SL_INVALID_ID;
    if (d.hasUniforms)
    {
        sl_std_uniforms u;
        std::memcpy(&u, c->ubos[sh.ubo].bytes, sizeof(u));
        if (d.tex) texId = u.texture;
    }
    B.draws.push_back(d);
    B.drawVbo.push_back(vbo ? vao.vbo : SL_INVALID_ID);
    B.drawTex.push_back(texId);
    B.totalTris += numTris;
    B.maxVary = std::max<uint32_t>(B.maxVary, d.nvary);
    return SL_OK;
}

int fill_texture(sl_context* c, Tex& t, const void* value, uint32_t bpt)
{
    const size_t n = (size_t)t.w * t.h * std::max<uint16_t>(t.d, 1);
    if (bpt == 4)
    {
        uint32_t v;
        std::memcpy(&v, value, 4);
        return check(c, launch_fill_u32(reinterpret_cast<uint32_t*>(t.buf.p), v, n, c->stream), "fill texture");
    }
    if (bpt % 4 == 0)
    {
        uint32_t v[4];
        std::memcpy(v, value, bpt);
        bool same = true;
        for (uint32_t k = 1; k < bpt / 4; ++k) same = same && v[k] == v[0];
        if (same)
            return check(c, launch_fill_u32(reinterpret_cast<uint32_t*>(t.buf.p), v[0], (bpt / 4) * n, c->stream), "fill texture");
    }
    if (bpt == 0 || bpt > 16) return fail(c, SL_ERR_UNSUPPORTED, "clear of this texture format");
    std::vector<uint8_t> host(n * bpt);
    for (size_t i = 0; i < n; ++i) std::memcpy(host.data() + bpt * i, value, bpt);
    return upload(c, t.buf.p, host.data(), host.size(), "fill texture");
}

int record_clear(sl_context* c, size_t fboId, int att, const double* rgba, bool depth, double dval)
{
    if (fboId >= c->fbos.size()) return fail(c, SL_ERR_INVALID_HANDLE, "framebuffer handle");
    Fbo& F = c->fbos[fboId];
    if (rgba)
    {
        if (att < 0 || (size_t)att >= F.colors.size() || !tex_at(c, F.colors[att])) return fail(c, SL_ERR_INVALID_HANDLE, "colour attachment");
        const int ty = c->texs[F.colors[att]].type;
        if (!color_format(ty)) return fail(c, SL_ERR_UNSUPPORTED, "clear of this colour format");
    }
    if (depth)
    {
        Tex* dt = tex_at(c, F.depth);
        if (!dt) return fail(c, SL_ERR_STATE, "no depth attachment");
        if (!depth_format(dt->type)) return fail(c, SL_ERR_UNSUPPORTED, "depth format");
    }
    // a clear after draws, or of another framebuffer, ends the current batch
    if (!c->batch.empty() && (c->batch.fbo != fboId || !c->batch.draws.empty()))
    {
        const int rc = flush(c);
        if (rc != SL_OK) return rc;
    }
    if (rgba && att != 0)
    {
        // attachments other than 0 are cleared directly
        Tex& t = c->texs[F.colors[att]];
        if (const int n = u8_components(t.type))
        {
            uint8_t v[4];
            for (int k = 0; k < 4; ++k) v[k] = (uint8_t)(int64_t)(rgba[k] * 255.0);
            return fill_texture(c, t, v, (uint32_t)n);
        }
        if (packed_format(t.type).bytes)
        {
            const uint32_t w = packed_clear_word(t.type, rgba);
            return fill_texture(c, t, &w, (uint32_t)packed_format(t.type).bytes);
        }
        float v[4];
        for (int k = 0; k < 4; ++k) v[k] = (float)rgba[k];
        return fill_texture(c, t, v, 4u * (uint32_t)float_components(t.type));
    }
    Batch& B = c->batch;
    B.fbo = fboId;
    if (rgba)
    {
        B.clearColor = true;
        for (int k = 0; k < 4; ++k) B.rgba[k] = rgba[k];
    }
    if (depth)
    {
        B.clearDepth = true;
        B.depth = dval;
    }
    return SL_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int sl_context_create(sl_context** out)
{
    if (!out) return SL_ERR_INVALID_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return SL_ERR_DEVICE;
    sl_context* c = new sl_context();
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return SL_ERR_DEVICE; }
    for (auto& e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) { delete c; return SL_ERR_DEVICE; }
    for (auto& e : c->bwdEv)
        if (hipEventCreate(&e) != hipSuccess) { delete c; return SL_ERR_DEVICE; }
    for (auto& e : c->gEv)
        if (hipEventCreate(&e) != hipSuccess) { delete c; return SL_ERR_DEVICE; }
    if (hipHostMalloc(reinterpret_cast<void**>(&c->hCtrl), sizeof(FrameCtrl), hipHostMallocDefault) != hipSuccess) { delete c; return SL_ERR_DEVICE; }
    std::memset(c->hCtrl, 0, sizeof(FrameCtrl));
    *out = c;
    return SL_OK;
}

void sl_context_destroy(sl_context* c)
{
    if (!c) return;
    (void)flush(c);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& v : c->vbos) release(v.buf);
    for (auto& v : c->ibos) release(v.buf);
    for (auto& t : c->texs) release(t.buf);
    graph_release(c);
    for (DevBuf* b : {&c->res.draws, &c->res.uni, &c->res.recs, &c->res.vars, &c->res.keys, &c->res.vals, &c->res.sorted,
                      &c->res.clip, &c->res.status, &c->res.ranges, &c->res.ctrl, &c->res.binCounts, &c->res.binTotals, &c->res.order,
                      &c->res.items, &c->res.itemCount, &c->res.partDepth, &c->res.partRes, &c->res.partDone,
                      &c->res.positions, &c->res.vtxDraws}) release(*b);
    for (auto& e : c->ev) if (e) (void)hipEventDestroy(e);
    for (auto& f : c->fbos)
        for (DevBuf* b : {&f.gradDev, &f.smoothDev, &f.smoothUniDev}) release(*b);
    release(c->smoothBounds);
    for (auto& e : c->bwdEv) if (e) (void)hipEventDestroy(e);
    for (auto& e : c->gEv) if (e) (void)hipEventDestroy(e);
    if (c->hCtrl) (void)hipHostFree(c->hCtrl);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* sl_last_error(const sl_context* c) { return c ? c->err.c_str() : "null context"; }

unsigned sl_num_threads(const sl_context* c) { return c ? c->numThreads : 0u; }

unsigned sl_set_num_threads(sl_context* c, unsigned n)
{
    if (!c) return 0;
    c->numThreads = std::max(1u, n);   // the GPU path has no host worker pool; kept for API parity
    return c->numThreads;
}

int sl_set_raster_config(sl_context* c, int config)
{
    if (!c) return SL_ERR_INVALID_ARG;
    if (config < 0 || config >= kNumTileConfigs) return SL_ERR_INVALID_ARG;
    const int rc = flush(c);
    c->tileConfig = config;
    return rc;
}

int sl_set_raster_split(sl_context* c, uint32_t partRecords)
{
    if (!c) return SL_ERR_INVALID_ARG;
    const int rc = flush(c);
    if (rc != SL_OK) return rc;
    if (partRecords == SL_RASTER_SPLIT_DEFAULT)
    {
        c->splitRecs = kDefaultSplitRecords;
        c->splitAll = false;
        return SL_OK;
    }
    c->splitRecs = partRecords;
    c->splitAll = partRecords > 0;
    return SL_OK;
}

int sl_set_raster_occupancy(sl_context* c, uint32_t workgroupsPerCU)
{
    if (!c) return SL_ERR_INVALID_ARG;
    if (workgroupsPerCU > 32) return fail(c, SL_ERR_INVALID_ARG, "at most 32 raster workgroups per CU");
    const int rc = flush(c);
    if (rc != SL_OK) return rc;
    c->wgPerCU = workgroupsPerCU;
    return SL_OK;
}

int sl_set_raster_depth_reject(sl_context* c, int enable)
{
    if (!c) return SL_ERR_INVALID_ARG;
    const int rc = flush(c);
    if (rc != SL_OK) return rc;
    c->depthReject = enable != 0;
    return SL_OK;
}

int sl_get_raster_config(const sl_context* c, int config, int* tileW, int* tileH, int* segments)
{
    if (!c) return SL_ERR_INVALID_ARG;
    if (config < 0) config = c->tileConfig;
    if (config >= kNumTileConfigs) return SL_ERR_INVALID_ARG;
    if (tileW) *tileW = kTileConfigs[config].tileW;
    if (tileH) *tileH = kTileConfigs[config].tileH;
    if (segments) *segments = kTileConfigs[config].segs;
    return config;
}

int sl_set_viewport(sl_context* c, int32_t x, int32_t y, int32_t w, int32_t h)
{
    if (!c) return SL_ERR_INVALID_ARG;
    c->viewport[0] = std::min(std::max(x, -65536), 65535);   // SL_ViewportState.hpp:148-156
    c->viewport[1] = std::min(std::max(y, -65536), 65535);
    c->viewport[2] = std::min(w, 65535 - x);
    c->viewport[3] = std::min(h, 65535 - y);
    return SL_OK;
}

int sl_set_scissor(sl_context* c, int32_t x, int32_t y, int32_t w, int32_t h)
{
    if (!c) return SL_ERR_INVALID_ARG;
    c->scissor[0] = std::min(std::max(x, -65536), 65535);
    c->scissor[1] = std::min(std::max(y, -65536), 65535);
    c->scissor[2] = std::min(w, 65535 - x);
    c->scissor[3] = std::min(h, 65535 - y);
    return SL_OK;
}

size_t sl_create_vbo(sl_context* c) { c->vbos.emplace_back(); return c->vbos.size() - 1; }

int sl_vbo_init(sl_context* c, size_t vbo, size_t numBytes, const void* data)
{
    if (!c || vbo >= c->vbos.size()) return SL_ERR_INVALID_HANDLE;
    if (!numBytes) return -1;                                   // SL_VertexBuffer.cpp:98-101
    int rc = flush(c);
    if (rc != SL_OK) return rc;
    Vbo& v = c->vbos[vbo];
    if (v.valid)
    {
        // storage moves: the graph and the saved draws point at the old buffer
        if (c->graph.valid) { (void)graph_collect(c); graph_release(c); }
        stale_saved_draws(c);
    }
    if ((rc = check(c, hipStreamSynchronize(c->stream), "sync")) != SL_OK) return rc;
    release(v.buf);
    if ((rc = ensure(c, v.buf, numBytes)) != SL_OK) return rc;
    v.bytes = numBytes;
    v.valid = true;
    if (data) return upload(c, v.buf.p, data, numBytes, "vbo upload");
    return zero_fill(c, v.buf.p, numBytes, "vbo zero");
}

int sl_vbo_assign(sl_context* c, size_t vbo, const void* data, size_t offset, size_t numBytes)
{
    if (!c || vbo >= c->vbos.size() || !c->vbos[vbo].valid) return SL_ERR_INVALID_HANDLE;
    Vbo& v = c->vbos[vbo];
    if (offset + numBytes > v.bytes || !data) return SL_ERR_INVALID_ARG;
    int rc = flush(c);
    if (rc != SL_OK) return rc;
    // queued graph replays and backward passes may still read the buffer
    if ((rc = check(c, hipStreamSynchronize(c->stream), "sync")) != SL_OK) return rc;
    return upload(c, reinterpret_cast<uint8_t*>(v.buf.p) + offset, data, numBytes, "vbo assign");
}

int sl_vbo_read(sl_context* c, size_t vbo, void* out, size_t offset, size_t numBytes)
{
    if (!c || vbo >= c->vbos.size() || !c->vbos[vbo].valid) return SL_ERR_INVALID_HANDLE;
    Vbo& v = c->vbos[vbo];
    if (offset + numBytes > v.bytes || !out) return SL_ERR_INVALID_ARG;
    int rc = flush(c);
    if (rc != SL_OK) return rc;
    if ((rc = check(c, hipStreamSynchronize(c->stream), "sync")) != SL_OK) return rc;
    collect_backward(c);
    return download(c, out, reinterpret_cast<const uint8_t*>(v.buf.p) + offset, numBytes, "vbo read");
}

size_t sl_create_ibo(sl_context* c) { c->ibos.emplace_back(); return c->ibos.size() - 1; }

int sl_ibo_init(sl_context* c, size_t ibo, uint32_t numElements, int type, const void* data)
{
    if (!c || ibo >= c->ibos.size()) return SL_ERR_INVALID_HANDLE;
    if (!numElements) return -1;                                // SL_IndexBuffer.cpp:144-147
    const uint32_t bpe = bytes_per_index(type);
    if (!bpe) return SL_ERR_INVALID_ARG;
    int rc = flush(c);
    if (rc != SL_OK) return rc;
    Ibo& b = c->ibos[ibo];
    if (b.valid)
    {
        if (c->graph.valid) { (void)graph_collect(c); graph_release(c); }
        stale_saved_draws(c);
    }
    if ((rc = check(c, hipStreamSynchronize(c->stream), "sync")) != SL_OK) return rc;
    release(b.buf);
    const size_t bytes = (size_t)numElements * bpe;
    if ((rc = ensure(c, b.buf, bytes + 16)) != SL_OK) return rc;   // padded like SL_IndexBuffer.cpp:155
    if ((rc = zero_fill(c, b.buf.p, bytes + 16, "ibo zero")) != SL_OK) return rc;
    if (data && (rc = upload(c, b.buf.p, data, bytes, "ibo upload")) != SL_OK) return rc;
    b.type = type;
    b.bpe = bpe;
    b.count = numElements;
    b.valid = true;
    return SL_OK;
}

size_t sl_create_vao(sl_context* c) { c->vaos.emplace_back(); return c->vaos.size() - 1; }

int sl_vao_set_vertex_buffer(sl_context* c, size_t vao, size_t vbo)
{
    if (!c || vao >= c->vaos.size()) return SL_ERR_INVALID_HANDLE;
    c->vaos[vao].vbo = vbo;
    return SL_OK;
}

int sl_vao_set_index_buffer(sl_context* c, size_t vao, size_t ibo)
{
    if (!c || vao >= c->vaos.size()) return SL_ERR_INVALID_HANDLE;
    c->vaos[vao].ibo = ibo;
    return SL_OK;
}

int sl_vao_set_num_bindings(sl_context* c, size_t vao, size_t n)
{
    if (!c || vao >= c->vaos.size()) return SL_ERR_INVALID_HANDLE;
    if (n >= SL_VAO_MAX_BINDINGS) return -1;                    // SL_VertexArray.cpp:164-167
    Vao& v = c->vaos[vao];
    for (size_t i = v.nb; i < n; ++i) v.b[i] = Binding{};
    v.nb = n;
    return (int)n;
}

int sl_vao_set_binding(sl_context* c, size_t vao, size_t bindId, ptrdiff_t offset, ptrdiff_t stride, int dims, int type)
{
    if (!c || vao >= c->vaos.size()) return SL_ERR_INVALID_HANDLE;
    Vao& v = c->vaos[vao];
    if (bindId >= v.nb) return SL_ERR_INVALID_ARG;
    if (dims < VERTEX_DIMENSION_1 || dims > VERTEX_DIMENSION_4) return SL_ERR_INVALID_ARG;
    v.b[bindId].offset = offset;
    v.b[bindId].stride = stride;
    v.b[bindId].dims = dims;
    v.b[bindId].type = type;
    return SL_OK;
}

size_t sl_create_texture(sl_context* c) { c->texs.emplace_back(); return c->texs.size() - 1; }

int sl_texture_init(sl_context* c, size_t tex, int type, uint16_t w, uint16_t h, uint16_t d)
{
    if (!c || tex >= c->texs.size()) return SL_ERR_INVALID_HANDLE;
    const uint32_t bpt = bytes_per_color(type);
    if (!bpt) return SL_ERR_UNSUPPORTED;
    if (!w || !h || !d) return SL_ERR_INVALID_ARG;
    int rc = flush(c);
    if (rc != SL_OK) return rc;
    Tex& t = c->texs[tex];
    if (t.valid)
    {
        if (c->graph.valid) { (void)graph_collect(c); graph_release(c); }
        stale_saved_draws(c);
    }
    if ((rc = check(c, hipStreamSynchronize(c->stream), "sync")) != SL_OK) return rc;
    release(t.buf);
    // SL_Texture.cpp:28-55: each dimension padded by (4 - dim % 4), zero filled,
    // so the bilinear sampler's x = W / y = H neighbours read zeros.
    const size_t pw = (size_t)w + (4 - w % 4), ph = (size_t)h + (4 - h % 4), pd = (size_t)d + (4 - d % 4);
    const size_t bytes = pw * ph * pd * bpt + 4 * bpt;
    if ((rc = ensure(c, t.buf, bytes)) != SL_OK) return rc;
    if ((rc = zero_fill(c, t.buf.p, bytes, "texture zero")) != SL_OK) return rc;
    t.type = type; t.w = w; t.h = h; t.d = d; t.bpt = bpt;
    t.valid = true;
    return SL_OK;
}

int sl_texture_set_texels(sl_context* c, size_t tex, const void* data, size_t numBytes)
{
    if (!c) return SL_ERR_INVALID_ARG;
    Tex* t = tex_at(c, tex);
    if (!t) return SL_ERR_INVALID_HANDLE;
    const size_t n = (size_t)t->w * t->h * t->d * t->bpt;
    if (numBytes != n || !data) return SL_ERR_INVALID_ARG;
    int rc = flush(c);
    if (rc != SL_OK) return rc;
    if ((rc = check(c, hipStreamSynchronize(c->stream), "sync")) != SL_OK) return rc;
    return upload(c, t->buf.p, data, n, "texture upload");
}

int sl_texture_read(sl_context* c, size_t tex, void* out, size_t numBytes)
{
    if (!c) return SL_ERR_INVALID_ARG;
    Tex* t = tex_at(c, tex);
    if (!t) return SL_ERR_INVALID_HANDLE;
    const size_t n = (size_t)t->w * t->h * t->d * t->bpt;
    if (numBytes != n || !out) return SL_ERR_INVALID_ARG;
    int rc = flush(c);
    if (rc != SL_OK) return rc;
    if ((rc = check(c, hipStreamSynchronize(c->stream), "sync")) != SL_OK) return rc;
    collect_backward(c);
    return download(c, out, t->buf.p, n, "texture readback");
}

size_t sl_create_framebuffer(sl_context* c) { c->fbos.emplace_back(); return c->fbos.size() - 1; }

int sl_fbo_reserve_color_buffers(sl_context* c, size_t fbo, unsigned n)
{
    if (!c || fbo >= c->fbos.size()) return SL_ERR_INVALID_HANDLE;
    if (n > SL_SHADER_MAX_FRAG_OUTPUTS) return SL_ERR_INVALID_ARG;
    int rc = flush(c);
    if (rc != SL_OK) return rc;
    c->fbos[fbo].colors.assign(n, SL_INVALID_ID);
    return SL_OK;
}

int sl_fbo_attach_color_buffer(sl_context* c, size_t fbo, unsigned idx, size_t tex)
{
    if (!c || fbo >= c->fbos.size() || tex >= c->texs.size()) return SL_ERR_INVALID_HANDLE;
    Fbo& f = c->fbos[fbo];
    if (idx >= f.colors.size()) return SL_ERR_INVALID_ARG;
    int rc = flush(c);
    if (rc != SL_OK) return rc;
    f.colors[idx] = tex;
    return SL_OK;
}

int sl_fbo_attach_depth_buffer(sl_context* c, size_t fbo, size_t tex)
{
    if (!c || fbo >= c->fbos.size() || tex >= c->texs.size()) return SL_ERR_INVALID_HANDLE;
    int rc = flush(c);
    if (rc != SL_OK) return rc;
    c->fbos[fbo].depth = tex;
    return SL_OK;
}

int sl_fbo_attach_debug_buffers(sl_context* c, size_t fbo, size_t lastTex, size_t countTex)
{
    if (!c || fbo >= c->fbos.size()) return SL_ERR_INVALID_HANDLE;
    for (size_t t : {lastTex, countTex})
        if (t != SL_INVALID_ID && (t >= c->texs.size() || c->texs[t].type != SL_COLOR_R_32U)) return SL_ERR_INVALID_ARG;
    int rc = flush(c);
    if (rc != SL_OK) return rc;
    c->fbos[fbo].dbgLast = lastTex;
    c->fbos[fbo].dbgCount = countTex;
    return SL_OK;
}

int sl_fbo_valid(sl_context* c, size_t fbo)
{
    if (!c || fbo >= c->fbos.size()) return SL_ERR_INVALID_HANDLE;
    const Fbo& f = c->fbos[fbo];
    if (f.colors.empty()) return -1;                            // SL_Framebuffer.cpp:930-990
    for (size_t t : f.colors) if (!tex_at(c, t)) return -2;
    const Tex& c0 = c->texs[f.colors[0]];
    for (size_t t : f.colors)
    {
        if (c->texs[t].w != c0.w) return -3;
        if (c->texs[t].h != c0.h) return -4;
        if (c->texs[t].d != c0.d) return -5;
    }
    const Tex* dt = tex_at(c, f.depth);
    if (!dt) return -6;
    if (dt->w != c0.w) return -7;
    if (dt->h != c0.h) return -8;
    if (dt->d > 1) return -9;
    if (dt->type != SL_COLOR_R_HALF && dt->type != SL_COLOR_R_FLOAT && dt->type != SL_COLOR_R_DOUBLE) return -10;
    return 0;
}

size_t sl_create_ubo(sl_context* c)
{
    c->ubos.emplace_back();
    std::memset(c->ubos.back().bytes, 0, SL_MAX_UNIFORM_BUFFER_SIZE);
    return c->ubos.size() - 1;
}

int sl_ubo_assign(sl_context* c, size_t ubo, const void* data, size_t offset, size_t n)
{
    if (!c || ubo >= c->ubos.size()) return SL_ERR_INVALID_HANDLE;
    if (offset + n > SL_MAX_UNIFORM_BUFFER_SIZE || !data) return SL_ERR_INVALID_ARG;
    std::memcpy(c->ubos[ubo].bytes + offset, data, n);         // snapshotted by later draws
    return SL_OK;
}

size_t sl_create_shader(sl_context* c, const sl_vertex_shader* vs, const sl_fragment_shader* fs, size_t ubo)
{
    // SL_Context::create_shader (SL_Context.cpp:547-604)
    if (!c || !vs || !fs) return SL_INVALID_ID;
    if (ubo != SL_INVALID_ID && ubo >= c->ubos.size()) return SL_INVALID_ID;
    if (vs->numVaryings < fs->numVaryings) return SL_INVALID_ID;
    if (vs->numVaryings > SL_SHADER_MAX_VARYING_VECTORS) return SL_INVALID_ID;
    if (fs->numVaryings > SL_SHADER_MAX_VARYING_VECTORS) return SL_INVALID_ID;
    if (fs->numOutputs > SL_SHADER_MAX_FRAG_OUTPUTS) return SL_INVALID_ID;
    if (vs->cullMode > SL_CULL_OFF || fs->depthTest > SL_DEPTH_TEST_NOT_EQUAL || fs->blend > SL_BLEND_SCREEN) return SL_INVALID_ID;
    Shader s;
    s.cull = vs->cullMode;
    s.dtest = fs->depthTest;
    s.dmask = fs->depthMask;
    s.blend = fs->blend;
    s.nvary = fs->numVaryings;
    s.nout = fs->numOutputs;
    s.vsProg = vs->programId;
    s.fsProg = fs->programId;
    s.ubo = ubo;
    c->shaders.push_back(s);
    return c->shaders.size() - 1;
}

int sl_draw(sl_context* c, const sl_mesh* m, size_t shader, size_t fbo)
{
    if (!c || !m) return SL_ERR_INVALID_ARG;
    return record_draw(c, *m, 1, shader, fbo, m->mode);
}

int sl_draw_multiple(sl_context* c, const sl_mesh* m, size_t n, size_t shader, size_t fbo)
{
    if (!c) return SL_ERR_INVALID_ARG;
    if (m == nullptr || n == 0) return SL_OK;                   // SL_Context.cpp:726-730
    for (size_t i = 0; i < n; ++i)
    {
        const int rc = record_draw(c, m[i], 1, shader, fbo, m[0].mode);
        if (rc != SL_OK) return rc;
    }
    return SL_OK;
}

int sl_draw_instanced(sl_context* c, const sl_mesh* m, size_t numInstances, size_t shader, size_t fbo)
{
    if (!c || !m) return SL_ERR_INVALID_ARG;
    if (numInstances == 0) return SL_OK;
    return record_draw(c, *m, numInstances, shader, fbo, m->mode);
}

int sl_draw_static(sl_context* c, const sl_mesh* m, size_t shaderId, size_t fboId)
{
    if (!c || !m) return SL_ERR_INVALID_ARG;
    // a pending batch for another framebuffer runs first, through the array too
    if (!c->batch.empty() && c->batch.fbo != fboId)
    {
        const int rc = flush(c, true);
        if (rc != SL_OK) return rc;
    }
    int rc = record_draw(c, *m, 1, shaderId, fboId, (int)m->mode);
    if (rc != SL_OK)
    {
        (void)flush(c, true);
        return rc;
    }
    return flush(c, true);
}

int sl_clear_color_buffer(sl_context* c, size_t fbo, unsigned att, const double rgba[4])
{
    if (!c || !rgba) return SL_ERR_INVALID_ARG;
    return record_clear(c, fbo, (int)att, rgba, false, 0.0);
}

int sl_clear_depth_buffer(sl_context* c, size_t fbo, double depth)
{
    if (!c) return SL_ERR_INVALID_ARG;
    return record_clear(c, fbo, -1, nullptr, true, depth);
}

int sl_clear_framebuffer(sl_context* c, size_t fbo, unsigned att, const double rgba[4], double depth)
{
    if (!c || !rgba) return SL_ERR_INVALID_ARG;
    const int rc = record_clear(c, fbo, (int)att, rgba, att == 0, depth);
    if (rc != SL_OK || att == 0) return rc;
    return record_clear(c, fbo, -1, nullptr, true, depth);
}

int sl_finish(sl_context* c)
{
    if (!c) return SL_ERR_INVALID_ARG;
    int rc = flush(c);
    if (rc != SL_OK) return rc;
    rc = graph_collect(c);
    if (rc != SL_OK) return rc;
    rc = check(c, hipStreamSynchronize(c->stream), "finish");
    collect_backward(c);
    return rc;
}

int sl_frame_graph_capture(sl_context* c)
{
    if (!c) return SL_ERR_INVALID_ARG;
    int rc = graph_collect(c);
    if (rc != SL_OK) return rc;
    return graph_capture(c);
}

int sl_frame_graph_launch(sl_context* c)
{
    if (!c) return SL_ERR_INVALID_ARG;
    if (!c->graph.valid) return fail(c, SL_ERR_STATE, "no captured frame graph");
    int rc = flush(c);
    if (rc != SL_OK) return rc;
    if (c->graph.pending == 0) (void)hipEventRecord(c->ev[6], c->stream);
    rc = check(c, hipGraphLaunch(c->graph.exec, c->stream), "graph launch");
    if (rc != SL_OK) return rc;
    (void)hipEventRecord(c->ev[7], c->stream);
    c->graph.pending += 1;
    FrameGraph& G = c->graph;
    if (G.fbo < c->fbos.size())
    {
        Fbo& F = c->fbos[G.fbo];
        if (tex_at(c, F.vis))
        {
            if (G.clearDepth)
            {
                F.saved = G.draws;
                F.savedStale = false;
                F.savedBlend = false;
                F.savedFromClear = true;
                F.savedClearDepth = G.depth;
                for (const SavedDraw& d : G.draws)
                    if (d.d.blend != SL_BLEND_OFF) F.savedBlend = true;
            }
            else
            {
                F.savedStale = true;   // replays without a depth clear stack up
            }
            mark_visibility(c, G.fbo);
        }
    }
    return SL_OK;
}

int sl_frame_graph_release(sl_context* c)
{
    if (!c) return SL_ERR_INVALID_ARG;
    const int rc = graph_collect(c);
    graph_release(c);
    return rc;
}

// ---------------------------------------------------------------------------
// blit
// ---------------------------------------------------------------------------
int sl_blit_region(sl_context* c, size_t outTex, size_t inTex, uint16_t srcX0, uint16_t srcY0, uint16_t srcX1,
                   uint16_t srcY1, uint16_t dstX0, uint16_t dstY0, uint16_t dstX1, uint16_t dstY1)
{
    if (!c) return SL_ERR_INVALID_ARG;
    Tex* o = tex_at(c, outTex);
    Tex* i = tex_at(c, inTex);
    if (!o || !i) return SL_ERR_INVALID_HANDLE;
    if (i->type != SL_COLOR_RGBA_8U && i->type != SL_COLOR_RGBA_FLOAT && i->type != SL_COLOR_R_FLOAT)
        return fail(c, SL_ERR_UNSUPPORTED, "blit source format");
    if (o->type != SL_COLOR_RGBA_8U && o->type != SL_COLOR_RGBA_FLOAT) return fail(c, SL_ERR_UNSUPPORTED, "blit destination format");
    if (o == i) return fail(c, SL_ERR_INVALID_ARG, "blit onto its own source");
    int rc = flush(c);
    if (rc != SL_OK) return rc;
    // SL_BlitProcessor::blit_nearest (SL_BlitProcessor.cpp:620-665), in the
    // reference's 64-bit unsigned arithmetic
    const unsigned long long inW = (unsigned long long)srcX1 - srcX0, inH = (unsigned long long)srcY1 - srcY0;
    const unsigned long long outW = (unsigned long long)dstX1 - dstX0;
    const unsigned long long totalW = o->w, totalH = o->h;
    BlitParams bp{};
    bp.src = reinterpret_cast<const uint8_t*>(i->buf.p);
    bp.dst = reinterpret_cast<uint8_t*>(o->buf.p);
    bp.srcW = i->w; bp.srcH = i->h; bp.dstW = o->w; bp.dstH = o->h;
    bp.srcX0 = srcX0; bp.srcY0 = srcY0; bp.srcX1 = srcX1; bp.srcY1 = srcY1;
    const unsigned long long x0 = dstX0, x1 = std::min<unsigned long long>(totalW, x0 + outW);
    const unsigned long long y0 = dstY0, y1 = std::min<unsigned long long>(totalH, dstY1);   // rows past the height are skipped
    bp.x0 = (uint32_t)x0; bp.x1 = (uint32_t)std::max(x0, x1);
    bp.y0 = (uint32_t)y0; bp.y1 = (uint32_t)std::max(y0, y1);
    bp.foutW = ((inW << 16) / totalW) + 1ull;
    bp.foutH = ((inH << 16) / totalH) + 1ull;
    return check(c, launch_blit(bp, i->type, o->type, c->stream), "k_blit");
}

int sl_blit(sl_context* c, size_t outTex, size_t inTex)
{
    if (!c) return SL_ERR_INVALID_ARG;
    Tex* o = tex_at(c, outTex);
    Tex* i = tex_at(c, inTex);
    if (!o || !i) return SL_ERR_INVALID_HANDLE;
    return sl_blit_region(c, outTex, inTex, 0, 0, i->w, i->h, 0, 0, o->w, o->h);   // SL_Context.cpp:747-769
}

// ---------------------------------------------------------------------------
// differentiable draws
// ---------------------------------------------------------------------------
int sl_framebuffer_attach_visibility(sl_context* c, size_t fbo, size_t tex)
{
    if (!c || fbo >= c->fbos.size()) return SL_ERR_INVALID_HANDLE;
    if (tex != SL_INVALID_ID)
    {
        Tex* t = tex_at(c, tex);
        if (!t) return SL_ERR_INVALID_HANDLE;
        if (t->type != SL_COLOR_RGBA_FLOAT) return fail(c, SL_ERR_INVALID_ARG, "visibility attachment must be RGBA_FLOAT");
    }
    int rc = flush(c);
    if (rc != SL_OK) return rc;
    Fbo& F = c->fbos[fbo];
    F.vis = tex;
    F.saved.clear();
    F.savedStale = false;
    F.savedBlend = false;
    F.savedFromClear = false;
    F.visGen = 0;
    return SL_OK;
}

int sl_framebuffer_attach_positions(sl_context* c, size_t fbo, size_t gradVbo, size_t offsetVbo, uint32_t numVertices)
{
    if (!c || fbo >= c->fbos.size()) return SL_ERR_INVALID_HANDLE;
    if (gradVbo != SL_INVALID_ID && (gradVbo >= c->vbos.size() || !c->vbos[gradVbo].valid)) return SL_ERR_INVALID_HANDLE;
    if (offsetVbo != SL_INVALID_ID && (offsetVbo >= c->vbos.size() || !c->vbos[offsetVbo].valid)) return SL_ERR_INVALID_HANDLE;
    if (gradVbo != SL_INVALID_ID && gradVbo == offsetVbo) return fail(c, SL_ERR_INVALID_ARG, "position gradients and offsets need separate buffers");
    int rc = flush(c);
    if (rc != SL_OK) return rc;
    Fbo& F = c->fbos[fbo];
    F.posGrad = gradVbo;
    F.posOffset = offsetVbo;
    F.posVerts = numVertices;
    return SL_OK;
}

// Visibility recomputed instead of kept: framebuffers may share one
// visibility attachment (one W x H buffer for any number of them). When the
// attachment holds another framebuffer's draws, the backward passes of this
// one first draw its saved draws again from their depth clear: the same
// draws, depth tests and sequence numbers give the same visibility, depth and
// (unblended draws over the finished image) colour, bit for bit. Call after
// flush(); the rebuild is executed before returning.
int rebuild_visibility(sl_context* c, size_t fbo)
{
    Fbo& F = c->fbos[fbo];
    Tex* V = tex_at(c, F.vis);
    if (!V || F.saved.empty()) return SL_OK;
    if (V->visFbo == fbo && V->visGen == F.visGen) return SL_OK;
    if (!F.savedFromClear)
        return fail(c, SL_ERR_STATE, "the visibility attachment holds another framebuffer's draws and this framebuffer's draws "
                                     "do not start from a depth clear: draw again");
    if (F.savedBlend)
        return fail(c, SL_ERR_UNSUPPORTED, "the visibility attachment holds another framebuffer's draws and blended draws cannot "
                                           "be redrawn over the finished image: give the framebuffer its own attachment");
    const std::vector<SavedDraw> saved = F.saved;   // the batch's execution records them again
    Batch& B = c->batch;
    B.reset();
    B.fbo = fbo;
    B.clearDepth = true;
    B.depth = F.savedClearDepth;
    for (const SavedDraw& sd : saved)
    {
        DrawDesc d = sd.d;
        // the draws may come from several batches: lay them out as one
        B.totalTris = (B.totalTris + kSetupBlock - 1) / kSetupBlock * kSetupBlock;
        d.triBase = B.totalTris;
        d.posBase = kNoPositions;   // the position program per corner: the same bits as the vertex pass
        if (d.hasUniforms)
        {
            if (sd.uni.size() != SL_MAX_UNIFORM_BUFFER_SIZE) return fail(c, SL_ERR_STATE, "saved draw without its uniform snapshot");
            d.uniOffset = (uint32_t)B.uni.size();
            B.uni.insert(B.uni.end(), sd.uni.begin(), sd.uni.end());
        }
        B.draws.push_back(d);
        B.drawVbo.push_back(sd.vbo);
        B.drawTex.push_back(sd.tex);
        B.totalTris += d.numTris;
        B.maxVary = std::max<uint32_t>(B.maxVary, d.nvary);
    }
    c->stats.visibility_rebuilds += 1;
    return flush(c);
}

// The smoothing passes' draw table (uniform offsets into the saved draws'
// snapshots) and arguments; uploads the table when it changed.
int smooth_setup(sl_context* c, size_t fbo, size_t srcTex, size_t dstTex, float sigma, bool backward, SmoothParams& sp)
{
    if (!c || fbo >= c->fbos.size()) return SL_ERR_INVALID_HANDLE;
    if (!(sigma >= 0.f) || !std::isfinite(sigma)) return fail(c, SL_ERR_INVALID_ARG, "sigma must be finite and >= 0");
    int rc = flush(c);
    if (rc != SL_OK) return rc;
    Fbo& F = c->fbos[fbo];
    Tex* vis = tex_at(c, F.vis);
    if (!vis) return fail(c, SL_ERR_STATE, "framebuffer has no visibility attachment");
    Tex* col = F.colors.empty() ? nullptr : tex_at(c, F.colors[0]);
    if (!col || col->type != SL_COLOR_RGBA_FLOAT) return fail(c, SL_ERR_STATE, "smoothed coverage needs an RGBA_FLOAT colour attachment 0");
    Tex* src = tex_at(c, srcTex);
    Tex* dst = tex_at(c, dstTex);
    if (!src || !dst) return SL_ERR_INVALID_HANDLE;
    for (const Tex* t : {src, dst})
        if (t->type != SL_COLOR_RGBA_FLOAT || t->w != vis->w || t->h != vis->h)
            return fail(c, SL_ERR_INVALID_ARG, "smoothing images must be RGBA_FLOAT of the framebuffer's size");
    if (src == dst) return fail(c, SL_ERR_INVALID_ARG, "smoothing input and output must differ");
    if (F.savedStale) return fail(c, SL_ERR_STATE, "resources changed since the forward pass; draw again");
    if ((rc = rebuild_visibility(c, fbo)) != SL_OK) return rc;
    const size_t nd = F.saved.size();
    const uint64_t need = (uint64_t)nd * F.posVerts * 8u;
    Vbo* pg = (backward && F.posGrad != SL_INVALID_ID) ? &c->vbos[F.posGrad] : nullptr;
    Vbo* po = (F.posOffset != SL_INVALID_ID) ? &c->vbos[F.posOffset] : nullptr;
    if ((pg && (!pg->valid || pg->bytes < need)) || (po && (!po->valid || po->bytes < need)))
        return fail(c, SL_ERR_INVALID_ARG, "position buffers need 8 bytes per saved draw and vertex");
    std::vector<GradDesc> G(nd);
    std::vector<uint8_t> U;
    for (size_t i = 0; i < nd; ++i)
    {
        std::memset(&G[i], 0, sizeof(GradDesc));
        G[i].d = F.saved[i].d;
        G[i].d.uniOffset = (uint32_t)U.size();
        const std::vector<uint8_t>& u = F.saved[i].uni;
        U.insert(U.end(), u.begin(), u.end());
        if (u.empty()) U.resize(U.size() + SL_MAX_UNIFORM_BUFFER_SIZE, 0);
    }
    if (nd > 0 && (F.smoothHost.size() != nd || F.smoothUniHost != U ||
                   std::memcmp(F.smoothHost.data(), G.data(), sizeof(GradDesc) * nd) != 0))
    {
        if ((rc = ensure(c, F.smoothDev, sizeof(GradDesc) * nd)) != SL_OK) return rc;
        if ((rc = ensure(c, F.smoothUniDev, U.size())) != SL_OK) return rc;
        if ((rc = upload(c, F.smoothDev.p, G.data(), sizeof(GradDesc) * nd, "upload smoothing draws")) != SL_OK) return rc;
        if ((rc = upload(c, F.smoothUniDev.p, U.data(), U.size(), "upload smoothing uniforms")) != SL_OK) return rc;
        F.smoothHost = G;
        F.smoothUniHost = U;
    }
    sp = SmoothParams{};
    sp.draws = reinterpret_cast<const GradDesc*>(F.smoothDev.p);
    sp.numDraws = (uint32_t)nd;
    sp.uniforms = reinterpret_cast<const uint8_t*>(F.smoothUniDev.p);
    sp.color = reinterpret_cast<const float4*>(col->buf.p);
    sp.vis = reinterpret_cast<const float4*>(vis->buf.p);
    sp.grad = backward ? reinterpret_cast<const float4*>(src->buf.p) : nullptr;
    sp.out = reinterpret_cast<float4*>(dst->buf.p);
    sp.posGrad = pg ? reinterpret_cast<float2*>(pg->buf.p) : nullptr;
    sp.posOffset = po ? reinterpret_cast<const float2*>(po->buf.p) : nullptr;
    sp.numVerts = F.posVerts;
    sp.W = vis->w;
    sp.H = vis->h;
    sp.sigma = sigma;
    if (nd == 0) sp.sigma = 0.f;   // nothing drawn: no boundaries
    if (sp.sigma > 0.f)
    {
        if ((rc = ensure(c, c->smoothBounds, sizeof(float2) * (size_t)vis->w * vis->h)) != SL_OK) return rc;
        sp.bounds = reinterpret_cast<float2*>(c->smoothBounds.p);
    }
    return SL_OK;
}

int sl_smooth_coverage(sl_context* c, size_t fbo, size_t outTex, float sigma)
{
    if (!c || fbo >= c->fbos.size()) return SL_ERR_INVALID_HANDLE;
    const size_t colId = c->fbos[fbo].colors.empty() ? SL_INVALID_ID : c->fbos[fbo].colors[0];
    SmoothParams sp;
    int rc = smooth_setup(c, fbo, colId, outTex, sigma, false, sp);
    if (rc != SL_OK) return rc;
    return check(c, launch_smooth(sp, c->stream), "k_smooth");
}

int sl_smooth_coverage_backward(sl_context* c, size_t fbo, size_t gradSmoothTex, float sigma, size_t gradImageTex)
{
    SmoothParams sp;
    int rc = smooth_setup(c, fbo, gradSmoothTex, gradImageTex, sigma, true, sp);
    if (rc != SL_OK) return rc;
    return check(c, launch_smooth_backward(sp, c->stream), "k_smooth_backward");
}

int sl_vbo_attach_gradient(sl_context* c, size_t vbo, size_t grad)
{
    if (!c || vbo >= c->vbos.size()) return SL_ERR_INVALID_HANDLE;
    if (grad != SL_INVALID_ID && (grad >= c->vbos.size() || grad == vbo)) return SL_ERR_INVALID_HANDLE;
    int rc = flush(c);
    if (rc != SL_OK) return rc;
    c->vbos[vbo].grad = grad;
    return SL_OK;
}

int sl_texture_attach_gradient(sl_context* c, size_t tex, size_t grad)
{
    if (!c || tex >= c->texs.size()) return SL_ERR_INVALID_HANDLE;
    if (grad != SL_INVALID_ID)
    {
        if (grad >= c->texs.size() || grad == tex) return SL_ERR_INVALID_HANDLE;
        const Tex* g = tex_at(c, grad);
        const Tex& t = c->texs[tex];
        if (!g || !t.valid) return SL_ERR_INVALID_HANDLE;
        if (g->type != SL_COLOR_RGBA_FLOAT || g->w != t.w || g->h != t.h || g->d != t.d)
            return fail(c, SL_ERR_INVALID_ARG, "texture gradient must be RGBA_FLOAT of the texture's size");
    }
    int rc = flush(c);
    if (rc != SL_OK) return rc;
    c->texs[tex].grad = grad;
    return SL_OK;
}

int sl_gradients_zero(sl_context* c)
{
    if (!c) return SL_ERR_INVALID_ARG;
    int rc = flush(c);
    if (rc != SL_OK) return rc;
    for (const Vbo& v : c->vbos)
    {
        if (!v.valid || v.grad >= c->vbos.size() || !c->vbos[v.grad].valid) continue;
        const Vbo& g = c->vbos[v.grad];
        if ((rc = check(c, launch_fill_u32(reinterpret_cast<uint32_t*>(g.buf.p), 0u, g.bytes / 4, c->stream), "zero gradient")) != SL_OK) return rc;
    }
    for (const Tex& t : c->texs)
    {
        if (!t.valid || t.grad >= c->texs.size() || !c->texs[t.grad].valid) continue;
        const Tex& g = c->texs[t.grad];
        if ((rc = check(c, launch_fill_u32(reinterpret_cast<uint32_t*>(g.buf.p), 0u, g.buf.bytes / 4, c->stream), "zero gradient")) != SL_OK) return rc;
    }
    for (const Fbo& F : c->fbos)
    {
        if (F.posGrad >= c->vbos.size() || !c->vbos[F.posGrad].valid) continue;
        const Vbo& g = c->vbos[F.posGrad];
        if ((rc = check(c, launch_fill_u32(reinterpret_cast<uint32_t*>(g.buf.p), 0u, g.bytes / 4, c->stream), "zero gradient")) != SL_OK) return rc;
    }
    return SL_OK;
}

int sl_loss_l2_grad(sl_context* c, size_t colorTex, size_t targetTex, size_t gradTex, float scale)
{
    if (!c) return SL_ERR_INVALID_ARG;
    Tex* a = tex_at(c, colorTex);
    Tex* b = tex_at(c, targetTex);
    Tex* g = tex_at(c, gradTex);
    if (!a || !b || !g) return SL_ERR_INVALID_HANDLE;
    for (const Tex* t : {a, b, g})
        if (t->type != SL_COLOR_RGBA_FLOAT || t->w != a->w || t->h != a->h || t->d != a->d)
            return fail(c, SL_ERR_INVALID_ARG, "loss textures must be RGBA_FLOAT of one size");
    int rc = flush(c);
    if (rc != SL_OK) return rc;
    const uint64_t n = (uint64_t)a->w * a->h * std::max<uint16_t>(a->d, 1);
    return check(c, launch_loss_l2_grad(reinterpret_cast<const float4*>(a->buf.p), reinterpret_cast<const float4*>(b->buf.p),
                                        reinterpret_cast<float4*>(g->buf.p), n, scale, c->stream), "loss gradient");
}

int sl_framebuffer_backward(sl_context* c, size_t fbo, size_t gradImageTex)
{
    if (!c || fbo >= c->fbos.size()) return SL_ERR_INVALID_HANDLE;
    int rc = flush(c);
    if (rc != SL_OK) return rc;
    Fbo& F = c->fbos[fbo];
    Tex* vis = tex_at(c, F.vis);
    if (!vis) return fail(c, SL_ERR_STATE, "framebuffer has no visibility attachment");
    Tex* gi = tex_at(c, gradImageTex);
    if (!gi) return SL_ERR_INVALID_HANDLE;
    if (gi->type != SL_COLOR_RGBA_FLOAT || gi->w != vis->w || gi->h != vis->h)
        return fail(c, SL_ERR_INVALID_ARG, "image gradient must be RGBA_FLOAT of the framebuffer's size");
    if (F.savedStale) return fail(c, SL_ERR_STATE, "resources changed since the forward pass; draw again");
    if (F.savedBlend) return fail(c, SL_ERR_UNSUPPORTED, "fixed-visibility backward of blended draws");
    if (F.saved.empty()) return SL_OK;
    if ((rc = rebuild_visibility(c, fbo)) != SL_OK) return rc;
    std::vector<GradDesc> G(F.saved.size());
    for (size_t i = 0; i < F.saved.size(); ++i)
    {
        const SavedDraw& sd = F.saved[i];
        const bool ct = sd.d.vsProg == SL_PROGRAM_COLOR_TEXTURED && sd.d.fsProg == SL_PROGRAM_COLOR_TEXTURED;
        const bool pg = sd.d.vsProg == SL_PROGRAM_PASSTHROUGH_GOURAUD && sd.d.fsProg == SL_PROGRAM_PASSTHROUGH_GOURAUD;
        if (!ct && !pg) return fail(c, SL_ERR_UNSUPPORTED, "backward of this shader program");
        GradDesc& g = G[i];
        std::memset(&g, 0, sizeof(g));
        g.d = sd.d;
        if (sd.vbo < c->vbos.size())
        {
            const size_t gv = c->vbos[sd.vbo].grad;
            if (gv < c->vbos.size() && c->vbos[gv].valid)
            {
                g.gradCol = reinterpret_cast<float*>(c->vbos[gv].buf.p);
                g.gradColCount = c->vbos[gv].bytes / 16;
            }
        }
        if (ct && sd.tex < c->texs.size())
        {
            const Tex& t = c->texs[sd.tex];
            const Tex* gt = tex_at(c, t.grad);
            if (gt && gt->type == SL_COLOR_RGBA_FLOAT && gt->w == t.w && gt->h == t.h)
            {
                g.gradTex = reinterpret_cast<float*>(gt->buf.p);
                g.gradTexTexels = gt->buf.bytes / 16;
            }
        }
    }
    // the table only changes with the draws or the gradient attachments;
    // re-upload (synchronously, after all queued work) only then
    hipStream_t s = c->stream;
    const bool same = F.gradHost.size() == G.size() && F.gradDev.p &&
                      std::memcmp(F.gradHost.data(), G.data(), sizeof(GradDesc) * G.size()) == 0;
    if (!same)
    {
        if ((rc = ensure(c, F.gradDev, sizeof(GradDesc) * G.size())) != SL_OK) return rc;
        if ((rc = check(c, hipStreamSynchronize(s), "sync")) != SL_OK) return rc;
        if ((rc = upload(c, F.gradDev.p, G.data(), sizeof(GradDesc) * G.size(), "upload backward draws")) != SL_OK) return rc;
        F.gradHost = G;
    }
    BackwardParams bp{};
    bp.draws = reinterpret_cast<const GradDesc*>(F.gradDev.p);
    bp.numDraws = (uint32_t)G.size();
    bp.vis = reinterpret_cast<const float4*>(vis->buf.p);
    bp.gradImage = reinterpret_cast<const float4*>(gi->buf.p);
    bp.W = vis->w;
    bp.H = vis->h;
    if (c->bwdPending == 0) (void)hipEventRecord(c->bwdEv[0], s);
    if ((rc = check(c, launch_backward(bp, s), "k_backward")) != SL_OK) return rc;
    (void)hipEventRecord(c->bwdEv[1], s);
    c->bwdPending += 1;
    return SL_OK;
}

int sl_get_stats(const sl_context* c, sl_frame_stats* out)
{
    if (!c || !out) return SL_ERR_INVALID_ARG;
    *out = c->stats;
    return SL_OK;
}

int sl_reset_stats(sl_context* c)
{
    if (!c) return SL_ERR_INVALID_ARG;
    const int rc = flush(c);
    c->stats = sl_frame_stats{};
    c->triSeq = 0;
    return rc;
}

}  // extern "C"
