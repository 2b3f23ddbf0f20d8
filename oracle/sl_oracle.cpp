/*
 * sl_oracle.cpp -- CPU restatement of SoftLight's triangle draw path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the MI355X
 * pipeline in softlight_amd/: tests, the smoke check and the benchmark's CPU
 * baseline load it; the product never does.  It exports the same C API as
 * include/sl_gpu.h with the prefix `slo_` instead of `sl_`.
 *
 * It restates the reference's portable scalar path (the `#else` branches of
 * the ISA-specific code) with every floating-point operation written out in a
 * fixed order, because the reference's own results depend on its absent
 * LightMath submodule and on the host ISA.  The pinned choices (each cited
 * where it is used):
 *   rcp(x)        = 1.0f/x (IEEE)                    SL_TriProcessor.cpp:281,440
 *   fmadd(a,b,c)  = fmaf(a,b,c) (fused)              SL_PostVertexTransform.hpp:516-523, SL_TriRasterizer.cpp:857
 *   dot/sum       = left to right, unfused           SL_TriProcessor.cpp:115, SL_FragmentProcessor.cpp:491
 *   mat4*vec4     = ((c0*x + c1*y) + c2*z) + c3*w     SL_TriProcessor.cpp:615-617
 *   mix(a,b,t)    = a + t*(b-a)                      SL_TriProcessor.cpp:410,441
 *   float->int    = truncate; NaN/out of range -> INT32_MIN (x86 cvttss2si)
 *   max(a,b)      = (a < b) ? b : a                  SL_PostVertexTransform.hpp:520-523
 *   length(v)     = sqrtf(dot(v, v)), inversesqrt(x) = 1/sqrtf(x)   SL_LineProcessor.cpp:132,150; SL_LineRasterizer.cpp:39,61
 *   (half)f       = IEEE binary16, round to nearest even   SL_Context.cpp:946-948, SL_FragmentProcessor.cpp:526-529
 * The build runs with -ffp-contract=off so that no other operation fuses.
 *
 * Deviations from the reference (undefined behaviour there):
 *   - a triangle whose snapped screen x or y is not finite is dropped;
 *   - a trailing partial triangle (elementEnd-elementBegin not a multiple of 3)
 *     or half line is skipped instead of reading past elementEnd;
 *   - vertex reads beyond the end of a vertex buffer return zeros;
 *   - point and line texels at or past W*H (written outside the attachment)
 *     and line texels beyond column W are dropped;
 *   - the thread-pool mode splits point and line rows by the addressed
 *     texel's row (the reference's split by y races on wrapped texels).
 * Every primitive kind keeps submission order per pixel (the reference's
 * radix sorts of a flush's bins by primIndex or depth are not restated).
 *
 * Each part is pinned by a closed-form, exact-arithmetic or invariant check
 * in tests/test_closed_forms.py and tests/test_recurrence.py (list at the top
 * of test_closed_forms.py), and for the wireframe, point and line modes in
 * tests/test_wireframe.py and tests/test_points_lines.py; the reference ships
 * no golden images.  The
 * adjoints and the smoothed-coverage mode (no reference counterpart) are
 * pinned in tests/test_backward.py and tests/test_smooth.py: finite
 * differences, the adjoint identity, closed forms of the smoothing model and
 * its exact sigma = 0 limit.
 *
 * Parity unpinned (no reference output and no mathematical fact fixes them;
 * the pins listed above decide them, and the GPU path shares the choice):
 *   - summation order and fusion inside LightMath's vec4*mat4, dot, sum, mix
 *     and sum_inv (the LightMath submodule is absent);
 *   - truncation (generic build) rather than rounding (AVX build) of span ends;
 *   - rcp being the exact IEEE reciprocal;
 *   - float->u8 conversion of values outside [0,255] (low byte of the
 *     truncated int);
 *   - LightMath's length and inversesqrt (pinned above; an approximate
 *     reciprocal square root would move clipped line ends and line depths);
 *   - LightMath's float -> half rounding (pinned above to the IEEE default;
 *     pinned against numpy's conversion in tests/test_formats.py).
 */
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <thread>
#include <vector>

#include "../include/sl_gpu.h"

namespace {

/*-----------------------------------------------------------------------------
 * Pinned scalar arithmetic
-----------------------------------------------------------------------------*/
struct V4 { float v[4]; };

inline float rcpf(float x) { return 1.0f / x; }

inline int32_t f2i(float f)
{
    if (!(f >= -2147483648.0f && f < 2147483648.0f)) return INT32_MIN;
    return (int32_t)f;
}

inline float maxf_ref(float a, float b) { return (a < b) ? b : a; }

inline V4 v4(float a, float b, float c, float d) { V4 r; r.v[0] = a; r.v[1] = b; r.v[2] = c; r.v[3] = d; return r; }

// ls::math::mat4 * vec4, column-major, SL_TriProcessor.cpp:615-617
inline V4 mat_mul(const float* m, const V4& p)
{
    V4 r;
    for (int i = 0; i < 4; ++i)
    {
        float a = m[0 * 4 + i] * p.v[0];
        float b = m[1 * 4 + i] * p.v[1];
        float c = m[2 * 4 + i] * p.v[2];
        float d = m[3 * 4 + i] * p.v[3];
        r.v[i] = ((a + b) + c) + d;
    }
    return r;
}

inline float dot4(const V4& a, const V4& b)
{
    float p0 = a.v[0] * b.v[0];
    float p1 = a.v[1] * b.v[1];
    float p2 = a.v[2] * b.v[2];
    float p3 = a.v[3] * b.v[3];
    return ((p0 + p1) + p2) + p3;
}

inline V4 mix4(const V4& a, const V4& b, float t)
{
    V4 r;
    for (int i = 0; i < 4; ++i) r.v[i] = a.v[i] + t * (b.v[i] - a.v[i]);
    return r;
}

/*-----------------------------------------------------------------------------
 * Resources
-----------------------------------------------------------------------------*/
struct Vbo { bool valid = false; std::vector<uint8_t> bytes; size_t grad = SL_INVALID_ID; };

struct Ibo { bool valid = false; int type = VERTEX_DATA_INT; uint32_t count = 0; uint32_t bpe = 4; std::vector<uint8_t> bytes; };

struct Binding { int dims = VERTEX_DIMENSION_4; int type = VERTEX_DATA_FLOAT; ptrdiff_t offset = 0; ptrdiff_t stride = 0; };

struct Vao { size_t vbo = SL_INVALID_ID; size_t ibo = SL_INVALID_ID; size_t nb = 0; Binding b[SL_VAO_MAX_BINDINGS]; };

struct Tex
{
    bool valid = false;
    int type = SL_COLOR_RGBA_8U;
    uint16_t w = 0, h = 0, d = 0;
    uint32_t bpt = 0;
    std::vector<uint8_t> data;   // padded + zeroed as SL_Texture.cpp:28-55
    size_t grad = SL_INVALID_ID;
};

// One (mesh, instance) pass drawn with a visibility attachment, kept for the
// backward pass.
struct SavedPass
{
    sl_mesh mesh;
    size_t instance;
    size_t numTris;
    uint32_t seqBase;
    size_t shader;
    size_t tex;               // texture bound through the uniforms at draw time
    size_t drawIdx;           // the draw call it belongs to (instances of one call share it)
    std::vector<uint8_t> uni; // uniform snapshot, scissor matrix and viewport at draw time
    float scissorMat[16];
    float vpDims[4];
};

struct Fbo
{
    std::vector<size_t> colors;
    size_t depth = SL_INVALID_ID;
    size_t dbgLast = SL_INVALID_ID;
    size_t dbgCount = SL_INVALID_ID;
    size_t vis = SL_INVALID_ID;
    std::vector<SavedPass> saved;   // passes since the depth attachment was last cleared
    size_t savedDraws = 0;          // draw calls among them
    bool savedStale = false;
    bool savedBlend = false;
    size_t posGrad = SL_INVALID_ID, posOffset = SL_INVALID_ID;   // smoothed coverage buffers
    uint32_t posVerts = 0;
};

struct Ubo { uint8_t bytes[SL_MAX_UNIFORM_BUFFER_SIZE]; };

struct Shader
{
    uint8_t cull, dtest, dmask, blend, nvary, nout;
    uint32_t vsProg, fsProg;
    size_t ubo;
};

} // namespace

struct sl_context
{
    std::vector<Vbo> vbos;
    std::vector<Ibo> ibos;
    std::vector<Vao> vaos;
    std::vector<Tex> texs;
    std::vector<Fbo> fbos;
    std::vector<Ubo> ubos;
    std::vector<Shader> shaders;
    int32_t viewport[4] = {0, 0, 65535, 65535};   // SL_ViewportState.hpp:80-81
    int32_t scissor[4]  = {0, 0, 65535, 65535};
    unsigned numThreads = 1;
    sl_frame_stats stats{};
    uint32_t triSeq = 0;                          // submitted-triangle counter (debug buffers)
    std::string err;
};

namespace {

uint32_t bytes_per_color(int type)
{
    switch (type)
    {
        case SL_COLOR_R_8U: return 1;
        case SL_COLOR_RG_8U: return 2;
        case SL_COLOR_RGB_8U: return 3;
        case SL_COLOR_RGBA_8U: return 4;
        case SL_COLOR_R_32U: return 4;
        case SL_COLOR_R_HALF: return 2;
        case SL_COLOR_R_FLOAT: return 4;
        case SL_COLOR_RG_FLOAT: return 8;
        case SL_COLOR_RGB_FLOAT: return 12;
        case SL_COLOR_RGBA_FLOAT: return 16;
        case SL_COLOR_RGB_332: return 1;
        case SL_COLOR_RGB_565: case SL_COLOR_RGBA_5551: case SL_COLOR_RGBA_4444: return 2;
        case SL_COLOR_RGBA_1010102: return 4;
        default: return 0;
    }
}

// Colour attachments drawn to: 8-bit unsigned (R, RG, RGB, RGBA) and float
// (R, RG, RGB, RGBA).  Components per texel, or 0.
int u8_components(int type)
{
    return type == SL_COLOR_R_8U ? 1 : type == SL_COLOR_RG_8U ? 2 : type == SL_COLOR_RGB_8U ? 3 : type == SL_COLOR_RGBA_8U ? 4 : 0;
}

int float_components(int type)
{
    return type == SL_COLOR_R_FLOAT ? 1 : type == SL_COLOR_RG_FLOAT ? 2 : type == SL_COLOR_RGB_FLOAT ? 3 : type == SL_COLOR_RGBA_FLOAT ? 4 : 0;
}

// Packed formats (SL_ColorCompressed.hpp): bit-field widths of r, g, b, a
// (r in the lowest bits, as the structs' bit-fields are laid out) and texel
// bytes.  Zero widths for a format that is not packed.
struct PackedFmt { int bits[4]; int bytes; };

PackedFmt packed_format(int type)
{
    switch (type)
    {
        case SL_COLOR_RGB_332: return PackedFmt{{3, 3, 2, 0}, 1};            // :20-28
        case SL_COLOR_RGB_565: return PackedFmt{{5, 6, 5, 0}, 2};            // :198-206
        case SL_COLOR_RGBA_5551: return PackedFmt{{5, 5, 5, 1}, 2};          // :376-385
        case SL_COLOR_RGBA_4444: return PackedFmt{{4, 4, 4, 4}, 2};          // :558-567
        case SL_COLOR_RGBA_1010102: return PackedFmt{{10, 10, 10, 2}, 4};    // :735-744
        default: return PackedFmt{{0, 0, 0, 0}, 0};
    }
}

// rgba_cast<packed, float> (e.g. SL_ColorCompressed.hpp:300-310): each
// component times its maximum, converted to the field's integer type (the
// low bits of the truncated int, as for 8-bit colour) and stored in the
// bit-field (which keeps its low bits).
uint32_t packed_encode(const PackedFmt& f, const float c[4])
{
    uint32_t w = 0;
    int shift = 0;
    for (int k = 0; k < 4; ++k)
    {
        if (f.bits[k] == 0) continue;
        const uint32_t mx = (1u << f.bits[k]) - 1u;
        w |= ((uint32_t)f2i(c[k] * (float)mx) & mx) << shift;
        shift += f.bits[k];
    }
    return w;
}

// rgba_cast<float, packed> (e.g. :354-363): field / maximum, alpha 1 when the
// format has none.
void packed_decode(const PackedFmt& f, uint32_t w, float c[4])
{
    int shift = 0;
    for (int k = 0; k < 4; ++k)
    {
        if (f.bits[k] == 0) { c[k] = (k == 3) ? 1.f : 0.f; continue; }
        const uint32_t mx = (1u << f.bits[k]) - 1u;
        c[k] = (float)((w >> shift) & mx) / (float)mx;
        shift += f.bits[k];
    }
}

// IEEE binary16 <-> binary32 (ls::math::half, LightMath absent): the
// conversion to half rounds to nearest, ties to even; overflow gives
// infinity, NaN stays NaN.  Written from the two formats' definitions.
uint16_t float_to_half(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    const uint32_t sign = (u >> 16) & 0x8000u;
    const uint32_t ex = (u >> 23) & 0xFFu;
    uint32_t man = u & 0x7FFFFFu;
    if (ex == 0xFFu) return (uint16_t)(sign | 0x7C00u | (man ? 0x200u | (man >> 13) : 0u));
    const int e = (int)ex - 127 + 15;
    if (e >= 31) return (uint16_t)(sign | 0x7C00u);
    if (e <= 0)
    {
        // subnormal half (or zero): value = man' * 2^-24
        if (e < -10) return (uint16_t)sign;
        man |= 0x800000u;
        const int shift = 14 - e;          // 24-bit significand down to the subnormal grid
        const uint32_t q = man >> shift, rem = man & ((1u << shift) - 1u), half = 1u << (shift - 1);
        const uint32_t r = q + ((rem > half || (rem == half && (q & 1u))) ? 1u : 0u);
        return (uint16_t)(sign | r);
    }
    const uint32_t q = man >> 13, rem = man & 0x1FFFu;
    uint32_t r = ((uint32_t)e << 10) | q;
    if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) ++r;   // may carry into the exponent (to infinity)
    return (uint16_t)(sign | r);
}

float half_to_float(uint16_t h)
{
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t ex = (h >> 10) & 0x1Fu, man = h & 0x3FFu;
    uint32_t u;
    if (ex == 0)
    {
        if (man == 0) u = sign;
        else
        {
            int e = -1;
            uint32_t m = man;
            do { m <<= 1; ++e; } while (!(m & 0x400u));
            u = sign | ((uint32_t)(127 - 15 - e) << 23) | ((m & 0x3FFu) << 13);
        }
    }
    else if (ex == 0x1Fu) u = sign | 0x7F800000u | (man << 13);
    else u = sign | ((ex - 15 + 127) << 23) | (man << 13);
    float f;
    std::memcpy(&f, &u, 4);
    return f;
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

/*-----------------------------------------------------------------------------
 * Per-draw state
-----------------------------------------------------------------------------*/
struct TVert { V4 pos; V4 var[SL_SHADER_MAX_VARYING_VECTORS]; };

struct Bin
{
    V4 p[3];                   // screen x, y (snapped), ndc z, 1/w
    V4 bx, by, bz;             // barycentric planes ddx*denom, ddy*denom, ddz*denom
    V4 var[3][SL_SHADER_MAX_VARYING_VECTORS];
    uint32_t seq;
};

struct Target
{
    int W, H;
    float* depth;             // R32F depth attachment, or null
    uint16_t* depthH;         // R16F depth attachment, or null
    uint8_t* color8;          // 8-bit attachment 0 (comps components), or null
    float* colorF;            // float attachment 0 (comps components), or null
    int comps;                // components per colour texel (bytes per texel for a packed format)
    int packed;               // attachment 0's packed format (SL_COLOR_RGB_332 ...), or 0
    // attachments 1..3 of a multiple-output fragment program, same layout
    uint8_t* extra8[3];
    float* extraF[3];
    int extraComps[3];
    int extraPacked[3];
    uint32_t* dbgLast;
    uint32_t* dbgCount;
    float* vis;               // visibility attachment {seq bits, b0, b1, b2}, or null
};

struct Draw
{
    sl_context* ctx;
    const Shader* sh;
    const uint8_t* uni;       // uniform snapshot (may be null)
    Target tgt;
    float scissorMat[16];
    float vpDims[4];
    const Vao* vao;
    const Vbo* vbo;
    const Ibo* ibo;
    unsigned nvary;           // the shader's varyings, plus the weight slot with a visibility attachment
    int visSlot;              // varying slot of the source-vertex weights, or -1
    int prim = 0;             // primitive kind of the draw call (kPrimFill, kPrimWire, ...)
    uint64_t tested = 0, shaded = 0, binned = 0;
};

const sl_std_uniforms* uniforms(const Draw& d)
{
    return reinterpret_cast<const sl_std_uniforms*>(d.uni);
}

// Primitive kinds.  A draw call's processor and rasteriser follow the render
// mode of its first mesh (sl_processor_type_for_draw_mode,
// SL_ShaderProcessor.cpp:12-37; rasterizer.mMode = mRenderMode,
// SL_VertexProcessor.cpp:154; ProcessorPool passes meshes[0].mode,
// SL_ProcessorPool.cpp:286,339); whether a mesh reads its index buffer is the
// processor's own test of that mesh's mode (SL_TriProcessor.cpp:682,697).
enum { kPrimFill = 0, kPrimWire = 1, kPrimLine = 2, kPrimPoint = 3, kPrimBad = -1 };

int prim_of(int mode)
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
bool mesh_indexed(int prim, int mode)
{
    if (prim == kPrimFill || prim == kPrimWire) return mode == RENDER_MODE_INDEXED_TRIANGLES || mode == RENDER_MODE_INDEXED_TRI_WIRE;
    if (prim == kPrimLine) return mode == RENDER_MODE_INDEXED_LINES;
    return mode == RENDER_MODE_INDEXED_POINTS;
}

size_t verts_per_prim(int prim)
{
    return prim == kPrimPoint ? 1u : prim == kPrimLine ? 2u : 3u;
}

// LS_EPSILON (LightSetup, absent).  The line clipper compares differences of
// snapped screen coordinates (whole numbers) against it, so any value in
// (0, 1] gives the same results.
constexpr float kLineEpsilon = 1.0e-5f;

/*-----------------------------------------------------------------------------
 * Textures and samplers (SL_Sampler.hpp:51-218, SL_Texture.hpp:190)
-----------------------------------------------------------------------------*/
// SL_WrapModeRepeat (SL_Sampler.hpp:111-114): (u < 0 ? 1 : 0) + fmod_1(u),
// fmod_1 taken as the sign-preserving fractional part u - trunc(u).
inline float wrap_repeat(float u)
{
    return (u < 0.f ? 1.f : 0.f) + (u - truncf(u));
}

inline uint16_t to_u16(float f) { return (uint16_t)(uint32_t)f2i(f); }

// sl_sample_bilinear<SL_ColorRGBA8, REPEAT> (SL_Sampler.hpp:186-218): the
// x+1/y+1 neighbours clamp to W/H, so x == W aliases the next row's first
// texel and y == H reads the zeroed allocation padding.
void sample_bilinear_rgba8(const Tex& t, float u, float v, uint8_t out[4])
{
    const float xf = wrap_repeat(u) * (float)t.w;
    const float yf = wrap_repeat(v) * (float)t.h;
    const uint16_t xi0 = to_u16(xf);
    const uint16_t yi0 = to_u16(yf);
    const uint16_t xi1 = (uint16_t)std::min<uint32_t>((uint32_t)xi0 + 1u, t.w);
    const uint16_t yi1 = (uint16_t)std::min<uint32_t>((uint32_t)yi0 + 1u, t.h);
    const float dx = xf - (float)xi0;
    const float dy = yf - (float)yi0;
    const float omdx = 1.f - dx;
    const float omdy = 1.f - dy;
    const uint8_t* px = t.data.data();
    const size_t i0 = (size_t)xi0 + (size_t)t.w * yi0;
    const size_t i1 = (size_t)xi0 + (size_t)t.w * yi1;
    const size_t i2 = (size_t)xi1 + (size_t)t.w * yi0;
    const size_t i3 = (size_t)xi1 + (size_t)t.w * yi1;
    for (int c = 0; c < 4; ++c)
    {
        const float p0 = (float)px[i0 * 4 + c] * 0.00392156862745f;
        const float p1 = (float)px[i1 * 4 + c] * 0.00392156862745f;
        const float p2 = (float)px[i2 * 4 + c] * 0.00392156862745f;
        const float p3 = (float)px[i3 * 4 + c] * 0.00392156862745f;
        const float w0 = (p0 * omdx) * omdy;
        const float w1 = (p1 * omdx) * dy;
        const float w2 = (p2 * dx) * omdy;
        const float w3 = (p3 * dx) * dy;
        const float s = ((w0 + w1) + w2) + w3;
        out[c] = (uint8_t)(uint32_t)f2i(s * 255.f);   // color_cast<uint8_t,float>, SL_Color.hpp:621-626
    }
}

// sl_sample_nearest<color, REPEAT> (SL_Sampler.hpp:122-148): the texel at
// xi = (uint32_t)(W * wrap(u)), yi = (uint32_t)(H * wrap(v)), ordered layout
// xi + W * yi.  wrap(u) can be 1 (a tiny negative u rounds up), so xi = W
// reads texel 0 of the next row and yi = H the zeroed padding.
size_t nearest_index(const Tex& t, float u, float v)
{
    const uint32_t xi = (uint32_t)f2i((float)t.w * wrap_repeat(u));
    const uint32_t yi = (uint32_t)f2i((float)t.h * wrap_repeat(v));
    return (size_t)xi + (size_t)t.w * yi;
}

void sample_nearest_rgba8(const Tex& t, float u, float v, uint8_t out[4])
{
    std::memcpy(out, t.data.data() + 4 * nearest_index(t, u, v), 4);
}

void sample_nearest_rgbaf(const Tex& t, float u, float v, float out[4])
{
    std::memcpy(out, t.data.data() + 16 * nearest_index(t, u, v), 16);
}

// Same sampler for an RGBA32F texture: color_cast<float,float> is the identity.
void sample_bilinear_rgbaf(const Tex& t, float u, float v, float out[4])
{
    const float xf = wrap_repeat(u) * (float)t.w;
    const float yf = wrap_repeat(v) * (float)t.h;
    const uint16_t xi0 = to_u16(xf);
    const uint16_t yi0 = to_u16(yf);
    const uint16_t xi1 = (uint16_t)std::min<uint32_t>((uint32_t)xi0 + 1u, t.w);
    const uint16_t yi1 = (uint16_t)std::min<uint32_t>((uint32_t)yi0 + 1u, t.h);
    const float dx = xf - (float)xi0;
    const float dy = yf - (float)yi0;
    const float omdx = 1.f - dx;
    const float omdy = 1.f - dy;
    const float* px = reinterpret_cast<const float*>(t.data.data());
    const size_t i0 = (size_t)xi0 + (size_t)t.w * yi0;
    const size_t i1 = (size_t)xi0 + (size_t)t.w * yi1;
    const size_t i2 = (size_t)xi1 + (size_t)t.w * yi0;
    const size_t i3 = (size_t)xi1 + (size_t)t.w * yi1;
    for (int c = 0; c < 4; ++c)
    {
        const float w0 = (px[i0 * 4 + c] * omdx) * omdy;
        const float w1 = (px[i1 * 4 + c] * omdx) * dy;
        const float w2 = (px[i2 * 4 + c] * dx) * omdy;
        const float w3 = (px[i3 * 4 + c] * dx) * dy;
        out[c] = ((w0 + w1) + w2) + w3;
    }
}

/*-----------------------------------------------------------------------------
 * Built-in shader programs (the device programs of include/sl_gpu.h)
-----------------------------------------------------------------------------*/
// SL_VertexArray::offset(bind, vert) = offset + stride*vert (SL_VertexArray.hpp:120)
void read_attr(const Draw& d, int bind, size_t vertId, float out[4], float defW)
{
    out[0] = out[1] = out[2] = 0.f;
    out[3] = defW;
    if ((size_t)bind >= d.vao->nb || d.vbo == nullptr) return;
    const Binding& b = d.vao->b[bind];
    const ptrdiff_t off = b.offset + b.stride * (ptrdiff_t)vertId;
    const int n = b.dims + 1;
    const uint8_t* base = d.vbo->bytes.data();
    const size_t size = d.vbo->bytes.size();
    if (b.type == VERTEX_DATA_FLOAT)
    {
        if (off < 0 || (size_t)off + 4u * n > size) return;
        for (int i = 0; i < n; ++i) std::memcpy(&out[i], base + off + 4 * i, 4);
    }
    else if (b.type == VERTEX_DATA_BYTE)
    {
        if (off < 0 || (size_t)off + n > size) return;
        for (int i = 0; i < n; ++i) out[i] = (float)base[off + i] * 0.00392156862745f;  // SL_Color.hpp:666-669
    }
}

void model_matrix(const Draw& d, size_t instanceId, float m[16])
{
    const sl_std_uniforms* u = uniforms(d);
    if (u->instance_vbo != SL_NO_RESOURCE && u->instance_vbo < d.ctx->vbos.size())
    {
        const Vbo& ib = d.ctx->vbos[u->instance_vbo];
        const size_t off = instanceId * 64u;
        if (off + 64u <= ib.bytes.size())
        {
            std::memcpy(m, ib.bytes.data() + off, 64);
            return;
        }
        std::memset(m, 0, 64);
        return;
    }
    std::memcpy(m, u->model, 64);
}

V4 run_vertex_program(const Draw& d, size_t vertId, size_t instanceId, V4* var)
{
    switch (d.sh->vsProg)
    {
        case SL_PROGRAM_PASSTHROUGH_GOURAUD:
        {
            float p[4], c[4];
            read_attr(d, 0, vertId, p, 1.f);
            read_attr(d, 1, vertId, c, 1.f);
            var[0] = v4(c[0], c[1], c[2], c[3]);
            return v4(p[0], p[1], p[2], p[3]);
        }
        case SL_PROGRAM_LAMBERT_TEXTURED:
        case SL_PROGRAM_LAMBERT_MRT:
        case SL_PROGRAM_COLOR_TEXTURED:
        {
            float p[4], a[4], uv[4], m[16];
            read_attr(d, 0, vertId, p, 1.f);
            read_attr(d, 1, vertId, a, 0.f);
            read_attr(d, 2, vertId, uv, 0.f);
            model_matrix(d, instanceId, m);
            const V4 world = mat_mul(m, v4(p[0], p[1], p[2], 1.f));
            if (d.sh->vsProg != SL_PROGRAM_COLOR_TEXTURED)
            {
                const V4 n = mat_mul(m, v4(a[0], a[1], a[2], 0.f));
                var[0] = v4(uv[0], uv[1], 0.f, 0.f);
                var[1] = v4(n.v[0], n.v[1], n.v[2], 0.f);
                if (d.sh->vsProg == SL_PROGRAM_LAMBERT_MRT) var[2] = v4(world.v[0], world.v[1], world.v[2], 1.f);
            }
            else
            {
                var[0] = v4(a[0], a[1], a[2], 1.f);
                var[1] = v4(uv[0], uv[1], 0.f, 0.f);
            }
            return mat_mul(uniforms(d)->mvp, world);
        }
        default:
            return v4(0.f, 0.f, 0.f, 0.f);
    }
}

// math::clamp(x, lo, hi) as min(max(x, lo), hi) with max(a,b) = (a < b) ? b : a
// and min(a,b) = (b < a) ? b : a (NaN stays NaN).
inline float clampf_ref(float x, float lo, float hi)
{
    const float m = (x < lo) ? lo : x;
    return (hi < m) ? hi : m;
}

inline V4 clamp01(const V4& v)
{
    return v4(clampf_ref(v.v[0], 0.f, 1.f), clampf_ref(v.v[1], 0.f, 1.f), clampf_ref(v.v[2], 0.f, 1.f), clampf_ref(v.v[3], 0.f, 1.f));
}

bool run_fragment_program(const Draw& d, const V4* var, V4* outs)
{
    V4& out = outs[0];
    switch (d.sh->fsProg)
    {
        case SL_PROGRAM_PASSTHROUGH_GOURAUD:
            out = var[0];
            return true;

        case SL_PROGRAM_LAMBERT_TEXTURED:
        case SL_PROGRAM_LAMBERT_MRT:
        {
            const sl_std_uniforms* u = uniforms(d);
            uint8_t texel[4] = {0, 0, 0, 0};
            if (u->texture < d.ctx->texs.size() && d.ctx->texs[u->texture].type == SL_COLOR_RGBA_8U)
            {
                if (u->flags & SL_UNIFORM_SAMPLE_NEAREST)
                    sample_nearest_rgba8(d.ctx->texs[u->texture], var[0].v[0], var[0].v[1], texel);
                else
                    sample_bilinear_rgba8(d.ctx->texs[u->texture], var[0].v[0], var[0].v[1], texel);
            }
            float albedo[4];
            for (int c = 0; c < 4; ++c) albedo[c] = (float)texel[c] * 0.00392156862745f;
            const float nx = var[1].v[0], ny = var[1].v[1], nz = var[1].v[2];
            const float len2 = ((nx * nx) + (ny * ny)) + (nz * nz);
            const float inv = 1.0f / sqrtf(len2);
            const float mx = nx * inv, my = ny * inv, mz = nz * inv;
            const float ndl = -(((mx * u->light_dir[0]) + (my * u->light_dir[1])) + (mz * u->light_dir[2]));
            const float lam = (ndl > 0.f) ? ndl : 0.f;
            for (int c = 0; c < 3; ++c)
            {
                const float light = u->ambient[c] + (u->diffuse[c] * lam);
                const float v = (albedo[c] * light) * u->tint[c];
                out.v[c] = (v < 1.f) ? v : 1.f;
            }
            out.v[3] = albedo[3] * u->tint[3];
            if (d.sh->fsProg == SL_PROGRAM_LAMBERT_MRT)
            {
                // the G-buffer outputs of tests/sl_mrt_test.cpp:105-123: position, uv, normal
                outs[1] = clamp01(var[2]);
                outs[2] = clamp01(var[0]);
                outs[3] = clamp01(v4(mx, my, mz, 0.f));
            }
            return true;
        }

        case SL_PROGRAM_COLOR_TEXTURED:
        {
            const sl_std_uniforms* u = uniforms(d);
            float t[4] = {1.f, 1.f, 1.f, 1.f};
            if (u->texture < d.ctx->texs.size() && d.ctx->texs[u->texture].type == SL_COLOR_RGBA_FLOAT)
            {
                if (u->flags & SL_UNIFORM_SAMPLE_NEAREST)
                    sample_nearest_rgbaf(d.ctx->texs[u->texture], var[1].v[0], var[1].v[1], t);
                else
                    sample_bilinear_rgbaf(d.ctx->texs[u->texture], var[1].v[0], var[1].v[1], t);
            }
            for (int c = 0; c < 4; ++c) out.v[c] = var[0].v[c] * t[c];
            return true;
        }

        default:
            out = v4(0.f, 0.f, 0.f, 0.f);
            return true;
    }
}

/*-----------------------------------------------------------------------------
 * Output merger (SL_Framebuffer.cpp:44-80,153-166,270-368; SL_Color.hpp:600-670)
-----------------------------------------------------------------------------*/
inline float clamp01(float x) { return (x < 0.f) ? 0.f : ((x > 1.f) ? 1.f : x); }

void blend_rgba(int mode, const V4& s, float d[4])
{
    const float as = s.v[3];
    const float mod = 1.f - as;
    if (mode == SL_BLEND_ALPHA)                       // SL_Framebuffer.cpp:324-330
    {
        const float dstMod = mod * d[3];
        const float dstAlpha = dstMod + as;
        const float r = rcpf(dstAlpha);
        float o[4];
        for (int c = 0; c < 3; ++c) o[c] = fmaf(s.v[c], as, d[c] * dstMod) * r;
        o[3] = dstAlpha;
        for (int c = 0; c < 4; ++c) d[c] = o[c];
    }
    else if (mode == SL_BLEND_PREMULTIPLED_ALPHA)     // :331-334
    {
        for (int c = 0; c < 4; ++c) d[c] = fmaf(d[c], mod, s.v[c]);
    }
    else if (mode == SL_BLEND_ADDITIVE)               // :335-338
    {
        for (int c = 0; c < 4; ++c) d[c] = fmaf(s.v[c], as, d[c]);
    }
    else if (mode == SL_BLEND_SCREEN)                 // :339-342
    {
        for (int c = 0; c < 4; ++c) d[c] = (s.v[c] * as) + (d[c] * mod);
    }
    for (int c = 0; c < 4; ++c) d[c] = clamp01(d[c]);
}

// The depth texel as the depth test reads it ((float) of a half texel,
// _sl_get_depth_texel, SL_TriRasterizer.cpp:94-101), and its write
// ((depth_type)z, SL_FragmentProcessor.cpp:526-529).
float depth_read(const Target& t, size_t idx)
{
    return t.depthH ? half_to_float(t.depthH[idx]) : t.depth[idx];
}

void depth_write(Target& t, size_t idx, float z)
{
    if (t.depthH) t.depthH[idx] = float_to_half(z);
    else t.depth[idx] = z;
}

// assign_pixel / assign_alpha_pixel (SL_Framebuffer.cpp:42-80, 268-370) for
// 8-bit and float attachments of 1..4 components.  A blend reads the
// destination as (r, 0, 0, 1), (r, g, 0, 1), (r, r, r, 1) for one, two and
// three components (the three-component read repeats the first component,
// :306-307), blends all four, clamps and stores the texel's components.
void place_color_in(const Draw& d, uint8_t* color8, float* colorF, int n, int packed, int x, int y, const V4& c);

void place_color(const Draw& d, int x, int y, const V4& c)
{
    place_color_in(d, d.tgt.color8, d.tgt.colorF, d.tgt.comps, d.tgt.packed, x, y, c);
}

void place_color_in(const Draw& d, uint8_t* color8, float* colorF, int n, int packed, int x, int y, const V4& c)
{
    const Target& t = d.tgt;
    const size_t idx = (size_t)x + (size_t)t.W * (size_t)y;
    const int blend = d.sh->blend;
    if (packed)
    {
        // assign_pixel / assign_alpha_pixel for packed texels (SL_Framebuffer.cpp:82-141, 370-560):
        // the whole texel decoded, blended, clamped and encoded
        const PackedFmt f = packed_format(packed);
        uint8_t* px = color8 + idx * (size_t)f.bytes;
        uint32_t w = 0;
        std::memcpy(&w, px, (size_t)f.bytes);
        if (blend == SL_BLEND_OFF)
        {
            w = packed_encode(f, c.v);
        }
        else
        {
            float dst[4];
            packed_decode(f, w, dst);
            blend_rgba(blend, c, dst);
            w = packed_encode(f, dst);
        }
        std::memcpy(px, &w, (size_t)f.bytes);
        return;
    }
    auto read_dst = [&](const float v[4], float dst[4]) {
        if (n == 1) { dst[0] = v[0]; dst[1] = 0.f; dst[2] = 0.f; dst[3] = 1.f; }
        else if (n == 2) { dst[0] = v[0]; dst[1] = v[1]; dst[2] = 0.f; dst[3] = 1.f; }
        else if (n == 3) { dst[0] = v[0]; dst[1] = v[0]; dst[2] = v[0]; dst[3] = 1.f; }
        else for (int k = 0; k < 4; ++k) dst[k] = v[k];
    };
    if (color8)
    {
        uint8_t* px = color8 + idx * (size_t)n;
        if (blend == SL_BLEND_OFF)
        {
            for (int k = 0; k < n; ++k) px[k] = (uint8_t)(uint32_t)f2i(c.v[k] * 255.f);
        }
        else
        {
            float v[4] = {0.f, 0.f, 0.f, 0.f}, dst[4];
            for (int k = 0; k < n; ++k) v[k] = (float)px[k] * 0.00392156862745f;
            // the missing components are the type's min (0) and max (255) cast to float
            const float one = 255.f * 0.00392156862745f;
            read_dst(v, dst);
            if (n < 4) dst[3] = one;
            blend_rgba(blend, c, dst);
            for (int k = 0; k < n; ++k) px[k] = (uint8_t)(uint32_t)f2i(dst[k] * 255.f);
        }
    }
    else if (colorF)
    {
        float* px = colorF + idx * (size_t)n;
        if (blend == SL_BLEND_OFF)
        {
            for (int k = 0; k < n; ++k) px[k] = c.v[k];
        }
        else
        {
            float v[4] = {0.f, 0.f, 0.f, 0.f}, dst[4];
            for (int k = 0; k < n; ++k) v[k] = px[k];
            read_dst(v, dst);
            blend_rgba(blend, c, dst);
            for (int k = 0; k < n; ++k) px[k] = dst[k];
        }
    }
}

/*-----------------------------------------------------------------------------
 * Raster + fragment stage (SL_ScanlineBounds.hpp:137-166,310-322;
 * SL_TriRasterizer.cpp:806-923; SL_FragmentProcessor.cpp:315-330,416-531)
-----------------------------------------------------------------------------*/
inline bool depth_func(int f, float a, float b)      // SL_ShaderUtil.hpp:240-507, scalar operators
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

struct Scanline { float v0y, v1y, v0x, v1x, p20y, p20x, p21xy, p10xy; };

// SL_ScanlineBounds::init, generic branch (:143-166)
Scanline scanline_init(V4 p0, V4 p1, V4 p2)
{
    if (p0.v[1] < p1.v[1]) std::swap(p0, p1);
    if (p0.v[1] < p2.v[1]) std::swap(p0, p2);
    if (p1.v[1] < p2.v[1]) std::swap(p1, p2);
    Scanline s;
    s.v0y = p0.v[1];
    s.v1y = p1.v[1];
    s.v0x = p0.v[0];
    s.v1x = p1.v[0];
    s.p20y = rcpf(p2.v[1] - p0.v[1]);
    s.p20x = p2.v[0] - p0.v[0];
    s.p21xy = (p2.v[0] - p1.v[0]) * rcpf(p2.v[1] - p1.v[1]);
    s.p10xy = (p1.v[0] - p0.v[0]) * rcpf(p1.v[1] - p0.v[1]);
    return s;
}

// SL_ScanlineBounds::step, generic branch (:310-322).  Only the selected edge
// is converted to int: the other one may be +-inf or NaN.
inline void scanline_step(const Scanline& s, float yf, int32_t& xMin, int32_t& xMax)
{
    const float d0 = yf - s.v0y;
    const float d1 = yf - s.v1y;
    const float alpha = d0 * s.p20y;
    const int32_t lo = f2i((s.p20x * alpha) + s.v0x);
    const int32_t hi = (d1 < 0.f) ? f2i((s.p21xy * d1) + s.v1x) : f2i((s.p10xy * d0) + s.v0x);
    xMin = (lo >= hi) ? hi : lo;
    xMax = (lo >= hi) ? lo : hi;
}

// A fragment that passed the depth test, with its varyings: the fragment
// program, the colour output and the depth write (SL_FragmentProcessor.cpp:
// 365-402, 496-530; SL_PointRasterizer.cpp:64-96).
void finish_fragment(Draw& d, int x, int y, float z, const V4* var, uint32_t seq)
{
    V4 outs[SL_SHADER_MAX_FRAG_OUTPUTS];
    const bool haveOutputs = run_fragment_program(d, var, outs);
    if (haveOutputs && d.sh->nout > 0)
    {
        // one output per attachment (SL_FragmentProcessor.cpp:379-395)
        for (unsigned k = d.sh->nout; k-- > 1;)   // attachments 3, 2, 1, then 0 (the switch falls through)
            place_color_in(d, d.tgt.extra8[k - 1], d.tgt.extraF[k - 1], d.tgt.extraComps[k - 1], d.tgt.extraPacked[k - 1], x, y,
                           outs[k]);
        place_color(d, x, y, outs[0]);
    }

    const size_t idx = (size_t)x + (size_t)d.tgt.W * (size_t)y;
    if (d.sh->dmask == SL_DEPTH_MASK_ON) depth_write(d.tgt, idx, z);   // regardless of the shader's return
    if (d.tgt.dbgLast) d.tgt.dbgLast[idx] = seq;
    if (d.tgt.dbgCount) d.tgt.dbgCount[idx] += 1u;
    if (d.tgt.vis)
    {
        // the source triangle's vertex weights, interpolated like a varying
        float* v = d.tgt.vis + idx * 4;
        std::memcpy(&v[0], &seq, 4);
        for (int k = 0; k < 3; ++k) v[1 + k] = var[d.visSlot].v[k];
    }
    ++d.shaded;
}

void shade_fragment(Draw& d, const Bin& b, int x, int y, float z, const float bc[4])
{
    // perspective correction, generic branch SL_FragmentProcessor.cpp:486-494
    const float h0 = b.p[0].v[3], h1 = b.p[1].v[3], h2 = b.p[2].v[3];
    const float t0 = bc[0] * h0, t1 = bc[1] * h1, t2 = bc[2] * h2, t3 = bc[3] * 0.f;
    const float s = rcpf(((t0 + t1) + t2) + t3);
    const float c0 = t0 * s, c1 = t1 * s, c2 = t2 * s;

    // interpolate_tri_varyings, generic branch :315-330
    V4 var[SL_SHADER_MAX_VARYING_VECTORS];
    for (unsigned i = 0; i < d.nvary; ++i)
        for (int k = 0; k < 4; ++k)
            var[i].v[k] = ((b.var[0][i].v[k] * c0) + (b.var[1][i].v[k] * c1)) + (b.var[2][i].v[k] * c2);
    for (unsigned i = d.nvary; i < SL_SHADER_MAX_VARYING_VECTORS; ++i) var[i] = v4(0.f, 0.f, 0.f, 0.f);

    finish_fragment(d, x, y, z, var, b.seq);
}

// d0 / d1 of render_wireframe: max(|a - b|, 1) on int32.  The row above a
// triangle's first row can have a span end of INT32_MIN (an extrapolated
// horizontal edge); the reference's int subtraction then wraps, as it does on
// x86, and |INT32_MIN| stays INT32_MIN, so the result is 1.
int32_t wire_gap(int32_t a, int32_t b)
{
    const int32_t dd = (int32_t)((uint32_t)a - (uint32_t)b);
    const int32_t ad = (dd < 0) ? (int32_t)(0u - (uint32_t)dd) : dd;
    return (ad < 1) ? 1 : ad;
}

// SL_TriRasterizer::render_wireframe (SL_TriRasterizer.cpp:116-202): on each
// row, the pixels from the span's left end over the distance the left end
// moved since the row above (at least one), then a jump to the pixels over
// the distance the right end moved.  The jump target can lie inside the first
// run, so a pixel may be visited twice.  Barycentrics are evaluated directly
// per pixel (two fused multiply-adds), not stepped.
void raster_wire_bin(Draw& d, const Bin& b, int32_t rowThread, int32_t numRowThreads)
{
    const V4* P = b.p;
    const int32_t bboxMinY = f2i(std::min(std::min(P[0].v[1], P[1].v[1]), P[2].v[1]));
    const int32_t bboxMaxY = f2i(std::max(std::max(P[0].v[1], P[1].v[1]), P[2].v[1]));
    const int32_t n1 = numRowThreads - 1;
    const int32_t offset = n1 - ((bboxMinY + (n1 - rowThread)) % numRowThreads);
    const Scanline sc = scanline_init(P[0], P[1], P[2]);
    const float dep[3] = {P[0].v[2], P[1].v[2], P[2].v[2]};
    const int dfunc = d.sh->dtest;
    const int W = d.tgt.W;

    for (int32_t y = bboxMinY + offset; y < bboxMaxY; y += numRowThreads)
    {
        const float yf = (float)y;
        float bcY[4];
        for (int k = 0; k < 3; ++k) bcY[k] = fmaf(b.by.v[k], yf, b.bz.v[k]);
        bcY[3] = fmaf(b.by.v[3], 0.f, b.bz.v[3]);
        int32_t xMin0, xMax0, xMin1, xMax1;
        scanline_step(sc, yf, xMin0, xMax0);
        scanline_step(sc, yf - 1.f, xMin1, xMax1);
        const int32_t d0 = wire_gap(xMin0, xMin1);
        const int32_t d1 = wire_gap(xMax0, xMax1);
        for (int32_t ix = 0, x = xMin0; (uint32_t)x < (uint32_t)xMax0; ++ix, ++x)
        {
            if (ix == d0)
            {
                const int32_t t = (int32_t)((uint32_t)xMax0 - (uint32_t)d1 - 1u);
                x = (xMin0 < t) ? t : xMin0;
                continue;
            }
            const float xf = (float)x;
            float bc[4];
            for (int k = 0; k < 3; ++k) bc[k] = fmaf(b.bx.v[k], xf, bcY[k]);
            bc[3] = fmaf(b.bx.v[3], 0.f, bcY[3]);
            const float z = (((dep[0] * bc[0]) + (dep[1] * bc[1])) + (dep[2] * bc[2])) + (0.f * bc[3]);
            const float dv = depth_read(d.tgt, (size_t)x + (size_t)W * (size_t)y);
            ++d.tested;
            if (depth_func(dfunc, z, dv)) shade_fragment(d, b, x, y, z, bc);
        }
    }
}

// Pixel (x, y) of a point or line, addressed as the reference addresses its
// attachments: texel x + W*y (SL_Framebuffer.cpp:33-38, SL_PointRasterizer.cpp:
// 56,94).  A snapped coordinate can equal W (NDC +1, or a line clipped to the
// viewport's right edge): column W of row y is texel 0 of row y + 1.  False
// for an address at or past W*H, which the reference writes outside the
// attachment's texels (into its allocation padding); those fragments are
// dropped here, and so are columns beyond W (only reachable by a line end
// that float cancellation in the clipper moved past the viewport edge).
bool linear_pixel(const Draw& d, int x, int y, int& px, int& py)
{
    if (x > d.tgt.W) return false;
    const size_t idx = (size_t)x + (size_t)d.tgt.W * (size_t)y;
    if (idx >= (size_t)d.tgt.W * (size_t)d.tgt.H) return false;
    px = (int)(idx % (size_t)d.tgt.W);
    py = (int)(idx / (size_t)d.tgt.W);
    return true;
}

// SL_PointRasterizer::render_point (SL_PointRasterizer.cpp:27-98): varyings
// pass to the fragment program unchanged.
void raster_point_bin(Draw& d, const Bin& b, int32_t rowThread, int32_t numRowThreads)
{
    const uint16_t x = (uint16_t)f2i(b.p[0].v[0]);
    const uint16_t y = (uint16_t)f2i(b.p[0].v[1]);
    int px, py;
    if (!linear_pixel(d, x, y, px, py)) return;
    // rows are split over threads by the addressed texel's row (the
    // reference splits by y, which lets two threads race on a wrapped texel)
    if (py % numRowThreads != rowThread) return;
    const float z = b.p[0].v[2];
    ++d.tested;
    if (!depth_func(d.sh->dtest, z, depth_read(d.tgt, (size_t)px + (size_t)d.tgt.W * (size_t)py))) return;
    V4 var[SL_SHADER_MAX_VARYING_VECTORS];
    for (unsigned i = 0; i < SL_SHADER_MAX_VARYING_VECTORS; ++i) var[i] = (i < d.nvary) ? b.var[0][i] : v4(0.f, 0.f, 0.f, 0.f);
    finish_fragment(d, px, py, z, var, b.seq);
}

// SL_LineRasterizer::render_line (SL_LineRasterizer.cpp:27-80) over
// sl_draw_line_bresenham (SL_Geometry.hpp:314-366, 16-bit integers): each
// pixel's position along the line is its distance from the first end times
// the inverse length; depth and varyings are mixed linearly by it
// (interpolate_line_varyings, SL_FragmentProcessor.cpp:119-145).
void raster_line_bin(Draw& d, const Bin& b, int32_t rowThread, int32_t numRowThreads)
{
    const float z0 = b.p[0].v[2], z1 = b.p[1].v[2];
    const float s0x = b.p[0].v[0], s0y = b.p[0].v[1];
    const float ex = b.p[1].v[0] - s0x, ey = b.p[1].v[1] - s0y;
    const float dist = 1.f / sqrtf((ex * ex) + (ey * ey));   // inversesqrt(length_squared(sc1 - sc0))

    auto pixel = [&](uint16_t x, uint16_t y) {
        int qx, qy;
        if (!linear_pixel(d, x, y, qx, qy)) return;
        if (qy % numRowThreads != rowThread) return;   // by the addressed texel's row, as for points
        const float px = (float)x - s0x, py = (float)y - s0y;
        const float len = sqrtf(((px * px) + (py * py)) + (0.f * 0.f) + (0.f * 0.f));   // length(p - p0)
        const float interp = len * dist;
        const float z = z0 + interp * (z1 - z0);                                     // mix(z0, z1, interp)
        ++d.tested;
        if (!depth_func(d.sh->dtest, z, depth_read(d.tgt, (size_t)qx + (size_t)d.tgt.W * (size_t)qy))) return;
        V4 var[SL_SHADER_MAX_VARYING_VECTORS];
        for (unsigned i = 0; i < SL_SHADER_MAX_VARYING_VECTORS; ++i)
            var[i] = (i < d.nvary) ? mix4(b.var[0][i], b.var[1][i], interp) : v4(0.f, 0.f, 0.f, 0.f);
        finish_fragment(d, qx, qy, z, var, b.seq);
    };

    // (uint16_t) casts of the screen coordinates, passed as int16_t
    int16_t x1 = (int16_t)(uint16_t)f2i(s0x), y1 = (int16_t)(uint16_t)f2i(s0y);
    int16_t x2 = (int16_t)(uint16_t)f2i(b.p[1].v[0]), y2 = (int16_t)(uint16_t)f2i(b.p[1].v[1]);
    const bool steep = std::abs(x1 - x2) < std::abs(y1 - y2);
    if (steep) { std::swap(x1, y1); std::swap(x2, y2); }
    if (x1 > x2) { std::swap(x1, x2); std::swap(y1, y2); }
    const int16_t dx = (int16_t)(x2 - x1);
    const int16_t dy = (int16_t)(y2 - y1);
    const int16_t dErr = (int16_t)(std::abs(dy) * 2);
    const int16_t yErr = (y2 > y1) ? 1 : -1;
    int16_t err = 0;
    int16_t y = y1;
    for (int16_t x = x1; x <= x2; ++x)
    {
        if (steep) pixel((uint16_t)y, (uint16_t)x);
        else pixel((uint16_t)x, (uint16_t)y);
        err = (int16_t)(err + dErr);
        if (err > dx)
        {
            y = (int16_t)(y + yErr);
            err = (int16_t)(err - 2 * dx);
        }
    }
}

// SL_TriRasterizer::render_triangle_simd, generic branch (:806-923), for the
// rows y = minY + offset + k*step owned by one thread.
void raster_bin(Draw& d, const Bin& b, int32_t rowThread, int32_t numRowThreads)
{
    if (d.prim == kPrimWire)
    {
        raster_wire_bin(d, b, rowThread, numRowThreads);
        return;
    }
    if (d.prim == kPrimLine)
    {
        raster_line_bin(d, b, rowThread, numRowThreads);
        return;
    }
    if (d.prim == kPrimPoint)
    {
        raster_point_bin(d, b, rowThread, numRowThreads);
        return;
    }
    const V4* P = b.p;
    const int32_t bboxMinY = f2i(std::min(std::min(P[0].v[1], P[1].v[1]), P[2].v[1]));
    const int32_t bboxMaxY = f2i(std::max(std::max(P[0].v[1], P[1].v[1]), P[2].v[1]));
    const int32_t n1 = numRowThreads - 1;
    const int32_t offset = n1 - ((bboxMinY + (n1 - rowThread)) % numRowThreads);  // sl_scanline_offset, SL_ScanlineBounds.hpp:31-50
    int32_t y = bboxMinY + offset;
    if (y >= bboxMaxY) return;

    const float dep[3] = {P[0].v[2], P[1].v[2], P[2].v[2]};
    const Scanline sc = scanline_init(P[0], P[1], P[2]);
    const int dfunc = d.sh->dtest;
    const int W = d.tgt.W;

    do
    {
        const float yf = (float)y;
        int32_t xMin, xMax;
        scanline_step(sc, yf, xMin, xMax);
        if ((uint32_t)xMin < (uint32_t)xMax)
        {
            float bcY[4], bc[4][4], bcX[4];
            for (int k = 0; k < 4; ++k)
            {
                bcY[k] = fmaf(b.by.v[k], yf, b.bz.v[k]);
                bcX[k] = b.bx.v[k] * 4.f;
            }
            for (int i = 0; i < 4; ++i)
                for (int k = 0; k < 4; ++k)
                    bc[i][k] = ((float)(xMin + i) * b.bx.v[k]) + bcY[k];

            int32_t x0 = xMin;
            do
            {
                for (int i = 0; i < 4; ++i)
                {
                    const int32_t x = x0 + i;
                    if (x < xMax)
                    {
                        // z = depth * bc (SL_TriRasterizer.cpp:868), 4th term = 0 * bc[3]
                        const float z = (((dep[0] * bc[i][0]) + (dep[1] * bc[i][1])) + (dep[2] * bc[i][2])) + (0.f * bc[i][3]);
                        const float dv = depth_read(d.tgt, (size_t)x + (size_t)W * (size_t)y);
                        ++d.tested;
                        if (depth_func(dfunc, z, dv)) shade_fragment(d, b, x, y, z, bc[i]);
                    }
                }
                for (int i = 0; i < 4; ++i)
                    for (int k = 0; k < 4; ++k)
                        bc[i][k] = bc[i][k] + bcX[k];
                x0 += 4;
            }
            while (x0 < xMax);
        }
        y += numRowThreads;
    }
    while (y < bboxMaxY);
}

/*-----------------------------------------------------------------------------
 * Vertex stage (SL_TriProcessor.cpp:31-118,248-342,370-519,527-640)
-----------------------------------------------------------------------------*/
// face_determinant, generic branch (:101-116)
float face_determinant(const V4& p0, const V4& p1, const V4& p2)
{
    const float w0 = rcpf(p0.v[3]);
    const float w1 = rcpf(p1.v[3]);
    const float w2 = rcpf(p2.v[3]);
    V4 v0, v1, v2;
    for (int k = 0; k < 4; ++k) { v0.v[k] = p0.v[k] * w0; v1.v[k] = p1.v[k] * w1; v2.v[k] = p2.v[k] * w2; }
    const V4 xABC = v4(v0.v[0], v1.v[0], v2.v[0], 0.f);
    const V4 yBCA = v4(v1.v[1], v2.v[1], v0.v[1], 0.f);
    const V4 yCAB = v4(v2.v[1], v0.v[1], v1.v[1], 0.f);
    V4 diff;
    for (int k = 0; k < 4; ++k) diff.v[k] = yBCA.v[k] - yCAB.v[k];
    return dot4(xABC, diff);
}

// sl_ndc_clip_status, generic branch (SL_PostVertexTransform.hpp:653-673)
int clip_status(const V4& c0, const V4& c1, const V4& c2)
{
    auto inside = [](const V4& c) {
        const float w = c.v[3];
        bool le = true, ge = true;
        for (int k = 0; k < 4; ++k) { le = le && (c.v[k] <= w); ge = ge && (c.v[k] >= -w); }
        return le && ge;
    };
    const int vis = (inside(c0) && inside(c1) && inside(c2)) ? 3 : 0;
    const int part = (c0.v[3] > 0.f || c1.v[3] > 0.f || c2.v[3] > 0.f) ? 1 : 0;
    return vis | part;
}

// sl_perspective_divide (:239-249) + sl_ndc_to_screen_coords (:508-524)
void divide_and_snap(V4& p, const float vp[4])
{
    const float wInv = rcpf(p.v[3]);
    const float x = p.v[0] * wInv, y = p.v[1] * wInv, z = p.v[2] * wInv;
    const float hw = vp[2] * 0.5f;
    const float hh = vp[3] * 0.5f;
    p.v[0] = maxf_ref(floorf(fmaf(x + 1.f, hw, vp[0])), 0.f);
    p.v[1] = maxf_ref(floorf(fmaf(y + 1.f, hh, vp[1])), 0.f);
    p.v[2] = z;
    p.v[3] = wInv;
}

// The polygon loop of SL_TriProcessor::clip_and_process_tris (:370-499):
// clips the triangle nverts[0..2] (with nv varyings per vertex) against the
// six planes in place, visiting the planes in the reference's order; returns
// the polygon's size, 0 when a plane leaves nothing.
unsigned clip_polygon(V4 nverts[9], V4 nvar[9][SL_SHADER_MAX_VARYING_VECTORS], unsigned nv)
{
    const V4 edges[6] = {
        v4(1.f, 0.f, 0.f, 1.f), v4(-1.f, 0.f, 0.f, 1.f),
        v4(0.f, 1.f, 0.f, 1.f), v4(0.f, -1.f, 0.f, 1.f),
        v4(0.f, 0.f, 1.f, 1.f), v4(0.f, 0.f, -1.f, 1.f)};
    unsigned n = 3;
    V4 tverts[9];
    V4 tvar[9][SL_SHADER_MAX_VARYING_VECTORS];
    for (const V4& edge : edges)
    {
        unsigned nn = 0;
        unsigned j = n - 1;
        V4 q0 = nverts[n - 1];
        float t0 = dot4(q0, edge);
        bool vis0 = !std::signbit(t0);
        for (unsigned k = 0; k < n; ++k)
        {
            const V4 q1 = nverts[k];
            const float t1 = dot4(q1, edge);
            const bool vis1 = !std::signbit(t1);
            if (vis0 != vis1)
            {
                const float t = t0 * rcpf(t0 - t1);
                tverts[nn] = mix4(q0, q1, t);
                for (unsigned i = 0; i < nv; ++i) tvar[nn][i] = mix4(nvar[j][i], nvar[k][i], t);
                ++nn;
            }
            if (vis1)
            {
                tverts[nn] = q1;
                for (unsigned i = 0; i < nv; ++i) tvar[nn][i] = nvar[k][i];
                ++nn;
            }
            j = k; q0 = q1; t0 = t1; vis0 = vis1;
        }
        if (!nn) return 0;
        n = nn;
        for (unsigned k = 0; k < n; ++k)
        {
            nverts[k] = tverts[k];
            for (unsigned i = 0; i < nv; ++i) nvar[k][i] = tvar[k][i];
        }
    }
    return n;
}

struct VertexStage
{
    Draw& d;
    std::vector<Bin>* out;     // when non-null, bins are collected instead of rasterised
    explicit VertexStage(Draw& dr) : d(dr), out(nullptr) {}

    void emit(Bin& b)
    {
        ++d.binned;
        if (out) out->push_back(b);
        else raster_bin(d, b, 0, 1);
    }

    // SL_TriProcessor::push_bin (:248-342)
    void push_bin(uint32_t seq, const TVert& a, const TVert& b, const TVert& c)
    {
        const V4& p0 = a.pos;
        const V4& p1 = b.pos;
        const V4& p2 = c.pos;
        for (int k = 0; k < 2; ++k)
            if (!std::isfinite(p0.v[k]) || !std::isfinite(p1.v[k]) || !std::isfinite(p2.v[k])) return;

        const float minX = std::min(std::min(p0.v[0], p1.v[0]), p2.v[0]);
        const float maxX = std::max(std::max(p0.v[0], p1.v[0]), p2.v[0]);
        const float minY = std::min(std::min(p0.v[1], p1.v[1]), p2.v[1]);
        const float maxY = std::max(std::max(p0.v[1], p1.v[1]), p2.v[1]);
        if ((maxX - minX < 1.f) || (maxY - minY < 1.f)) return;

        // 2D cross of (p2-p1) and (p0-p1) (:271-281)
        const float zyx = p2.v[0] - p1.v[0], zyy = p2.v[1] - p1.v[1];
        const float xyx = p0.v[0] - p1.v[0], xyy = p0.v[1] - p1.v[1];
        const float denom = rcpf((zyx * xyy) - (zyy * xyx));

        // ddx = cross(Y, 1), ddy = cross(1, X), ddz = cross(X, Y) (:285-287)
        const float X[3] = {p0.v[0], p1.v[0], p2.v[0]};
        const float Y[3] = {p0.v[1], p1.v[1], p2.v[1]};
        const float ddx[4] = {(Y[1] * 1.f) - (Y[2] * 1.f), (Y[2] * 1.f) - (Y[0] * 1.f), (Y[0] * 1.f) - (Y[1] * 1.f), 0.f};
        const float ddy[4] = {(1.f * X[2]) - (1.f * X[1]), (1.f * X[0]) - (1.f * X[2]), (1.f * X[1]) - (1.f * X[0]), 0.f};
        const float ddz[4] = {(X[1] * Y[2]) - (X[2] * Y[1]), (X[2] * Y[0]) - (X[0] * Y[2]), (X[0] * Y[1]) - (X[1] * Y[0]), 0.f};

        Bin bin;
        bin.p[0] = p0; bin.p[1] = p1; bin.p[2] = p2;
        for (int k = 0; k < 4; ++k)
        {
            bin.bx.v[k] = ddx[k] * denom;
            bin.by.v[k] = ddy[k] * denom;
            bin.bz.v[k] = ddz[k] * denom;
        }
        const unsigned nv = d.nvary;
        for (unsigned i = 0; i < nv; ++i)
        {
            bin.var[0][i] = a.var[i];
            bin.var[1][i] = b.var[i];
            bin.var[2][i] = c.var[i];
        }
        bin.seq = seq;
        emit(bin);
    }

    // SL_TriProcessor::clip_and_process_tris (:370-519)
    void clip_and_process(uint32_t seq, const TVert& a, const TVert& b, const TVert& c)
    {
        const unsigned nv = d.nvary;
        V4 nverts[9];
        V4 nvar[9][SL_SHADER_MAX_VARYING_VECTORS];
        nverts[0] = a.pos; nverts[1] = b.pos; nverts[2] = c.pos;
        for (unsigned i = 0; i < nv; ++i) { nvar[0][i] = a.var[i]; nvar[1][i] = b.var[i]; nvar[2][i] = c.var[i]; }
        const unsigned n = clip_polygon(nverts, nvar, nv);
        if (n < 3) return;

        for (unsigned k = 0; k < n; ++k) divide_and_snap(nverts[k], d.vpDims);

        TVert f0, f1, f2;
        f0.pos = nverts[0];
        for (unsigned i = 0; i < nv; ++i) f0.var[i] = nvar[0][i];
        for (unsigned i = n - 2; i--;)                // reverse fan sharing one primIndex (:505-518)
        {
            f1.pos = nverts[i + 1];
            f2.pos = nverts[i + 2];
            for (unsigned v = 0; v < nv; ++v) { f1.var[v] = nvar[i + 1][v]; f2.var[v] = nvar[i + 2][v]; }
            push_bin(seq, f0, f1, f2);
        }
    }

    // SL_PointProcessor::process_verts (SL_PointProcessor.cpp:64-113): one
    // vertex, visible when w > 0 and every component lies in [-w, w]
    // (sl_ndc_clip_status, SL_PostVertexTransform.hpp:536-554)
    void process_point(uint32_t seq, size_t vid, size_t instanceId)
    {
        TVert v;
        for (auto& x : v.var) x = v4(0.f, 0.f, 0.f, 0.f);
        v.pos = mat_mul(d.scissorMat, run_vertex_program(d, vid, instanceId, v.var));
        const float w = v.pos.v[3];
        bool in = w > 0.f;
        for (int k = 0; k < 4; ++k) in = in && v.pos.v[k] <= w && v.pos.v[k] >= -w;
        if (!in) return;
        divide_and_snap(v.pos, d.vpDims);
        Bin bin;
        bin.p[0] = v.pos;
        for (unsigned i = 0; i < d.nvary; ++i) bin.var[0][i] = v.var[i];
        bin.seq = seq;
        emit(bin);
    }

    void push_line(uint32_t seq, const TVert& a, const TVert& b)   // SL_LineProcessor::push_bin (:26-62)
    {
        Bin bin;
        bin.p[0] = a.pos;
        bin.p[1] = b.pos;
        for (unsigned i = 0; i < d.nvary; ++i)
        {
            bin.var[0][i] = a.var[i];
            bin.var[1][i] = b.var[i];
        }
        bin.seq = seq;
        emit(bin);
    }

    // SL_LineProcessor::clip_and_process_lines (:69-169): Liang-Barsky in
    // screen space against viewportDims read as (xMin, yMin, xMax, yMax),
    // i.e. the rectangle [x, w] x [y, h] of the viewport rect (x, y, w, h).
    void clip_line(uint32_t seq, const TVert& a, const TVert& b)
    {
        const float dx = b.pos.v[0] - a.pos.v[0];
        const float dy = b.pos.v[1] - a.pos.v[1];
        const float xMin = d.vpDims[0], yMin = d.vpDims[1], xMax = d.vpDims[2], yMax = d.vpDims[3];
        if (fabsf(dx) < kLineEpsilon && fabsf(dy) < kLineEpsilon)
        {
            if (a.pos.v[0] >= xMin && a.pos.v[0] <= xMax && a.pos.v[1] >= yMin && a.pos.v[1] <= yMax)
            {
                push_line(seq, a, b);
                return;
            }
        }
        auto clip_segment = [](float num, float denom, float& tE, float& tL) -> bool {
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
        };
        TVert p0 = a, p1 = b;
        const float ex = p0.pos.v[0] - p1.pos.v[0], ey = p0.pos.v[1] - p1.pos.v[1];
        const float dist = 1.f / sqrtf((ex * ex) + (ey * ey));   // inversesqrt(length_squared(...))
        float tE = 0.f, tL = 1.f;
        const float x1 = p0.pos.v[0], y1 = p0.pos.v[1];
        if (clip_segment(xMin - x1, dx, tE, tL) && clip_segment(x1 - xMax, -dx, tE, tL) &&
            clip_segment(yMin - y1, dy, tE, tL) && clip_segment(y1 - yMax, -dy, tE, tL))
        {
            if (tL < 1.f)
            {
                p1.pos.v[0] = x1 + tL * dx;
                p1.pos.v[1] = y1 + tL * dy;
            }
            if (tE > 0.f)
            {
                p0.pos.v[0] = x1 + tE * dx;
                p0.pos.v[1] = y1 + tE * dy;
            }
            V4 d0, d1;
            for (int k = 0; k < 4; ++k) { d0.v[k] = p0.pos.v[k] - a.pos.v[k]; d1.v[k] = p1.pos.v[k] - b.pos.v[k]; }
            const float interp0 = sqrtf(dot4(d0, d0)) * dist;   // length(p0 - a) * dist
            const float interp1 = sqrtf(dot4(d1, d1)) * dist;
            p0.pos.v[2] = a.pos.v[2] + interp0 * (b.pos.v[2] - a.pos.v[2]);
            p0.pos.v[3] = a.pos.v[3] + interp0 * (b.pos.v[3] - a.pos.v[3]);
            p1.pos.v[2] = b.pos.v[2] + interp1 * (a.pos.v[2] - b.pos.v[2]);
            p1.pos.v[3] = b.pos.v[3] + interp1 * (a.pos.v[3] - b.pos.v[3]);
            for (unsigned i = 0; i < d.nvary; ++i)
            {
                p0.var[i] = mix4(a.var[i], b.var[i], interp0);
                p1.var[i] = mix4(b.var[i], a.var[i], interp1);
            }
            push_line(seq, p0, p1);
        }
    }

    // SL_LineProcessor::process_verts (:196-276)
    void process_line(uint32_t seq, size_t vid0, size_t vid1, size_t instanceId)
    {
        TVert v[2];
        const size_t ids[2] = {vid0, vid1};
        for (int k = 0; k < 2; ++k)
        {
            for (auto& x : v[k].var) x = v4(0.f, 0.f, 0.f, 0.f);
            v[k].pos = mat_mul(d.scissorMat, run_vertex_program(d, ids[k], instanceId, v[k].var));
        }
        if (v[0].pos.v[3] < 0.f || v[1].pos.v[3] < 0.f) return;
        // sl_ndc_clip_status, two vertices (SL_PostVertexTransform.hpp:560-587)
        bool all = true;
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 4; ++k) all = all && v[j].pos.v[k] <= v[j].pos.v[3] && v[j].pos.v[k] >= -v[j].pos.v[3];
        const int status = (all ? 3 : 0) | ((v[0].pos.v[3] > 0.f || v[1].pos.v[3] > 0.f) ? 1 : 0);
        if (status == 0) return;
        divide_and_snap(v[0].pos, d.vpDims);
        divide_and_snap(v[1].pos, d.vpDims);
        if (status == 3) push_line(seq, v[0], v[1]);
        else clip_line(seq, v[0], v[1]);
    }

    void process_triangle(uint32_t seq, size_t vid0, size_t vid1, size_t vid2, size_t instanceId)
    {
        TVert tv[3];
        const size_t ids[3] = {vid0, vid1, vid2};
        for (int k = 0; k < 3; ++k)
        {
            for (auto& v : tv[k].var) v = v4(0.f, 0.f, 0.f, 0.f);
            tv[k].pos = run_vertex_program(d, ids[k], instanceId, tv[k].var);
            if (d.visSlot >= 0) tv[k].var[d.visSlot] = v4(k == 0 ? 1.f : 0.f, k == 1 ? 1.f : 0.f, k == 2 ? 1.f : 0.f, 0.f);
        }
        const int cull = d.sh->cull;
        if (cull != SL_CULL_OFF)                       // (:597-612)
        {
            const float det = face_determinant(tv[0].pos, tv[1].pos, tv[2].pos);
            const bool culled = (cull == SL_CULL_FRONT_FACE) ^ (bool)std::signbit(det);
            if (culled) return;
        }
        for (int k = 0; k < 3; ++k) tv[k].pos = mat_mul(d.scissorMat, tv[k].pos);   // (:615-617)

        const int status = clip_status(tv[0].pos, tv[1].pos, tv[2].pos);
        if (status == 3)
        {
            for (int k = 0; k < 3; ++k) divide_and_snap(tv[k].pos, d.vpDims);
            push_bin(seq, tv[0], tv[1], tv[2]);
        }
        else if (status == 1)
        {
            clip_and_process(seq, tv[0], tv[1], tv[2]);
        }
    }
};

size_t index_at(const Ibo& ibo, size_t i)
{
    if (i >= ibo.count) return 0;
    const uint8_t* p = ibo.bytes.data() + i * ibo.bpe;
    switch (ibo.type)
    {
        case VERTEX_DATA_BYTE: return p[0];
        case VERTEX_DATA_SHORT: { uint16_t v; std::memcpy(&v, p, 2); return v; }
        default: { uint32_t v; std::memcpy(&v, p, 4); return v; }
    }
}

// One (mesh, instance) pass of a draw call, as the reference loops them in
// SL_TriProcessor::execute (SL_TriProcessor.cpp:677-713).
struct Pass { const sl_mesh* mesh; size_t instance; size_t triBase; size_t numTris; uint32_t seqBase; };

// Primitive g of a pass: its element indices, through the index buffer when
// the mesh's mode says so, and the processor of the draw's kind.
void process_prim(VertexStage& vs, const Draw& d, const Pass& P, size_t g)
{
    const size_t t = g - P.triBase;
    const uint32_t seq = P.seqBase + (uint32_t)t + 1u;
    const bool indexed = mesh_indexed(d.prim, P.mesh->mode);
    auto vid = [&](size_t e) { return indexed ? index_at(*d.ibo, e) : e; };
    const size_t i = P.mesh->elementBegin + verts_per_prim(d.prim) * t;
    if (d.prim == kPrimPoint) vs.process_point(seq, vid(i), P.instance);
    else if (d.prim == kPrimLine) vs.process_line(seq, vid(i), vid(i + 1), P.instance);
    else vs.process_triangle(seq, vid(i), vid(i + 1), vid(i + 2), P.instance);
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

    const float inv[4] = {rcpf(fboW), rcpf(fboH), rcpf(fboW), rcpf(fboH)};
    float x = sc[0] * inv[0], y = sc[1] * inv[1], w = sc[2] * inv[2], h = sc[3] * inv[3];
    if (x < 0.f) { w += x; x = 0.f; }
    if (y < 0.f) { h += y; y = 0.f; }
    w = std::min(1.f - x, w);
    h = std::min(1.f - y, h);
    const float nm00 = rcpf(w), nm11 = rcpf(h);
    const float nm30 = nm00 - 1.f, nm31 = nm11 - 1.f;
    const float m30 = fmaf(x, (-2.f * nm00), nm30);
    const float m31 = fmaf(y, (-2.f * nm11), nm31);
    const float m[16] = {nm00, 0.f, 0.f, 0.f, 0.f, nm11, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f, m30, m31, 0.f, 1.f};
    std::memcpy(mat, m, sizeof(m));
}

int setup_draw(sl_context* c, size_t shaderId, size_t fboId, Draw& d)
{
    if (shaderId >= c->shaders.size() || fboId >= c->fbos.size()) return SL_ERR_INVALID_HANDLE;
    const Shader& sh = c->shaders[shaderId];
    Fbo& f = c->fbos[fboId];
    if (f.depth >= c->texs.size() || !c->texs[f.depth].valid) return SL_ERR_STATE;
    Tex& dt = c->texs[f.depth];
    if (dt.type != SL_COLOR_R_FLOAT && dt.type != SL_COLOR_R_HALF) return SL_ERR_UNSUPPORTED;
    d.ctx = c;
    d.sh = &sh;
    d.uni = (sh.ubo < c->ubos.size()) ? c->ubos[sh.ubo].bytes : nullptr;
    if (d.uni == nullptr && (sh.vsProg != SL_PROGRAM_PASSTHROUGH_GOURAUD || sh.fsProg != SL_PROGRAM_PASSTHROUGH_GOURAUD)) return SL_ERR_STATE;
    d.tgt.W = dt.w;
    d.tgt.H = dt.h;
    d.tgt.depth = (dt.type == SL_COLOR_R_FLOAT) ? reinterpret_cast<float*>(dt.data.data()) : nullptr;
    d.tgt.depthH = (dt.type == SL_COLOR_R_HALF) ? reinterpret_cast<uint16_t*>(dt.data.data()) : nullptr;
    d.tgt.color8 = nullptr;
    d.tgt.colorF = nullptr;
    d.tgt.comps = 4;
    d.tgt.packed = 0;
    for (int k = 0; k < 3; ++k) { d.tgt.extra8[k] = nullptr; d.tgt.extraF[k] = nullptr; d.tgt.extraComps[k] = 4; d.tgt.extraPacked[k] = 0; }
    // several outputs only from the program that writes them
    if (sh.nout > 1 && sh.fsProg != SL_PROGRAM_LAMBERT_MRT) return SL_ERR_UNSUPPORTED;
    for (unsigned k = 0; k < sh.nout; ++k)
    {
        if (f.colors.size() <= k || f.colors[k] >= c->texs.size() || !c->texs[f.colors[k]].valid) return SL_ERR_STATE;
        Tex& ct = c->texs[f.colors[k]];
        if (ct.w != dt.w || ct.h != dt.h) return SL_ERR_STATE;
        uint8_t* p8 = nullptr;
        float* pF = nullptr;
        int n = 0, packed = 0;
        if (u8_components(ct.type)) { p8 = ct.data.data(); n = u8_components(ct.type); }
        else if (float_components(ct.type)) { pF = reinterpret_cast<float*>(ct.data.data()); n = float_components(ct.type); }
        else if (packed_format(ct.type).bytes) { p8 = ct.data.data(); n = packed_format(ct.type).bytes; packed = ct.type; }
        else return SL_ERR_UNSUPPORTED;
        if (k == 0) { d.tgt.color8 = p8; d.tgt.colorF = pF; d.tgt.comps = n; d.tgt.packed = packed; }
        else { d.tgt.extra8[k - 1] = p8; d.tgt.extraF[k - 1] = pF; d.tgt.extraComps[k - 1] = n; d.tgt.extraPacked[k - 1] = packed; }
    }
    d.tgt.dbgLast = (f.dbgLast < c->texs.size()) ? reinterpret_cast<uint32_t*>(c->texs[f.dbgLast].data.data()) : nullptr;
    d.tgt.dbgCount = (f.dbgCount < c->texs.size()) ? reinterpret_cast<uint32_t*>(c->texs[f.dbgCount].data.data()) : nullptr;
    d.tgt.vis = nullptr;
    d.nvary = sh.nvary;
    d.visSlot = -1;
    if (f.vis < c->texs.size() && c->texs[f.vis].valid)
    {
        Tex& vt = c->texs[f.vis];
        if (vt.type != SL_COLOR_RGBA_FLOAT || vt.w != dt.w || vt.h != dt.h) return SL_ERR_STATE;
        if ((d.tgt.colorF == nullptr || d.tgt.comps != 4) && sh.nout > 0) return SL_ERR_UNSUPPORTED;
        if (sh.nout > 1) return SL_ERR_UNSUPPORTED;
        if (d.tgt.depthH) return SL_ERR_UNSUPPORTED;
        if (sh.nvary >= SL_SHADER_MAX_VARYING_VECTORS) return SL_ERR_UNSUPPORTED;
        d.tgt.vis = reinterpret_cast<float*>(vt.data.data());
        d.visSlot = sh.nvary;
        d.nvary = sh.nvary + 1u;
    }
    viewport_setup(c, (float)dt.w, (float)dt.h, d.scissorMat, d.vpDims);
    return SL_OK;
}

int bind_mesh(sl_context* c, const sl_mesh& m, Draw& d)
{
    if (m.vaoId >= c->vaos.size()) return SL_ERR_INVALID_HANDLE;
    d.vao = &c->vaos[m.vaoId];
    d.vbo = (d.vao->vbo < c->vbos.size()) ? &c->vbos[d.vao->vbo] : nullptr;
    d.ibo = (d.vao->ibo < c->ibos.size()) ? &c->ibos[d.vao->ibo] : nullptr;
    if (prim_of(m.mode) == kPrimBad) return SL_ERR_UNSUPPORTED;
    if (mesh_indexed(d.prim, m.mode) && d.ibo == nullptr) return SL_ERR_STATE;
    return SL_OK;
}

/*-----------------------------------------------------------------------------
 * Multithreaded timing mode: the SL_ProcessorPool scheme (a persistent pool,
 * strided vertex work, row-interleaved rasterisation, barriers between the
 * phases of each batch, SL_ProcessorPool.cpp:280-327,
 * SL_VertexProcessor.cpp:84-181).  Bins are merged in submission order before
 * rasterisation, so the output equals the single-threaded one.
-----------------------------------------------------------------------------*/

size_t index_at(const Ibo& ibo, size_t i);
int bind_mesh(sl_context* c, const sl_mesh& m, Draw& d);

class SpinBarrier
{
  public:
    explicit SpinBarrier(unsigned n) : mN(n) {}
    void wait()
    {
        const unsigned gen = mGen.load(std::memory_order_acquire);
        if (mCount.fetch_add(1, std::memory_order_acq_rel) + 1 == mN)
        {
            mCount.store(0, std::memory_order_relaxed);
            mGen.fetch_add(1, std::memory_order_acq_rel);
        }
        else
        {
            while (mGen.load(std::memory_order_acquire) == gen) std::this_thread::yield();
        }
    }
  private:
    const unsigned mN;
    std::atomic<unsigned> mCount{0};
    std::atomic<unsigned> mGen{0};
};

void run_draw_threads(sl_context* c, Draw& proto, const std::vector<Pass>& passes, size_t totalTris)
{
    const unsigned N = std::max(1u, c->numThreads);
    const size_t window = 16384 * (size_t)N;
    std::vector<std::vector<Bin>> perThread(N);
    std::vector<const Bin*> merged;
    std::vector<uint64_t> tested(N, 0), shaded(N, 0), binned(N, 0);
    SpinBarrier bar(N);

    // triangle g of the draw -> its pass (passes are in order, triBase ascending)
    auto pass_of = [&](size_t g, size_t hint) {
        while (hint + 1 < passes.size() && passes[hint + 1].triBase <= g) ++hint;
        return hint;
    };

    auto worker = [&](unsigned t) {
        Draw d = proto;
        size_t hint = 0;
        for (size_t w0 = 0; w0 < totalTris; w0 += window)
        {
            const size_t w1 = std::min(totalTris, w0 + window);
            {
                perThread[t].clear();
                for (size_t g = w0 + t; g < w1; g += N)
                {
                    hint = pass_of(g, hint);
                    const Pass& P = passes[hint];
                    bind_mesh(c, *P.mesh, d);
                    VertexStage vs(d);
                    vs.out = &perThread[t];
                    process_prim(vs, d, P, g);
                }
            }
            bar.wait();
            if (t == 0)
            {
                // merge: bins of one triangle are contiguous in its thread's
                // list and carry its sequence number; walk triangles in order.
                merged.clear();
                std::vector<size_t> pos(N, 0);
                size_t mh = 0;
                for (size_t g = w0; g < w1; ++g)
                {
                    const unsigned o = (unsigned)((g - w0) % N);
                    const std::vector<Bin>& L = perThread[o];
                    if (pos[o] >= L.size()) continue;
                    mh = pass_of(g, mh);
                    const uint32_t sq = passes[mh].seqBase + (uint32_t)(g - passes[mh].triBase) + 1u;
                    while (pos[o] < L.size() && L[pos[o]].seq == sq) merged.push_back(&L[pos[o]++]);
                }
            }
            bar.wait();
            for (const Bin* b : merged) raster_bin(d, *b, (int32_t)t, (int32_t)N);
            bar.wait();
        }
        tested[t] = d.tested;
        shaded[t] = d.shaded;
        binned[t] = d.binned;
    };
    std::vector<std::thread> th;
    for (unsigned t = 1; t < N; ++t) th.emplace_back(worker, t);
    worker(0);
    for (auto& x : th) x.join();
    for (unsigned t = 0; t < N; ++t) { proto.tested += tested[t]; proto.shaded += shaded[t]; proto.binned += binned[t]; }
}

int run_draws(sl_context* c, const sl_mesh* meshes, size_t numMeshes, size_t numInstances, size_t shaderId, size_t fboId)
{
    Draw proto;
    int rc = setup_draw(c, shaderId, fboId, proto);
    if (rc != SL_OK) return rc;
    proto.prim = prim_of(meshes[0].mode);
    if (proto.prim == kPrimBad) return SL_ERR_UNSUPPORTED;
    // visibility records, and what is built on them, cover filled triangles
    if (proto.tgt.vis && proto.prim != kPrimFill) return SL_ERR_UNSUPPORTED;
    auto t0 = std::chrono::steady_clock::now();
    const size_t nIter = (numInstances == 1) ? numMeshes : numInstances;
    std::vector<Pass> passes;
    size_t total = 0;
    for (size_t it = 0; it < nIter; ++it)
    {
        const sl_mesh& m = (numInstances == 1) ? meshes[it] : meshes[0];
        rc = bind_mesh(c, m, proto);
        if (rc != SL_OK) return rc;
        // whole primitives only (SL_TriProcessor.cpp:569-574; a trailing
        // partial line or triangle is skipped)
        const size_t n = (m.elementEnd > m.elementBegin) ? (m.elementEnd - m.elementBegin) / verts_per_prim(proto.prim) : 0;
        passes.push_back(Pass{&m, (numInstances == 1) ? 0 : it, total, n, c->triSeq});
        if (proto.tgt.vis && n > 0)
        {
            Fbo& f = c->fbos[fboId];
            size_t tex = SL_INVALID_ID;
            if (proto.uni) tex = reinterpret_cast<const sl_std_uniforms*>(proto.uni)->texture;
            SavedPass sp{m, (numInstances == 1) ? 0 : it, n, c->triSeq, shaderId, tex, f.savedDraws, {}, {}, {}};
            if (numInstances == 1 || it + 1 == nIter) f.savedDraws += 1;
            if (proto.uni) sp.uni.assign(proto.uni, proto.uni + SL_MAX_UNIFORM_BUFFER_SIZE);
            std::memcpy(sp.scissorMat, proto.scissorMat, sizeof(sp.scissorMat));
            std::memcpy(sp.vpDims, proto.vpDims, sizeof(sp.vpDims));
            f.saved.push_back(std::move(sp));
            if (proto.sh->blend != SL_BLEND_OFF) f.savedBlend = true;
        }
        c->triSeq += (uint32_t)n;
        c->stats.triangles_submitted += n;
        total += n;
    }
    if (c->numThreads > 1)
    {
        run_draw_threads(c, proto, passes, total);
    }
    else
    {
        for (const Pass& P : passes)
        {
            bind_mesh(c, *P.mesh, proto);
            VertexStage vs(proto);
            for (size_t t = 0; t < P.numTris; ++t) process_prim(vs, proto, P, P.triBase + t);
        }
    }
    c->stats.fragments_tested += proto.tested;
    c->stats.fragments_shaded += proto.shaded;
    c->stats.triangles_binned += proto.binned;
    auto t1 = std::chrono::steady_clock::now();
    c->stats.ms_total += std::chrono::duration<double, std::milli>(t1 - t0).count();
    c->stats.frames += 1;
    return SL_OK;
}

void clear_texture_bytes(Tex& t, const void* value, uint32_t bpt)
{
    const size_t n = (size_t)t.w * t.h * std::max<uint16_t>(t.d, 1);
    uint8_t* p = t.data.data();
    for (size_t i = 0; i < n; ++i) std::memcpy(p + i * bpt, value, bpt);
}

int clear_color(sl_context* c, size_t fbo, unsigned att, const double rgba[4])
{
    if (fbo >= c->fbos.size()) return SL_ERR_INVALID_HANDLE;
    Fbo& f = c->fbos[fbo];
    if (att >= f.colors.size() || f.colors[att] >= c->texs.size()) return SL_ERR_INVALID_HANDLE;
    Tex& t = c->texs[f.colors[att]];
    // sl_match_color_for_type (SL_Color.cpp:137-180): the clear colour cast to the texel type
    if (const int n = u8_components(t.type))
    {
        uint8_t v[4];
        for (int k = 0; k < 4; ++k) v[k] = (uint8_t)(int64_t)(rgba[k] * 255.0);   // color_cast<uint8_t,double>
        clear_texture_bytes(t, v, (uint32_t)n);
    }
    else if (const int n = float_components(t.type))
    {
        float v[4];
        for (int k = 0; k < 4; ++k) v[k] = (float)rgba[k];
        clear_texture_bytes(t, v, 4u * (uint32_t)n);
    }
    else if (packed_format(t.type).bytes)
    {
        // rgba_cast<packed, double>: the products in double, then the field's integer type
        const PackedFmt f = packed_format(t.type);
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
        clear_texture_bytes(t, &w, (uint32_t)f.bytes);
    }
    else return SL_ERR_UNSUPPORTED;
    return SL_OK;
}

int clear_depth(sl_context* c, size_t fbo, double depth)
{
    if (fbo >= c->fbos.size()) return SL_ERR_INVALID_HANDLE;
    Fbo& f = c->fbos[fbo];
    if (f.depth >= c->texs.size()) return SL_ERR_STATE;
    Tex& t = c->texs[f.depth];
    if (t.type == SL_COLOR_R_FLOAT)
    {
        const float v = (float)depth;
        clear_texture_bytes(t, &v, 4);
    }
    else if (t.type == SL_COLOR_R_HALF)
    {
        const uint16_t v = float_to_half((float)depth);   // (half)(float)depth, SL_Context.cpp:946-948
        clear_texture_bytes(t, &v, 2);
    }
    else return SL_ERR_UNSUPPORTED;
    if (f.dbgLast < c->texs.size()) { const uint32_t z = 0; clear_texture_bytes(c->texs[f.dbgLast], &z, 4); }
    if (f.dbgCount < c->texs.size()) { const uint32_t z = 0; clear_texture_bytes(c->texs[f.dbgCount], &z, 4); }
    if (f.vis < c->texs.size()) { const float z[4] = {0.f, 0.f, 0.f, 0.f}; clear_texture_bytes(c->texs[f.vis], z, 16); }
    f.saved.clear();
    f.savedDraws = 0;
    f.savedStale = false;
    f.savedBlend = false;
    return SL_OK;
}

// Resources changed under the saved passes: the backward pass refuses them
// (the GPU library keeps device pointers in its saved draws).
void stale_saved(sl_context* c)
{
    for (Fbo& f : c->fbos)
        if (!f.saved.empty()) f.savedStale = true;
}

/*-----------------------------------------------------------------------------
 * Fixed-visibility adjoint (no reference counterpart: the reference renders
 * forward only).  The scalar restatement of the backward kernels: for every
 * pixel whose visibility record names a triangle, rebuild the winning
 * fragment's varyings from the stored vertex weights b in the forward's
 * operation order, then add
 *   dL/d colour[v_k] += b_k * (dL/dC . texel)
 *   dL/d texel[tap_j] += w_j * (dL/dC . colour)
 * Each product is formed in double from the float forward values and summed
 * in double in pixel order; the sums are added to the float gradient buffers
 * at the end.
-----------------------------------------------------------------------------*/
struct Taps { size_t i[4]; float w[4]; };

Taps bilinear_taps(const Tex& t, float u, float v)
{
    // same arithmetic as sample_bilinear_rgbaf (SL_Sampler.hpp:186-218)
    const float xf = wrap_repeat(u) * (float)t.w;
    const float yf = wrap_repeat(v) * (float)t.h;
    const uint16_t xi0 = to_u16(xf);
    const uint16_t yi0 = to_u16(yf);
    const uint16_t xi1 = (uint16_t)std::min<uint32_t>((uint32_t)xi0 + 1u, t.w);
    const uint16_t yi1 = (uint16_t)std::min<uint32_t>((uint32_t)yi0 + 1u, t.h);
    const float dx = xf - (float)xi0;
    const float dy = yf - (float)yi0;
    const float omdx = 1.f - dx;
    const float omdy = 1.f - dy;
    Taps r;
    r.i[0] = (size_t)xi0 + (size_t)t.w * yi0;
    r.i[1] = (size_t)xi0 + (size_t)t.w * yi1;
    r.i[2] = (size_t)xi1 + (size_t)t.w * yi0;
    r.i[3] = (size_t)xi1 + (size_t)t.w * yi1;
    r.w[0] = omdx; r.w[1] = omdx; r.w[2] = dx; r.w[3] = dx;      // x factor of each tap
    (void)omdy; (void)dy;
    return r;
}

int backward(sl_context* c, size_t fboId, size_t gradImageTex)
{
    if (fboId >= c->fbos.size()) return SL_ERR_INVALID_HANDLE;
    Fbo& f = c->fbos[fboId];
    if (f.vis >= c->texs.size() || !c->texs[f.vis].valid) return SL_ERR_STATE;
    const Tex& vt = c->texs[f.vis];
    if (gradImageTex >= c->texs.size() || !c->texs[gradImageTex].valid) return SL_ERR_INVALID_HANDLE;
    const Tex& gt = c->texs[gradImageTex];
    if (gt.type != SL_COLOR_RGBA_FLOAT || gt.w != vt.w || gt.h != vt.h) return SL_ERR_INVALID_ARG;
    if (f.savedStale) return SL_ERR_STATE;
    if (f.savedBlend) return SL_ERR_UNSUPPORTED;
    if (f.saved.empty()) return SL_OK;
    for (const SavedPass& sp : f.saved)
    {
        const Shader& sh = c->shaders[sp.shader];
        const bool ct = sh.vsProg == SL_PROGRAM_COLOR_TEXTURED && sh.fsProg == SL_PROGRAM_COLOR_TEXTURED;
        const bool pg = sh.vsProg == SL_PROGRAM_PASSTHROUGH_GOURAUD && sh.fsProg == SL_PROGRAM_PASSTHROUGH_GOURAUD;
        if (!ct && !pg) return SL_ERR_UNSUPPORTED;
    }
    // double accumulators, one per destination buffer
    std::vector<std::vector<double>> colAcc(c->vbos.size()), texAcc(c->texs.size());
    const float* vis = reinterpret_cast<const float*>(vt.data.data());
    const float* gimg = reinterpret_cast<const float*>(gt.data.data());
    const size_t npx = (size_t)vt.w * vt.h;
    for (size_t p = 0; p < npx; ++p)
    {
        uint32_t seq;
        std::memcpy(&seq, &vis[4 * p], 4);
        if (seq == 0) continue;
        // last pass whose first sequence number is below seq
        size_t lo = 0, hi = f.saved.size();
        while (hi - lo > 1)
        {
            const size_t mid = (lo + hi) / 2;
            if (f.saved[mid].seqBase < seq) lo = mid;
            else hi = mid;
        }
        const SavedPass& sp = f.saved[lo];
        if (seq <= sp.seqBase) continue;
        const size_t t = (size_t)(seq - 1u - sp.seqBase);
        if (t >= sp.numTris) continue;
        Draw d;
        d.ctx = c;
        d.sh = &c->shaders[sp.shader];
        if (bind_mesh(c, sp.mesh, d) != SL_OK) continue;
        const size_t e = sp.mesh.elementBegin + 3 * t;
        size_t vid[3] = {e, e + 1, e + 2};
        if (mesh_indexed(kPrimFill, sp.mesh.mode))
            for (int k = 0; k < 3; ++k) vid[k] = index_at(*d.ibo, e + (size_t)k);
        const float b[3] = {vis[4 * p + 1], vis[4 * p + 2], vis[4 * p + 3]};
        const float* g = &gimg[4 * p];
        const bool ct = d.sh->fsProg == SL_PROGRAM_COLOR_TEXTURED;

        float a[3][4], uv[3][4];
        for (int k = 0; k < 3; ++k)
        {
            read_attr(d, 1, vid[k], a[k], ct ? 0.f : 1.f);
            if (ct) a[k][3] = 1.f;                                 // varying0 = (rgb, 1)
            if (ct) read_attr(d, 2, vid[k], uv[k], 0.f);
        }
        float col[4];
        for (int ch = 0; ch < 4; ++ch) col[ch] = ((a[0][ch] * b[0]) + (a[1][ch] * b[1])) + (a[2][ch] * b[2]);
        float tex[4] = {1.f, 1.f, 1.f, 1.f};
        const Tex* T = nullptr;
        Taps tp{};
        float dy = 0.f;
        if (ct && sp.tex < c->texs.size() && c->texs[sp.tex].valid && c->texs[sp.tex].type == SL_COLOR_RGBA_FLOAT)
        {
            T = &c->texs[sp.tex];
            const float u = ((uv[0][0] * b[0]) + (uv[1][0] * b[1])) + (uv[2][0] * b[2]);
            const float v = ((uv[0][1] * b[0]) + (uv[1][1] * b[1])) + (uv[2][1] * b[2]);
            sl_std_uniforms su;
            std::memset(&su, 0, sizeof(su));
            if (sp.uni.size() >= sizeof(su)) std::memcpy(&su, sp.uni.data(), sizeof(su));
            if (su.flags & SL_UNIFORM_SAMPLE_NEAREST)
            {
                // one tap of weight 1 (SL_Sampler.hpp:122-148)
                sample_nearest_rgbaf(*T, u, v, tex);
                tp.i[0] = tp.i[1] = tp.i[2] = tp.i[3] = nearest_index(*T, u, v);
                tp.w[0] = 1.f; tp.w[1] = tp.w[2] = tp.w[3] = 0.f;
                dy = 0.f;
            }
            else
            {
                sample_bilinear_rgbaf(*T, u, v, tex);
                tp = bilinear_taps(*T, u, v);
                dy = wrap_repeat(v) * (float)T->h;
                dy = dy - (float)to_u16(dy);
            }
        }
        // colour attribute gradients
        const Vbo& V = c->vbos[c->vaos[sp.mesh.vaoId].vbo];
        if (V.grad < c->vbos.size() && c->vbos[V.grad].valid)
        {
            const size_t gv = V.grad;
            const size_t cnt = c->vbos[gv].bytes.size() / 16;
            std::vector<double>& acc = colAcc[gv];
            if (acc.empty()) acc.assign(cnt * 4, 0.0);
            for (int k = 0; k < 3; ++k)
            {
                if (vid[k] >= cnt) continue;
                for (int ch = 0; ch < (ct ? 3 : 4); ++ch)
                    acc[vid[k] * 4 + ch] += (double)b[k] * ((double)g[ch] * (double)tex[ch]);
            }
        }
        // texel gradients
        if (T && T->grad < c->texs.size() && c->texs[T->grad].valid)
        {
            const size_t gi = T->grad;
            const size_t cnt = c->texs[gi].data.size() / 16;
            std::vector<double>& acc = texAcc[gi];
            if (acc.empty()) acc.assign(cnt * 4, 0.0);
            const double omdy = 1.0 - (double)dy;
            const double wy[4] = {omdy, (double)dy, omdy, (double)dy};
            for (int j = 0; j < 4; ++j)
            {
                if (tp.i[j] >= cnt) continue;
                const double wj = (double)tp.w[j] * wy[j];
                for (int ch = 0; ch < 4; ++ch)
                    acc[tp.i[j] * 4 + ch] += wj * ((double)g[ch] * (double)col[ch]);
            }
        }
    }
    for (size_t i = 0; i < colAcc.size(); ++i)
    {
        if (colAcc[i].empty()) continue;
        float* dst = reinterpret_cast<float*>(c->vbos[i].bytes.data());
        for (size_t k = 0; k < colAcc[i].size(); ++k) dst[k] += (float)colAcc[i][k];
    }
    for (size_t i = 0; i < texAcc.size(); ++i)
    {
        if (texAcc[i].empty()) continue;
        float* dst = reinterpret_cast<float*>(c->texs[i].data.data());
        for (size_t k = 0; k < texAcc[i].size(); ++k) dst[k] += (float)texAcc[i][k];
    }
    c->stats.backward_launches += 1;
    return SL_OK;
}

} // namespace

/*-----------------------------------------------------------------------------
 * C API (prefix slo_), one function per entry point of include/sl_gpu.h
-----------------------------------------------------------------------------*/
extern "C" {

int slo_context_create(sl_context** out)
{
    if (!out) return SL_ERR_INVALID_ARG;
    *out = new sl_context();
    return SL_OK;
}

void slo_context_destroy(sl_context* c) { delete c; }

const char* slo_last_error(const sl_context* c) { return c ? c->err.c_str() : "null context"; }

unsigned slo_num_threads(const sl_context* c) { return c ? c->numThreads : 0u; }

unsigned slo_set_num_threads(sl_context* c, unsigned n)
{
    if (!c) return 0;
    c->numThreads = std::max(1u, n);
    return c->numThreads;
}

// The oracle has no tiles: any raster configuration index is accepted and
// ignored.
int slo_set_raster_config(sl_context* c, int config)
{
    if (!c) return SL_ERR_INVALID_ARG;
    return config >= 0 ? SL_OK : SL_ERR_INVALID_ARG;
}

// Heavy-tile split of the GPU raster: no effect on results, none here.
int slo_set_raster_split(sl_context* c, uint32_t partRecords)
{
    (void)partRecords;
    return c ? SL_OK : SL_ERR_INVALID_ARG;
}

// Raster occupancy of the GPU path: no effect on results, none here.
int slo_set_raster_occupancy(sl_context* c, uint32_t workgroupsPerCU)
{
    if (!c) return SL_ERR_INVALID_ARG;
    return workgroupsPerCU <= 32 ? SL_OK : SL_ERR_INVALID_ARG;
}

// GPU raster variant choice: nothing to do on the CPU (results never depend on it)
int slo_set_raster_depth_reject(sl_context* c, int enable)
{
    (void)enable;
    return c ? SL_OK : SL_ERR_INVALID_ARG;
}

int slo_get_raster_config(const sl_context* c, int config, int* tileW, int* tileH, int* segments)
{
    if (!c) return SL_ERR_INVALID_ARG;
    if (tileW) *tileW = 0;
    if (tileH) *tileH = 0;
    if (segments) *segments = 0;
    return config < 0 ? 0 : config;
}

int slo_set_viewport(sl_context* c, int32_t x, int32_t y, int32_t w, int32_t h)
{
    if (!c) return SL_ERR_INVALID_ARG;
    // SL_ViewportState::viewport (SL_ViewportState.hpp:148-156)
    c->viewport[0] = std::min(std::max(x, -65536), 65535);
    c->viewport[1] = std::min(std::max(y, -65536), 65535);
    c->viewport[2] = std::min(w, 65535 - x);
    c->viewport[3] = std::min(h, 65535 - y);
    return SL_OK;
}

int slo_set_scissor(sl_context* c, int32_t x, int32_t y, int32_t w, int32_t h)
{
    if (!c) return SL_ERR_INVALID_ARG;
    c->scissor[0] = std::min(std::max(x, -65536), 65535);
    c->scissor[1] = std::min(std::max(y, -65536), 65535);
    c->scissor[2] = std::min(w, 65535 - x);
    c->scissor[3] = std::min(h, 65535 - y);
    return SL_OK;
}

size_t slo_create_vbo(sl_context* c) { c->vbos.emplace_back(); return c->vbos.size() - 1; }

int slo_vbo_init(sl_context* c, size_t vbo, size_t numBytes, const void* data)
{
    if (vbo >= c->vbos.size()) return SL_ERR_INVALID_HANDLE;
    if (!numBytes) return -1;
    if (c->vbos[vbo].valid) stale_saved(c);
    Vbo& v = c->vbos[vbo];
    v.bytes.assign(numBytes, 0);
    if (data) std::memcpy(v.bytes.data(), data, numBytes);
    v.valid = true;
    return SL_OK;
}

int slo_vbo_assign(sl_context* c, size_t vbo, const void* data, size_t offset, size_t numBytes)
{
    if (vbo >= c->vbos.size()) return SL_ERR_INVALID_HANDLE;
    Vbo& v = c->vbos[vbo];
    if (offset + numBytes > v.bytes.size()) return SL_ERR_INVALID_ARG;
    std::memcpy(v.bytes.data() + offset, data, numBytes);
    return SL_OK;
}

int slo_vbo_read(sl_context* c, size_t vbo, void* out, size_t offset, size_t numBytes)
{
    if (vbo >= c->vbos.size() || !c->vbos[vbo].valid) return SL_ERR_INVALID_HANDLE;
    const Vbo& v = c->vbos[vbo];
    if (offset + numBytes > v.bytes.size() || !out) return SL_ERR_INVALID_ARG;
    std::memcpy(out, v.bytes.data() + offset, numBytes);
    return SL_OK;
}

size_t slo_create_ibo(sl_context* c) { c->ibos.emplace_back(); return c->ibos.size() - 1; }

int slo_ibo_init(sl_context* c, size_t ibo, uint32_t numElements, int type, const void* data)
{
    if (ibo >= c->ibos.size()) return SL_ERR_INVALID_HANDLE;
    if (!numElements) return -1;
    if (c->ibos[ibo].valid) stale_saved(c);
    const uint32_t bpe = bytes_per_index(type);
    if (!bpe) return SL_ERR_INVALID_ARG;
    Ibo& b = c->ibos[ibo];
    b.type = type;
    b.bpe = bpe;
    b.count = numElements;
    b.bytes.assign((size_t)numElements * bpe + 16, 0);   // padded like SL_IndexBuffer.cpp:155
    if (data) std::memcpy(b.bytes.data(), data, (size_t)numElements * bpe);
    b.valid = true;
    return SL_OK;
}

size_t slo_create_vao(sl_context* c) { c->vaos.emplace_back(); return c->vaos.size() - 1; }

int slo_vao_set_vertex_buffer(sl_context* c, size_t vao, size_t vbo)
{
    if (vao >= c->vaos.size()) return SL_ERR_INVALID_HANDLE;
    c->vaos[vao].vbo = vbo;
    return SL_OK;
}

int slo_vao_set_index_buffer(sl_context* c, size_t vao, size_t ibo)
{
    if (vao >= c->vaos.size()) return SL_ERR_INVALID_HANDLE;
    c->vaos[vao].ibo = ibo;
    return SL_OK;
}

int slo_vao_set_num_bindings(sl_context* c, size_t vao, size_t n)
{
    if (vao >= c->vaos.size()) return SL_ERR_INVALID_HANDLE;
    if (n >= SL_VAO_MAX_BINDINGS) return -1;                   // SL_VertexArray.cpp:164-167
    Vao& v = c->vaos[vao];
    for (size_t i = v.nb; i < n; ++i) v.b[i] = Binding{};
    v.nb = n;
    return (int)n;
}

int slo_vao_set_binding(sl_context* c, size_t vao, size_t bindId, ptrdiff_t offset, ptrdiff_t stride, int dims, int type)
{
    if (vao >= c->vaos.size()) return SL_ERR_INVALID_HANDLE;
    Vao& v = c->vaos[vao];
    if (bindId >= v.nb) return SL_ERR_INVALID_ARG;
    v.b[bindId].offset = offset;
    v.b[bindId].stride = stride;
    v.b[bindId].dims = dims;
    v.b[bindId].type = type;
    return SL_OK;
}

size_t slo_create_texture(sl_context* c) { c->texs.emplace_back(); return c->texs.size() - 1; }

int slo_texture_init(sl_context* c, size_t tex, int type, uint16_t w, uint16_t h, uint16_t d)
{
    if (tex >= c->texs.size()) return SL_ERR_INVALID_HANDLE;
    const uint32_t bpt = bytes_per_color(type);
    if (!bpt) return SL_ERR_UNSUPPORTED;
    if (!w || !h || !d) return SL_ERR_INVALID_ARG;
    if (c->texs[tex].valid) stale_saved(c);
    Tex& t = c->texs[tex];
    t.type = type; t.w = w; t.h = h; t.d = d; t.bpt = bpt;
    // SL_Texture.cpp:28-55: each dimension padded by (4 - dim % 4), zero filled
    const size_t pw = (size_t)w + (4 - w % 4), ph = (size_t)h + (4 - h % 4), pd = (size_t)d + (4 - d % 4);
    t.data.assign(pw * ph * pd * bpt + 4 * bpt, 0);
    t.valid = true;
    return SL_OK;
}

int slo_texture_set_texels(sl_context* c, size_t tex, const void* data, size_t numBytes)
{
    if (tex >= c->texs.size() || !c->texs[tex].valid) return SL_ERR_INVALID_HANDLE;
    Tex& t = c->texs[tex];
    const size_t n = (size_t)t.w * t.h * t.d * t.bpt;
    if (numBytes != n) return SL_ERR_INVALID_ARG;
    std::memcpy(t.data.data(), data, n);
    return SL_OK;
}

int slo_texture_read(sl_context* c, size_t tex, void* out, size_t numBytes)
{
    if (tex >= c->texs.size() || !c->texs[tex].valid) return SL_ERR_INVALID_HANDLE;
    Tex& t = c->texs[tex];
    const size_t n = (size_t)t.w * t.h * t.d * t.bpt;
    if (numBytes != n) return SL_ERR_INVALID_ARG;
    std::memcpy(out, t.data.data(), n);
    return SL_OK;
}

size_t slo_create_framebuffer(sl_context* c) { c->fbos.emplace_back(); return c->fbos.size() - 1; }

int slo_fbo_reserve_color_buffers(sl_context* c, size_t fbo, unsigned n)
{
    if (fbo >= c->fbos.size()) return SL_ERR_INVALID_HANDLE;
    if (n > SL_SHADER_MAX_FRAG_OUTPUTS) return SL_ERR_INVALID_ARG;
    c->fbos[fbo].colors.assign(n, SL_INVALID_ID);
    return SL_OK;
}

int slo_fbo_attach_color_buffer(sl_context* c, size_t fbo, unsigned idx, size_t tex)
{
    if (fbo >= c->fbos.size() || tex >= c->texs.size()) return SL_ERR_INVALID_HANDLE;
    Fbo& f = c->fbos[fbo];
    if (idx >= f.colors.size()) return SL_ERR_INVALID_ARG;
    f.colors[idx] = tex;
    return SL_OK;
}

int slo_fbo_attach_depth_buffer(sl_context* c, size_t fbo, size_t tex)
{
    if (fbo >= c->fbos.size() || tex >= c->texs.size()) return SL_ERR_INVALID_HANDLE;
    c->fbos[fbo].depth = tex;
    return SL_OK;
}

int slo_fbo_attach_debug_buffers(sl_context* c, size_t fbo, size_t lastTex, size_t countTex)
{
    if (fbo >= c->fbos.size()) return SL_ERR_INVALID_HANDLE;
    Fbo& f = c->fbos[fbo];
    for (size_t t : {lastTex, countTex})
        if (t != SL_INVALID_ID && (t >= c->texs.size() || c->texs[t].type != SL_COLOR_R_32U)) return SL_ERR_INVALID_ARG;
    f.dbgLast = lastTex;
    f.dbgCount = countTex;
    return SL_OK;
}

int slo_fbo_valid(sl_context* c, size_t fbo)
{
    if (fbo >= c->fbos.size()) return SL_ERR_INVALID_HANDLE;
    const Fbo& f = c->fbos[fbo];
    if (f.colors.empty()) return -1;                            // SL_Framebuffer.cpp:930-990
    for (size_t t : f.colors) if (t >= c->texs.size() || !c->texs[t].valid) return -2;
    const Tex& c0 = c->texs[f.colors[0]];
    for (size_t t : f.colors)
    {
        if (c->texs[t].w != c0.w) return -3;
        if (c->texs[t].h != c0.h) return -4;
        if (c->texs[t].d != c0.d) return -5;
    }
    if (f.depth >= c->texs.size() || !c->texs[f.depth].valid) return -6;
    const Tex& dt = c->texs[f.depth];
    if (dt.w != c0.w) return -7;
    if (dt.h != c0.h) return -8;
    if (dt.d > 1) return -9;
    if (dt.type != SL_COLOR_R_HALF && dt.type != SL_COLOR_R_FLOAT && dt.type != SL_COLOR_R_DOUBLE) return -10;
    return 0;
}

size_t slo_create_ubo(sl_context* c)
{
    c->ubos.emplace_back();
    std::memset(c->ubos.back().bytes, 0, SL_MAX_UNIFORM_BUFFER_SIZE);
    return c->ubos.size() - 1;
}

int slo_ubo_assign(sl_context* c, size_t ubo, const void* data, size_t offset, size_t n)
{
    if (ubo >= c->ubos.size()) return SL_ERR_INVALID_HANDLE;
    if (offset + n > SL_MAX_UNIFORM_BUFFER_SIZE) return SL_ERR_INVALID_ARG;
    std::memcpy(c->ubos[ubo].bytes + offset, data, n);
    return SL_OK;
}

size_t slo_create_shader(sl_context* c, const sl_vertex_shader* vs, const sl_fragment_shader* fs, size_t ubo)
{
    // SL_Context::create_shader (SL_Context.cpp:547-604)
    if (!vs || !fs) return SL_INVALID_ID;
    if (ubo != SL_INVALID_ID && ubo >= c->ubos.size()) return SL_INVALID_ID;
    if (vs->numVaryings < fs->numVaryings) return SL_INVALID_ID;
    if (vs->numVaryings > SL_SHADER_MAX_VARYING_VECTORS) return SL_INVALID_ID;
    if (fs->numVaryings > SL_SHADER_MAX_VARYING_VECTORS) return SL_INVALID_ID;
    if (fs->numOutputs > SL_SHADER_MAX_FRAG_OUTPUTS) return SL_INVALID_ID;
    Shader s;
    s.cull = vs->cullMode;
    s.dtest = fs->depthTest;
    s.dmask = fs->depthMask;
    s.blend = fs->blend;
    s.nvary = fs->numVaryings;          // the pipeline's varying count comes from the fragment shader
    s.nout = fs->numOutputs;
    s.vsProg = vs->programId;
    s.fsProg = fs->programId;
    s.ubo = ubo;
    c->shaders.push_back(s);
    return c->shaders.size() - 1;
}

int slo_draw(sl_context* c, const sl_mesh* m, size_t shader, size_t fbo)
{
    if (!c || !m) return SL_ERR_INVALID_ARG;
    return run_draws(c, m, 1, 1, shader, fbo);
}

int slo_draw_multiple(sl_context* c, const sl_mesh* m, size_t n, size_t shader, size_t fbo)
{
    if (!c) return SL_ERR_INVALID_ARG;
    if (m == nullptr || n == 0) return SL_OK;                     // SL_Context.cpp:726-730
    return run_draws(c, m, n, 1, shader, fbo);
}

int slo_draw_instanced(sl_context* c, const sl_mesh* m, size_t numInstances, size_t shader, size_t fbo)
{
    if (!c || !m) return SL_ERR_INVALID_ARG;
    if (numInstances == 0) return SL_OK;
    return run_draws(c, m, 1, numInstances, shader, fbo);
}

// sl_draw_static (no reference counterpart): the oracle draws synchronously
// into host memory, so it is sl_draw.
int slo_draw_static(sl_context* c, const sl_mesh* m, size_t shader, size_t fbo) { return slo_draw(c, m, shader, fbo); }

int slo_clear_color_buffer(sl_context* c, size_t fbo, unsigned att, const double rgba[4]) { return clear_color(c, fbo, att, rgba); }

int slo_clear_depth_buffer(sl_context* c, size_t fbo, double depth) { return clear_depth(c, fbo, depth); }

int slo_clear_framebuffer(sl_context* c, size_t fbo, unsigned att, const double rgba[4], double depth)
{
    int rc = clear_color(c, fbo, att, rgba);
    if (rc != SL_OK) return rc;
    return clear_depth(c, fbo, depth);
}

int slo_finish(sl_context*) { return SL_OK; }

// The oracle executes every draw immediately; there is nothing to capture.
int slo_frame_graph_capture(sl_context*) { return SL_OK; }
int slo_frame_graph_launch(sl_context*) { return SL_ERR_UNSUPPORTED; }
int slo_frame_graph_release(sl_context*) { return SL_OK; }

/* blit: SL_BlitProcessor::blit_nearest (SL_BlitProcessor.cpp:620-665) with the
 * colour operators of :289-416; pixels where the reference would read or
 * write outside a texture are skipped */
int slo_blit_region(sl_context* c, size_t outTex, size_t inTex, uint16_t srcX0, uint16_t srcY0, uint16_t srcX1,
                    uint16_t srcY1, uint16_t dstX0, uint16_t dstY0, uint16_t dstX1, uint16_t dstY1)
{
    if (!c) return SL_ERR_INVALID_ARG;
    if (outTex >= c->texs.size() || inTex >= c->texs.size() || !c->texs[outTex].valid || !c->texs[inTex].valid)
        return SL_ERR_INVALID_HANDLE;
    Tex& o = c->texs[outTex];
    const Tex& in = c->texs[inTex];
    if (in.type != SL_COLOR_RGBA_8U && in.type != SL_COLOR_RGBA_FLOAT && in.type != SL_COLOR_R_FLOAT) return SL_ERR_UNSUPPORTED;
    if (o.type != SL_COLOR_RGBA_8U && o.type != SL_COLOR_RGBA_FLOAT) return SL_ERR_UNSUPPORTED;
    if (outTex == inTex) return SL_ERR_INVALID_ARG;
    // uint_fast32_t is 64 bits wide on the reference's x86-64 targets
    const uint64_t inW = (uint64_t)srcX1 - (uint64_t)srcX0;
    const uint64_t inH = (uint64_t)srcY1 - (uint64_t)srcY0;
    const uint64_t outW = (uint64_t)dstX1 - (uint64_t)dstX0;
    const uint64_t totalOutW = o.w, totalOutH = o.h;
    const uint64_t x0 = dstX0;
    const uint64_t x1 = std::min<uint64_t>(totalOutW, x0 + outW);
    const uint64_t y1 = std::min<uint64_t>(dstY1, totalOutH);
    const uint64_t finW = inW << 16, finH = inH << 16;
    const uint64_t foutW = (finW / totalOutW) + 1u;        // "account for rounding errors"
    const uint64_t foutH = (finH / totalOutH) + 1u;
    for (uint64_t y = dstY0; y < y1; ++y)
    {
        const uint64_t yf = (y * foutH) >> 16;
        const uint64_t srcY = (uint64_t)srcY1 - ((uint64_t)srcY0 + yf) - 1u;
        for (uint64_t x = x0; x < x1; ++x)
        {
            const uint64_t srcX = (x * foutW) >> 16;
            if (srcX >= in.w || srcY >= in.h) continue;
            const size_t si = srcX + (size_t)in.w * srcY;
            const size_t di = x + (size_t)o.w * y;
            float col[4];
            if (in.type == SL_COLOR_RGBA_8U && o.type == SL_COLOR_RGBA_8U)
            {
                std::memcpy(o.data.data() + 4 * di, in.data.data() + 4 * si, 4);
                continue;
            }
            if (in.type == SL_COLOR_RGBA_8U)
                for (int k = 0; k < 4; ++k) col[k] = (float)in.data[4 * si + k] * 0.00392156862745f;   // color_cast<float, uint8_t>
            else if (in.type == SL_COLOR_RGBA_FLOAT)
                std::memcpy(col, in.data.data() + 16 * si, 16);
            else
            {
                // SL_Blit_R_to_RGBA: (min, min, r, max) of the source type (:303-306)
                col[0] = 0.f; col[1] = 0.f; std::memcpy(&col[2], in.data.data() + 4 * si, 4); col[3] = 1.f;
            }
            if (o.type == SL_COLOR_RGBA_8U)
                for (int k = 0; k < 4; ++k) o.data[4 * di + k] = (uint8_t)(uint32_t)f2i(col[k] * 255.f);   // color_cast<uint8_t, float>
            else
                std::memcpy(o.data.data() + 16 * di, col, 16);
        }
    }
    return SL_OK;
}

int slo_blit(sl_context* c, size_t outTex, size_t inTex)
{
    if (!c) return SL_ERR_INVALID_ARG;
    if (outTex >= c->texs.size() || inTex >= c->texs.size() || !c->texs[outTex].valid || !c->texs[inTex].valid)
        return SL_ERR_INVALID_HANDLE;
    const Tex& o = c->texs[outTex];
    const Tex& in = c->texs[inTex];
    return slo_blit_region(c, outTex, inTex, 0, 0, in.w, in.h, 0, 0, o.w, o.h);   // SL_Context.cpp:747-769
}

/* differentiable draws (no reference counterpart) */
int slo_framebuffer_attach_visibility(sl_context* c, size_t fbo, size_t tex)
{
    if (!c || fbo >= c->fbos.size()) return SL_ERR_INVALID_HANDLE;
    if (tex != SL_INVALID_ID)
    {
        if (tex >= c->texs.size() || !c->texs[tex].valid) return SL_ERR_INVALID_HANDLE;
        if (c->texs[tex].type != SL_COLOR_RGBA_FLOAT) return SL_ERR_INVALID_ARG;
    }
    Fbo& f = c->fbos[fbo];
    f.vis = tex;
    f.saved.clear();
    f.savedDraws = 0;
    f.savedStale = false;
    f.savedBlend = false;
    return SL_OK;
}

int slo_vbo_attach_gradient(sl_context* c, size_t vbo, size_t grad)
{
    if (!c || vbo >= c->vbos.size()) return SL_ERR_INVALID_HANDLE;
    if (grad != SL_INVALID_ID && (grad >= c->vbos.size() || grad == vbo)) return SL_ERR_INVALID_HANDLE;
    c->vbos[vbo].grad = grad;
    return SL_OK;
}

int slo_texture_attach_gradient(sl_context* c, size_t tex, size_t grad)
{
    if (!c || tex >= c->texs.size()) return SL_ERR_INVALID_HANDLE;
    if (grad != SL_INVALID_ID)
    {
        if (grad >= c->texs.size() || grad == tex || !c->texs[grad].valid || !c->texs[tex].valid) return SL_ERR_INVALID_HANDLE;
        const Tex& g = c->texs[grad];
        const Tex& t = c->texs[tex];
        if (g.type != SL_COLOR_RGBA_FLOAT || g.w != t.w || g.h != t.h || g.d != t.d) return SL_ERR_INVALID_ARG;
    }
    c->texs[tex].grad = grad;
    return SL_OK;
}

int slo_gradients_zero(sl_context* c)
{
    if (!c) return SL_ERR_INVALID_ARG;
    for (const Vbo& v : c->vbos)
        if (v.valid && v.grad < c->vbos.size() && c->vbos[v.grad].valid)
            std::fill(c->vbos[v.grad].bytes.begin(), c->vbos[v.grad].bytes.end(), 0);
    for (const Tex& t : c->texs)
        if (t.valid && t.grad < c->texs.size() && c->texs[t.grad].valid)
            std::fill(c->texs[t.grad].data.begin(), c->texs[t.grad].data.end(), 0);
    for (const Fbo& f : c->fbos)
        if (f.posGrad < c->vbos.size() && c->vbos[f.posGrad].valid)
            std::fill(c->vbos[f.posGrad].bytes.begin(), c->vbos[f.posGrad].bytes.end(), 0);
    return SL_OK;
}

int slo_loss_l2_grad(sl_context* c, size_t colorTex, size_t targetTex, size_t gradTex, float scale)
{
    if (!c) return SL_ERR_INVALID_ARG;
    for (size_t id : {colorTex, targetTex, gradTex})
        if (id >= c->texs.size() || !c->texs[id].valid) return SL_ERR_INVALID_HANDLE;
    const Tex& a = c->texs[colorTex];
    const Tex& b = c->texs[targetTex];
    Tex& g = c->texs[gradTex];
    for (const Tex* t : {&a, &b, (const Tex*)&g})
        if (t->type != SL_COLOR_RGBA_FLOAT || t->w != a.w || t->h != a.h || t->d != a.d) return SL_ERR_INVALID_ARG;
    const size_t n = (size_t)a.w * a.h * std::max<uint16_t>(a.d, 1) * 4;
    const float* pa = reinterpret_cast<const float*>(a.data.data());
    const float* pb = reinterpret_cast<const float*>(b.data.data());
    float* pg = reinterpret_cast<float*>(g.data.data());
    const float s2 = scale * 2.f;
    for (size_t i = 0; i < n; ++i) pg[i] = s2 * (pa[i] - pb[i]);
    return SL_OK;
}

} // extern "C"

/*-----------------------------------------------------------------------------
 * Smoothed coverage (no reference counterpart; the differentiable-coverage
 * mode).  Pixel (x, y) samples the point (x + 1, y): the span rule covers it
 * when e_min < x + 1 <= e_max at row y (SL_ScanlineBounds.hpp:310-322).
 * Where the visibility winners of two neighbouring pixels differ and an edge
 * of one of the winners separates their samples, the colour before the edge
 * leaks into the pixel past it with the weight
 *     w(d) = 2 s(-d / sigma) = 1 - d / sqrt(d^2 + sigma^2),
 * s(u) = (1 + u / sqrt(1 + u^2)) / 2 the algebraic sigmoid and d > 0 the
 * perpendicular distance of the receiving pixel's sample from the edge's
 * line.  w is 1 on the edge, so the image stays continuous where a pixel's
 * winner changes, and goes to 0 for every d > 0 as sigma -> 0; its tail
 * falls as (sigma / d)^2 / 2, so the weight and its derivative stay well
 * inside float range for sigma far below the pixel spacing.  Two separable
 * passes over the hard colour image C:
 *   X(x, y) = (1 - wx) C(x, y) + wx C(x - 1, y): the edge between the samples
 *     of x - 1 and x is a row-span end in [x, x + 1) -- the right end of the
 *     left pixel's winner if it lies there, else the left end of the right
 *     pixel's;
 *   O(x, y) = X(x, y) + sum wy (X(x, y') - X(x, y)) over the column edges
 *     that leak the neighbour row y' into the pixel: the edge between the
 *     samples of rows y and y + 1 in column x + 1 is a column-span end v in
 *     [y, y + 1] -- the high end of row y's winner if it lies there, else the
 *     low end of row y + 1's.  Row y receives (distance v - y) unless its
 *     sample lies on the edge (v = y, a tie the span rule gave to row y);
 *     then row y + 1 receives (distance 1).  The row pass needs no such
 *     rule: the span rule puts every tie on the left pixel of a pair.
 * A clipped winner is the polygon the clipper leaves: its row spans are the
 * union of the fan pieces the raster drew, its column spans come from its
 * edges; a vertex the clipper made has no offset and takes no position
 * gradient (it has no screen position of its own), the edge's other end
 * does.  sigma = 0 returns C.  The adjoint holds the winners fixed and sums
 * in double.
 *---------------------------------------------------------------------------*/
struct OShape
{
    bool ok = false;
    bool clipped = false;
    size_t k = 0;             // draw call of the framebuffer's saved batch
    int n = 0;                // vertices: the triangle's 3 or the clipped polygon's
    float sx[9], sy[9];       // snapped screen positions plus the offsets
    float rx[9], ry[9];       // snapped screen positions as rasterised
    size_t vid[9];
    bool orig[9];             // a vertex of the triangle, not one the clipper made
};

OShape winner_shape(sl_context* c, const Fbo& f, uint32_t seq, const float* offsets)
{
    OShape S;
    if (seq == 0 || f.saved.empty()) return S;
    size_t lo = 0, hi = f.saved.size();
    while (hi - lo > 1)
    {
        const size_t mid = (lo + hi) / 2;
        if (f.saved[mid].seqBase < seq) lo = mid;
        else hi = mid;
    }
    const SavedPass& sp = f.saved[lo];
    if (seq <= sp.seqBase) return S;
    const size_t t = (size_t)(seq - 1u - sp.seqBase);
    if (t >= sp.numTris) return S;
    Draw d;
    d.ctx = c;
    d.sh = &c->shaders[sp.shader];
    d.uni = sp.uni.empty() ? nullptr : sp.uni.data();
    std::memcpy(d.scissorMat, sp.scissorMat, sizeof(d.scissorMat));
    std::memcpy(d.vpDims, sp.vpDims, sizeof(d.vpDims));
    if (bind_mesh(c, sp.mesh, d) != SL_OK) return S;
    const size_t e = sp.mesh.elementBegin + 3 * t;
    size_t vid[3] = {e, e + 1, e + 2};
    if (mesh_indexed(kPrimFill, sp.mesh.mode))
        for (int k = 0; k < 3; ++k) vid[k] = index_at(*d.ibo, e + (size_t)k);
    V4 pos[3];
    for (int k = 0; k < 3; ++k)
    {
        V4 var[SL_SHADER_MAX_VARYING_VECTORS];
        pos[k] = mat_mul(d.scissorMat, run_vertex_program(d, vid[k], sp.instance, var));
    }
    const int status = clip_status(pos[0], pos[1], pos[2]);
    V4 poly[9];
    for (int k = 0; k < 3; ++k) poly[k] = pos[k];
    int n = 3;
    if (status == 1)
    {
        V4 var[9][SL_SHADER_MAX_VARYING_VECTORS];
        n = (int)clip_polygon(poly, var, 0);
        if (n < 3) return S;
        S.clipped = true;
    }
    else if (status != 3)
    {
        return S;
    }
    S.k = sp.drawIdx;
    S.n = n;
    for (int k = 0; k < n; ++k)
    {
        // the clipper copies the vertices it keeps unchanged
        int from = -1;
        for (int j = 0; j < 3 && from < 0; ++j)
            if (std::memcmp(&poly[k], &pos[j], sizeof(V4)) == 0) from = j;
        V4 p = poly[k];
        divide_and_snap(p, d.vpDims);
        S.rx[k] = p.v[0];
        S.ry[k] = p.v[1];
        S.orig[k] = from >= 0;
        S.vid[k] = from >= 0 ? vid[from] : 0;
        float ox = 0.f, oy = 0.f;
        if (offsets && from >= 0 && vid[from] < f.posVerts)
        {
            ox = offsets[2 * (S.k * f.posVerts + vid[from])];
            oy = offsets[2 * (S.k * f.posVerts + vid[from]) + 1];
        }
        S.sx[k] = p.v[0] + ox;
        S.sy[k] = p.v[1] + oy;
    }
    S.ok = true;
    return S;
}

// A span end and the edge (two vertex slots of the shape) it lies on.
struct OEnd { float e = 0.f; int p = 0, q = 0; };

// Which edge bounds a winner is decided on the positions as rasterised (the
// same float arithmetic as the raster, so the choice agrees with the hard
// winners, ties included); only the weights see the offsets.
//
// One triangle's span ends at row y with SL_ScanlineBounds::init's vertex
// order and :310-322's arithmetic: the left and the right end.
void tri_row_ends(const float* X, const float* Y, const int id[3], float yf, OEnd& left, OEnd& right)
{
    int a = id[0], b = id[1], cc = id[2];
    if (Y[a] < Y[b]) std::swap(a, b);
    if (Y[a] < Y[cc]) std::swap(a, cc);
    if (Y[b] < Y[cc]) std::swap(b, cc);
    const float ax = X[a], ay = Y[a], bx = X[b], by = Y[b], cx = X[cc], cy = Y[cc];
    const float d0 = yf - ay, d1 = yf - by;
    const float alpha = d0 * rcpf(cy - ay);
    OEnd lo{((cx - ax) * alpha) + ax, a, cc}, hi;
    if (d1 < 0.f) hi = OEnd{(((cx - bx) * rcpf(cy - by)) * d1) + bx, b, cc};
    else hi = OEnd{(((bx - ax) * rcpf(by - ay)) * d0) + ax, a, b};
    const bool loIsMax = lo.e >= hi.e;
    right = loIsMax ? lo : hi;
    left = loIsMax ? hi : lo;
}

// The shape's row span at y; a clipped polygon's is the union of the fan
// pieces (0, j + 1, j + 2) the raster drew whose rows hold y.
bool row_ends(const OShape& S, int y, OEnd& left, OEnd& right)
{
    const float yf = (float)y;
    if (!S.clipped)
    {
        const int id[3] = {0, 1, 2};
        tri_row_ends(S.rx, S.ry, id, yf, left, right);
        return true;
    }
    bool any = false;
    for (int j = S.n - 3; j >= 0; --j)
    {
        const int id[3] = {0, j + 1, j + 2};
        const float minX = std::min(std::min(S.rx[id[0]], S.rx[id[1]]), S.rx[id[2]]);
        const float maxX = std::max(std::max(S.rx[id[0]], S.rx[id[1]]), S.rx[id[2]]);
        const float minY = std::min(std::min(S.ry[id[0]], S.ry[id[1]]), S.ry[id[2]]);
        const float maxY = std::max(std::max(S.ry[id[0]], S.ry[id[1]]), S.ry[id[2]]);
        if ((maxX - minX < 1.f) || (maxY - minY < 1.f)) continue;      // push_bin's reject
        if (y < (int)minY || y >= (int)maxY) continue;
        OEnd l, r;
        tri_row_ends(S.rx, S.ry, id, yf, l, r);
        if (!any || l.e < left.e) left = l;
        if (!any || r.e > right.e) right = r;
        any = true;
    }
    return any;
}

// The crossing of edge (p, q) with the line b = c, along a:
// Pa + ((Qa - Pa) (c - Pb)) / (Qb - Pb), exact for vertices on the pixel
// lattice whenever the crossing is.
float edge_crossing(float pa, float pb, float qa, float qb, float c)
{
    return pa + ((qa - pa) * (c - pb)) / (qb - pb);
}

// The shape's column span at u: the lowest and highest crossing of its edges
// with the vertical line x = u.
bool col_ends(const OShape& S, float u, OEnd& low, OEnd& high)
{
    bool any = false;
    for (int i = 0; i < S.n; ++i)
    {
        const int p = i, q = (i + 1 == S.n) ? 0 : i + 1;
        const float px = S.rx[p], qx = S.rx[q];
        if (px == qx || u < std::min(px, qx) || u > std::max(px, qx)) continue;
        const float v = edge_crossing(S.ry[p], px, S.ry[q], qx, u);
        if (!any || v < low.e) low = OEnd{v, p, q};
        if (!any || v > high.e) high = OEnd{v, p, q};
        any = true;
    }
    return any;
}

// A smoothed boundary: the edge (P, Q) with the offsets, in pass coordinates
// -- a along the pass (x for the row pass, y for the column pass), b across
// it -- and the receiving sample (a = sample, b = cross).  The distance along
// a is dir (sample - crossing): positive on the rasterised positions, and a
// smooth function of the offsets (negative values extrapolate the weight).
struct OBnd
{
    bool ok = false;
    bool gp = false, gq = false;   // the end is a vertex of the triangle (has an offset and a gradient)
    bool recvLow = false;     // column pass: row y receives (else row y + 1)
    float sample = 0.f, cross = 0.f, dir = 1.f;
    float pa = 0.f, pb = 0.f, qa = 0.f, qb = 0.f;
    size_t k = 0, vp = 0, vq = 0;
};

OBnd make_bnd(const OShape& S, const OEnd& E, bool rowPass, float sample, float cross, float dir)
{
    OBnd B;
    B.ok = true;
    B.gp = S.orig[E.p];
    B.gq = S.orig[E.q];
    B.sample = sample;
    B.cross = cross;
    B.dir = dir;
    B.pa = rowPass ? S.sx[E.p] : S.sy[E.p];
    B.pb = rowPass ? S.sy[E.p] : S.sx[E.p];
    B.qa = rowPass ? S.sx[E.q] : S.sy[E.q];
    B.qb = rowPass ? S.sy[E.q] : S.sx[E.q];
    B.k = S.k;
    B.vp = S.vid[E.p];
    B.vq = S.vid[E.q];
    return B;
}

uint32_t vis_seq(const float* vis, int W, int x, int y)
{
    uint32_t s;
    std::memcpy(&s, &vis[4 * ((size_t)y * W + x)], 4);
    return s;
}

// The row-pass boundary between pixels x - 1 and x of row y: pixel x receives.
OBnd boundary_x(sl_context* c, const Fbo& f, const float* vis, int W, int x, int y, const float* offsets)
{
    if (x <= 0 || x >= W) return OBnd{};
    const uint32_t sl = vis_seq(vis, W, x - 1, y), sr = vis_seq(vis, W, x, y);
    if (sl == sr) return OBnd{};
    const float lo = (float)x, hi = (float)(x + 1), yf = (float)y;
    OEnd l, r;
    if (sl != 0)
    {
        const OShape L = winner_shape(c, f, sl, offsets);
        if (L.ok && row_ends(L, y, l, r) && r.e >= lo && r.e < hi) return make_bnd(L, r, true, hi, yf, 1.f);
    }
    if (sr != 0)
    {
        const OShape R = winner_shape(c, f, sr, offsets);
        if (R.ok && row_ends(R, y, l, r) && l.e >= lo && l.e < hi) return make_bnd(R, l, true, hi, yf, 1.f);
    }
    return OBnd{};
}

// The column-pass boundary between rows y and y + 1 of column x (sample
// x + 1): row y receives unless its sample lies on the edge.
OBnd col_bnd(const OShape& S, const OEnd& E, int y, int x)
{
    const float yl = (float)y, u = (float)(x + 1);
    const bool low = E.e > yl;
    OBnd B = make_bnd(S, E, false, low ? yl : yl + 1.f, u, low ? -1.f : 1.f);
    B.recvLow = low;
    return B;
}

OBnd boundary_y(sl_context* c, const Fbo& f, const float* vis, int W, int H, int x, int y, const float* offsets)
{
    if (y < 0 || y + 1 >= H) return OBnd{};
    const uint32_t sa = vis_seq(vis, W, x, y), sb = vis_seq(vis, W, x, y + 1);
    if (sa == sb) return OBnd{};
    const float lo = (float)y, hi = (float)(y + 1), u = (float)(x + 1);
    OEnd l, h;
    if (sa != 0)
    {
        const OShape A = winner_shape(c, f, sa, offsets);
        if (A.ok && col_ends(A, u, l, h) && h.e >= lo && h.e <= hi) return col_bnd(A, h, y, x);
    }
    if (sb != 0)
    {
        const OShape B = winner_shape(c, f, sb, offsets);
        if (B.ok && col_ends(B, u, l, h) && l.e >= lo && l.e <= hi) return col_bnd(B, l, y, x);
    }
    return OBnd{};
}

// Signed perpendicular distance of the receiving sample from the edge's line:
// the distance along a times |sin| of the edge against a.
float bnd_distance(const OBnd& B)
{
    // sample - crossing, measured from P so that no large coordinates cancel
    const float daxis = B.dir * ((B.sample - B.pa) - ((B.qa - B.pa) * (B.cross - B.pb)) / (B.qb - B.pb));
    const float ea = B.qa - B.pa, eb = B.qb - B.pb;
    const float len = std::sqrt(std::fma(ea, ea, eb * eb));
    return daxis * (std::fabs(eb) / len);
}

// w(d) = 1 - d / r, r = sqrt(d^2 + sigma^2); for d >= 0 written as
// sigma^2 / (r (r + d)) so that it keeps its precision where it is small.
float smooth_weight(float d, float sigma)
{
    const float r = std::sqrt(std::fma(d, d, sigma * sigma));
    return d >= 0.f ? (sigma * sigma) / (r * (r + d)) : 1.f - d / r;
}

float bnd_weight(const OBnd& B, float sigma) { return B.ok ? smooth_weight(bnd_distance(B), sigma) : 0.f; }

// In double: the signed distance and its partials d/d(Pa, Pb, Qa, Qb), with
// the crossing ea = Pa + (Qa - Pa) t, t = (c - Pb) / (Qb - Pb), and
// d = dir (sample - ea) |Qb - Pb| / |Q - P|.
double bnd_partials(const OBnd& B, double out[4])
{
    const double pa = B.pa, pb = B.pb, qa = B.qa, qb = B.qb, dir = B.dir, c = B.cross;
    const double inv = 1.0 / (qb - pb);
    const double t = (c - pb) * inv, s = (qa - pa) * inv;
    const double da = dir * (((double)B.sample - pa) - (qa - pa) * t);
    const double ePa = dir * (1.0 - t), eQa = dir * t, ePb = -dir * (1.0 - t) * s, eQb = -dir * t * s;
    const double ea = qa - pa, eb = qb - pb;
    const double len2 = ea * ea + eb * eb, len = std::sqrt(len2), len3 = len2 * len;
    const double g = std::fabs(eb) / len, sg = eb > 0.0 ? 1.0 : -1.0;
    const double gQa = -std::fabs(eb) * ea / len3, gPa = -gQa;
    const double gQb = sg * ea * ea / len3, gPb = -gQb;
    out[0] = -ePa * g + da * gPa;
    out[1] = -ePb * g + da * gPb;
    out[2] = -eQa * g + da * gQa;
    out[3] = -eQb * g + da * gQb;
    return da * g;
}

// checks shared by both smoothing entry points; fills the images' data
int smooth_check(sl_context* c, size_t fbo, size_t inTex, size_t outTex, float sigma, Fbo*& F, const Tex*& col,
                 const Tex*& vis, const Tex*& in, Tex*& out, const float*& offsets, bool backward)
{
    if (!c || fbo >= c->fbos.size()) return SL_ERR_INVALID_HANDLE;
    if (!(sigma >= 0.f) || !std::isfinite(sigma)) return SL_ERR_INVALID_ARG;
    F = &c->fbos[fbo];
    if (F->vis >= c->texs.size() || !c->texs[F->vis].valid) return SL_ERR_STATE;
    if (F->colors.empty() || F->colors[0] >= c->texs.size() || c->texs[F->colors[0]].type != SL_COLOR_RGBA_FLOAT) return SL_ERR_STATE;
    if (inTex >= c->texs.size() || !c->texs[inTex].valid || outTex >= c->texs.size() || !c->texs[outTex].valid)
        return SL_ERR_INVALID_HANDLE;
    vis = &c->texs[F->vis];
    col = &c->texs[F->colors[0]];
    in = &c->texs[inTex];
    out = &c->texs[outTex];
    for (const Tex* t : {in, (const Tex*)out})
        if (t->type != SL_COLOR_RGBA_FLOAT || t->w != vis->w || t->h != vis->h) return SL_ERR_INVALID_ARG;
    if (inTex == outTex) return SL_ERR_INVALID_ARG;
    if (F->savedStale) return SL_ERR_STATE;
    const uint64_t need = (uint64_t)F->savedDraws * F->posVerts * 8u;
    offsets = nullptr;
    if (F->posOffset != SL_INVALID_ID)
    {
        const Vbo& v = c->vbos[F->posOffset];
        if (!v.valid || v.bytes.size() < need) return SL_ERR_INVALID_ARG;
        offsets = reinterpret_cast<const float*>(v.bytes.data());
    }
    if (backward && F->posGrad != SL_INVALID_ID)
    {
        const Vbo& v = c->vbos[F->posGrad];
        if (!v.valid || v.bytes.size() < need) return SL_ERR_INVALID_ARG;
    }
    return SL_OK;
}

extern "C" {

int slo_framebuffer_attach_positions(sl_context* c, size_t fbo, size_t gradVbo, size_t offsetVbo, uint32_t numVertices)
{
    if (!c || fbo >= c->fbos.size()) return SL_ERR_INVALID_HANDLE;
    if (gradVbo != SL_INVALID_ID && (gradVbo >= c->vbos.size() || !c->vbos[gradVbo].valid)) return SL_ERR_INVALID_HANDLE;
    if (offsetVbo != SL_INVALID_ID && (offsetVbo >= c->vbos.size() || !c->vbos[offsetVbo].valid)) return SL_ERR_INVALID_HANDLE;
    if (gradVbo != SL_INVALID_ID && gradVbo == offsetVbo) return SL_ERR_INVALID_ARG;
    Fbo& f = c->fbos[fbo];
    f.posGrad = gradVbo;
    f.posOffset = offsetVbo;
    f.posVerts = numVertices;
    return SL_OK;
}

int slo_smooth_coverage(sl_context* c, size_t fbo, size_t outTex, float sigma)
{
    if (!c || fbo >= c->fbos.size()) return SL_ERR_INVALID_HANDLE;
    const size_t colId = c->fbos[fbo].colors.empty() ? SL_INVALID_ID : c->fbos[fbo].colors[0];
    Fbo* F; const Tex *col, *vis, *in; Tex* out; const float* offsets;
    int rc = smooth_check(c, fbo, colId, outTex, sigma, F, col, vis, in, out, offsets, false);
    if (rc != SL_OK) return rc;
    const int W = vis->w, H = vis->h;
    const size_t n = (size_t)W * H;
    const float* C = reinterpret_cast<const float*>(col->data.data());
    const float* V = reinterpret_cast<const float*>(vis->data.data());
    float* O = reinterpret_cast<float*>(out->data.data());
    if (!(sigma > 0.f) || F->saved.empty())
    {
        std::memcpy(O, C, n * 16);
        return SL_OK;
    }
    // row pass into X, then the column pass from X
    std::vector<float> X(n * 4);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x)
        {
            const size_t i = (size_t)y * W + x;
            const OBnd B = boundary_x(c, *F, V, W, x, y, offsets);
            for (int ch = 0; ch < 4; ++ch) X[4 * i + ch] = C[4 * i + ch];
            if (B.ok)
            {
                const float w = bnd_weight(B, sigma), u = 1.f - w;
                for (int ch = 0; ch < 4; ++ch) X[4 * i + ch] = (u * C[4 * i + ch]) + (w * C[4 * (i - 1) + ch]);
            }
        }
    for (size_t i = 0; i < n * 4; ++i) O[i] = X[i];
    for (int y = 0; y + 1 < H; ++y)
        for (int x = 0; x < W; ++x)
        {
            const OBnd B = boundary_y(c, *F, V, W, H, x, y, offsets);
            if (!B.ok) continue;
            const float w = bnd_weight(B, sigma);
            const size_t r = (size_t)(B.recvLow ? y : y + 1) * W + x, s = (size_t)(B.recvLow ? y + 1 : y) * W + x;
            for (int ch = 0; ch < 4; ++ch) O[4 * r + ch] += w * (X[4 * s + ch] - X[4 * r + ch]);
        }
    return SL_OK;
}

int slo_smooth_coverage_backward(sl_context* c, size_t fbo, size_t gradSmoothTex, float sigma, size_t gradImageTex)
{
    Fbo* F; const Tex *col, *vis, *in; Tex* out; const float* offsets;
    int rc = smooth_check(c, fbo, gradSmoothTex, gradImageTex, sigma, F, col, vis, in, out, offsets, true);
    if (rc != SL_OK) return rc;
    const int W = vis->w, H = vis->h;
    const size_t n = (size_t)W * H;
    const float* C = reinterpret_cast<const float*>(col->data.data());
    const float* V = reinterpret_cast<const float*>(vis->data.data());
    const float* G = reinterpret_cast<const float*>(in->data.data());
    float* O = reinterpret_cast<float*>(out->data.data());
    if (!(sigma > 0.f) || F->saved.empty())
    {
        std::memcpy(O, G, n * 16);
        return SL_OK;
    }
    const double s2 = (double)sigma * (double)sigma;
    std::vector<OBnd> bx(n), by(n);
    std::vector<double> wx(n, 0.0), wy(n, 0.0), dx(n, 0.0), dy(n, 0.0);
    auto dist = [](const OBnd& B) { double p[4]; return bnd_partials(B, p); };
    auto weight = [&](double d) {
        const double r = std::sqrt(d * d + s2);
        return d >= 0.0 ? s2 / (r * (r + d)) : 1.0 - d / r;
    };
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x)
        {
            const size_t i = (size_t)y * W + x;
            bx[i] = boundary_x(c, *F, V, W, x, y, offsets);
            by[i] = boundary_y(c, *F, V, W, H, x, y, offsets);
            if (bx[i].ok) { dx[i] = dist(bx[i]); wx[i] = weight(dx[i]); }
            if (by[i].ok) { dy[i] = dist(by[i]); wy[i] = weight(dy[i]); }
        }
    // X = row pass of C; GX = dL/dX = column pass transposed; dL/dC = row pass transposed
    std::vector<double> X(n * 4), GX(n * 4);
    for (size_t i = 0; i < n; ++i)
        for (int ch = 0; ch < 4; ++ch)
            X[4 * i + ch] = (1.0 - wx[i]) * C[4 * i + ch] + (bx[i].ok ? wx[i] * C[4 * (i - 1) + ch] : 0.0);
    // by[i] is the pair (row y, row y + 1) of column x: it moves w of the
    // source row's X into the receiving row
    for (size_t i = 0; i < n * 4; ++i) GX[i] = G[i];
    for (int y = 0; y + 1 < H; ++y)
        for (int x = 0; x < W; ++x)
        {
            const size_t i = (size_t)y * W + x;
            if (!by[i].ok) continue;
            const size_t r = by[i].recvLow ? i : i + W, s = by[i].recvLow ? i + W : i;
            for (int ch = 0; ch < 4; ++ch)
            {
                GX[4 * r + ch] -= wy[i] * G[4 * r + ch];
                GX[4 * s + ch] += wy[i] * G[4 * r + ch];
            }
        }
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x)
        {
            const size_t i = (size_t)y * W + x;
            for (int ch = 0; ch < 4; ++ch)
            {
                double g = (1.0 - wx[i]) * GX[4 * i + ch];
                if (x + 1 < W && bx[i + 1].ok) g += wx[i + 1] * GX[4 * (i + 1) + ch];
                O[4 * i + ch] = (float)g;
            }
        }
    if (F->posGrad == SL_INVALID_ID) return SL_OK;
    std::vector<double> pg((size_t)F->savedDraws * F->posVerts * 2, 0.0);
    auto scatter = [&](const OBnd& B, double dLdd, bool rowPass) {
        if (dLdd == 0.0) return;
        double p[4];
        bnd_partials(B, p);
        const int ax = rowPass ? 0 : 1, bxo = rowPass ? 1 : 0;   // a and b as x / y
        if (B.gp && B.vp < F->posVerts)
        {
            const size_t o = 2 * (B.k * F->posVerts + B.vp);
            pg[o + ax] += dLdd * p[0];
            pg[o + bxo] += dLdd * p[1];
        }
        if (B.gq && B.vq < F->posVerts)
        {
            const size_t o = 2 * (B.k * F->posVerts + B.vq);
            pg[o + ax] += dLdd * p[2];
            pg[o + bxo] += dLdd * p[3];
        }
    };
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x)
        {
            const size_t i = (size_t)y * W + x;
            if (bx[i].ok)
            {
                double dLdw = 0.0;
                for (int ch = 0; ch < 4; ++ch) dLdw += GX[4 * i + ch] * ((double)C[4 * (i - 1) + ch] - (double)C[4 * i + ch]);
                const double r = std::sqrt(dx[i] * dx[i] + s2);
                scatter(bx[i], dLdw * (-s2 / (r * r * r)), true);
            }
            if (by[i].ok)
            {
                const size_t rr = by[i].recvLow ? i : i + W, ss = by[i].recvLow ? i + W : i;
                double dLdw = 0.0;
                for (int ch = 0; ch < 4; ++ch) dLdw += (double)G[4 * rr + ch] * (X[4 * ss + ch] - X[4 * rr + ch]);
                const double r = std::sqrt(dy[i] * dy[i] + s2);
                scatter(by[i], dLdw * (-s2 / (r * r * r)), false);
            }
        }
    float* dst = reinterpret_cast<float*>(c->vbos[F->posGrad].bytes.data());
    for (size_t k = 0; k < pg.size(); ++k) dst[k] += (float)pg[k];
    return SL_OK;
}

// The smoothed image's inner product with gradTex, sum_i G_i . O_i, with
// the model evaluated in double (weights, both passes, the sum): the loss
// whose central differences check the position gradients at sigma far below
// the pixel spacing, where a float forward would drown them in rounding.
// Oracle only (test infrastructure).
int slo_smooth_loss_f64(sl_context* c, size_t fbo, size_t gradTex, float sigma, double* loss)
{
    if (!loss) return SL_ERR_INVALID_ARG;
    if (!c || fbo >= c->fbos.size()) return SL_ERR_INVALID_HANDLE;
    const size_t colId = c->fbos[fbo].colors.empty() ? SL_INVALID_ID : c->fbos[fbo].colors[0];
    if (gradTex == colId) return SL_ERR_INVALID_ARG;
    Fbo* F; const Tex *col, *vis, *in; Tex* out; const float* offsets;
    int rc = smooth_check(c, fbo, gradTex, colId, sigma, F, col, vis, in, out, offsets, false);
    if (rc != SL_OK) return rc;
    const int W = vis->w, H = vis->h;
    const size_t n = (size_t)W * H;
    const float* C = reinterpret_cast<const float*>(col->data.data());
    const float* V = reinterpret_cast<const float*>(vis->data.data());
    const float* G = reinterpret_cast<const float*>(in->data.data());
    const bool smooth = sigma > 0.f && !F->saved.empty();
    const double s2 = (double)sigma * (double)sigma;
    auto weight = [&](const OBnd& B) {
        double p[4];
        const double d = bnd_partials(B, p), r = std::sqrt(d * d + s2);
        return d >= 0.0 ? s2 / (r * (r + d)) : 1.0 - d / r;
    };
    std::vector<double> X(n * 4), O(n * 4);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x)
        {
            const size_t i = (size_t)y * W + x;
            const OBnd B = smooth ? boundary_x(c, *F, V, W, x, y, offsets) : OBnd{};
            const double w = B.ok ? weight(B) : 0.0;
            for (int ch = 0; ch < 4; ++ch)
                X[4 * i + ch] = (1.0 - w) * C[4 * i + ch] + (B.ok ? w * C[4 * (i - 1) + ch] : 0.0);
        }
    O = X;
    for (int y = 0; smooth && y + 1 < H; ++y)
        for (int x = 0; x < W; ++x)
        {
            const OBnd B = boundary_y(c, *F, V, W, H, x, y, offsets);
            if (!B.ok) continue;
            const double w = weight(B);
            const size_t r = (size_t)(B.recvLow ? y : y + 1) * W + x, s = (size_t)(B.recvLow ? y + 1 : y) * W + x;
            for (int ch = 0; ch < 4; ++ch) O[4 * r + ch] += w * (X[4 * s + ch] - X[4 * r + ch]);
        }
    double acc = 0.0;
    for (size_t i = 0; i < n * 4; ++i) acc += (double)G[i] * O[i];
    *loss = acc;
    return SL_OK;
}

int slo_framebuffer_backward(sl_context* c, size_t fbo, size_t gradImageTex)
{
    if (!c) return SL_ERR_INVALID_ARG;
    return backward(c, fbo, gradImageTex);
}

int slo_get_stats(const sl_context* c, sl_frame_stats* out)
{
    if (!c || !out) return SL_ERR_INVALID_ARG;
    *out = c->stats;
    return SL_OK;
}

int slo_reset_stats(sl_context* c)
{
    if (!c) return SL_ERR_INVALID_ARG;
    c->stats = sl_frame_stats{};
    c->triSeq = 0;
    return SL_OK;
}

} // extern "C"
