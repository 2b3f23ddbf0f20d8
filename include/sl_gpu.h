/*
 * sl_gpu.h -- C ABI of the MI355X (gfx950) SoftLight triangle pipeline.
 *
 * This is the drop-in boundary for SoftLight's `SL_Context::draw` path.  Every
 * entry point below replaces one member of the reference's C++ API; the
 * citation on each names the reference declaration/definition it stands in
 * for (paths relative to the SoftLight source tree).  Resource handles keep the
 * reference's semantics: dense `size_t` indices returned by `sl_create_*`.
 * Error codes keep the reference's integer conventions (0 = ok, negative =
 * error) and add SL_ERR_* codes where the reference had undefined behaviour
 * (bad handles).
 *
 * Host function-pointer shaders (SL_Shader.hpp:58-98) cannot run on the GPU;
 * shaders select a built-in device program by `programId` instead, and take a
 * POD uniform block (no host pointers).  See sl_std_uniforms below.
 *
 * Execution model: draw and clear calls are recorded into a per-context
 * command list and executed as one batched frame on the context's HIP stream
 * at the next synchronisation point (sl_finish, a readback, a resource upload,
 * or a draw into a different framebuffer).  Uniform blocks are snapshotted at
 * draw time, so updating a uniform buffer between draws behaves as in the
 * reference's synchronous `draw`.
 */
#ifndef SL_GPU_H
#define SL_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------
 * Limits (SL_Config.hpp:36-58, SL_VertexArray.hpp:18, SL_UniformBuffer.hpp:12)
 * ------------------------------------------------------------------------- */
#define SL_SHADER_MAX_VARYING_VECTORS 4
#define SL_SHADER_MAX_FRAG_OUTPUTS    4
#define SL_VAO_MAX_BINDINGS           8
#define SL_MAX_UNIFORM_BUFFER_SIZE    1024
#define SL_INVALID_ID                 ((size_t)-1)
#define SL_NO_RESOURCE                0xFFFFFFFFu

/* ---------------------------------------------------------------------------
 * Error codes.  Reference init() calls return 0 / small negative numbers; the
 * codes below are used where the reference had no error path.
 * ------------------------------------------------------------------------- */
#define SL_OK                  0
#define SL_ERR_INVALID_HANDLE  (-100)
#define SL_ERR_INVALID_ARG     (-101)
#define SL_ERR_OUT_OF_MEMORY   (-102)
#define SL_ERR_DEVICE          (-103)
#define SL_ERR_UNSUPPORTED     (-104)
#define SL_ERR_CAPACITY        (-105)
#define SL_ERR_STATE           (-106)

/* ---------------------------------------------------------------------------
 * Enumerations, values identical to the reference.
 * ------------------------------------------------------------------------- */
/* SL_PipelineState.hpp: SL_CullMode */
enum { SL_CULL_BACK_FACE = 0, SL_CULL_FRONT_FACE = 1, SL_CULL_OFF = 2 };
/* SL_PipelineState.hpp: SL_DepthTest */
enum {
    SL_DEPTH_TEST_OFF = 0, SL_DEPTH_TEST_LESS_THAN, SL_DEPTH_TEST_LESS_EQUAL,
    SL_DEPTH_TEST_GREATER_THAN, SL_DEPTH_TEST_GREATER_EQUAL,
    SL_DEPTH_TEST_EQUAL, SL_DEPTH_TEST_NOT_EQUAL
};
/* SL_PipelineState.hpp: SL_DepthMask */
enum { SL_DEPTH_MASK_OFF = 0, SL_DEPTH_MASK_ON = 1 };
/* SL_PipelineState.hpp: SL_BlendMode */
enum {
    SL_BLEND_OFF = 0, SL_BLEND_ALPHA, SL_BLEND_PREMULTIPLED_ALPHA,
    SL_BLEND_ADDITIVE, SL_BLEND_SCREEN
};
/* SL_Mesh.hpp:13-23: SL_RenderMode */
enum {
    RENDER_MODE_POINTS = 0x01, RENDER_MODE_INDEXED_POINTS = 0x03,
    RENDER_MODE_LINES = 0x04, RENDER_MODE_INDEXED_LINES = 0x0C,
    RENDER_MODE_TRIANGLES = 0x10, RENDER_MODE_INDEXED_TRIANGLES = 0x30,
    RENDER_MODE_TRI_WIRE = 0x50, RENDER_MODE_INDEXED_TRI_WIRE = 0x70
};
/* SL_Geometry.hpp:17-39: SL_Dimension, SL_DataType */
enum { VERTEX_DIMENSION_1 = 0, VERTEX_DIMENSION_2, VERTEX_DIMENSION_3, VERTEX_DIMENSION_4 };
enum {
    VERTEX_DATA_BYTE = 0, VERTEX_DATA_SHORT, VERTEX_DATA_INT, VERTEX_DATA_LONG,
    VERTEX_DATA_FLOAT, VERTEX_DATA_DOUBLE, VERTEX_DATA_INVALID
};
/* SL_Color.hpp:31-73: SL_ColorDataType.  Textures of every format can be
 * created; draws write 8-bit and float R/RG/RGB/RGBA colour, the packed
 * RGB_332 / RGB_565 / RGBA_5551 / RGBA_4444 / RGBA_1010102 formats, and
 * R_FLOAT or R_HALF depth; R_32U holds the debug buffers; samplers read
 * RGBA_8U and RGBA_FLOAT; the blit reads RGBA_8U, RGBA_FLOAT and R_FLOAT. */
enum {
    SL_COLOR_R_8U = 0, SL_COLOR_R_16U, SL_COLOR_R_32U, SL_COLOR_R_64U,
    SL_COLOR_R_HALF, SL_COLOR_R_FLOAT, SL_COLOR_R_DOUBLE,
    SL_COLOR_RG_8U, SL_COLOR_RG_16U, SL_COLOR_RG_32U, SL_COLOR_RG_64U,
    SL_COLOR_RG_HALF, SL_COLOR_RG_FLOAT, SL_COLOR_RG_DOUBLE,
    SL_COLOR_RGB_8U, SL_COLOR_RGB_16U, SL_COLOR_RGB_32U, SL_COLOR_RGB_64U,
    SL_COLOR_RGB_HALF, SL_COLOR_RGB_FLOAT, SL_COLOR_RGB_DOUBLE,
    SL_COLOR_RGBA_8U, SL_COLOR_RGBA_16U, SL_COLOR_RGBA_32U, SL_COLOR_RGBA_64U,
    SL_COLOR_RGBA_HALF, SL_COLOR_RGBA_FLOAT, SL_COLOR_RGBA_DOUBLE,
    SL_COLOR_RGB_332, SL_COLOR_RGB_565, SL_COLOR_RGBA_5551, SL_COLOR_RGBA_4444, SL_COLOR_RGBA_1010102
};

/* ---------------------------------------------------------------------------
 * Built-in device shader programs (replace the host function pointers of
 * SL_VertexShader::shader / SL_FragmentShader::shader, SL_Shader.hpp:58-98).
 *
 * SL_PROGRAM_PASSTHROUGH_GOURAUD
 *   VS: position = binding 0 (float, 3 or 4 components; w = 1 when 3);
 *       varying0 = binding 1 colour (4 x u8 scaled by 1/255 as
 *       color_cast<float,uint8_t>, SL_Color.hpp:666-669, or 4 x float).
 *   FS: output0 = varying0.  (Pattern of tests/sl_draw_test.cpp:36-81.)
 * SL_PROGRAM_LAMBERT_TEXTURED
 *   VBO vertex = {float pos[3]; float normal[3]; float uv[2]} (bindings 0,1,2).
 *   VS: model = uniforms.model, or the instance's mat4 from uniforms.instance_vbo;
 *       world = model*(pos,1); n = model*(normal,0); clip = uniforms.mvp*world.
 *       varying0 = (u, v, 0, 0), varying1 = (n, 0).
 *   FS: albedo = bilinear REPEAT sample of an RGBA8 texture (SL_Sampler.hpp:186-218;
 *       nearest with SL_UNIFORM_SAMPLE_NEAREST);
 *       lambert = max(0, -dot(normalize(n), light_dir));
 *       rgb = min(albedo.rgb * (ambient + diffuse*lambert) * tint.rgb, 1);
 *       a = albedo.a * tint.a.  (Pattern of tests/sl_instancing_test.cpp:105-178.)
 * SL_PROGRAM_COLOR_TEXTURED
 *   VBO vertex = {float pos[3]; float rgb[3]; float uv[2]}.
 *   VS: clip = uniforms.mvp*(model*(pos,1)); varying0 = (rgb, 1); varying1 = (u, v, 0, 0).
 *   FS: output0 = varying0 * bilinear REPEAT sample of an RGBA32F texture
 *       (nearest with SL_UNIFORM_SAMPLE_NEAREST).
 *       This is the program of the differentiable mode.
 * SL_PROGRAM_LAMBERT_MRT
 *   SL_PROGRAM_LAMBERT_TEXTURED with the G-buffer outputs of the reference's
 *   multiple-render-target test (tests/sl_mrt_test.cpp:105-123):
 *   VS adds varying2 = (world, 1).  FS: output0 = the Lambert colour,
 *   output1 = clamp(varying2, 0, 1), output2 = clamp(varying0, 0, 1),
 *   output3 = clamp((normalize(n), 0), 0, 1); output k goes to colour
 *   attachment k for k < numOutputs (up to 4).  Use numVaryings = 3.
 * ------------------------------------------------------------------------- */
enum {
    SL_PROGRAM_NONE = 0,
    SL_PROGRAM_PASSTHROUGH_GOURAUD = 1,
    SL_PROGRAM_LAMBERT_TEXTURED = 2,
    SL_PROGRAM_COLOR_TEXTURED = 3,
    SL_PROGRAM_LAMBERT_MRT = 4
};

/* sl_std_uniforms::flags.  SL_UNIFORM_SAMPLE_NEAREST: the textured programs
 * sample with sl_sample_nearest<..., SL_WrapModeRepeat> (SL_Sampler.hpp:
 * 122-148: the texel at (uint)(W * wrap(u)), (uint)(H * wrap(v)) in ordered
 * layout) instead of bilinear filtering; the backward pass then sends each
 * fragment's texel gradient to that one texel. */
#define SL_UNIFORM_SAMPLE_NEAREST 0x1u

/* POD uniform block read by the built-in programs (lives in an SL_UniformBuffer,
 * SL_UniformBuffer.hpp:12; matrices column-major like ls::math::mat4). */
typedef struct sl_std_uniforms {
    float mvp[16];
    float model[16];
    float light_dir[4];
    float ambient[4];
    float diffuse[4];
    float tint[4];
    uint32_t texture;       /* texture handle or SL_NO_RESOURCE */
    uint32_t instance_vbo;  /* vbo holding one mat4 per instance, or SL_NO_RESOURCE */
    uint32_t flags;         /* SL_UNIFORM_* bits */
    uint32_t reserved;
} sl_std_uniforms;

/* SL_Mesh.hpp:30-39 */
typedef struct sl_mesh {
    size_t vaoId;
    size_t elementBegin;
    size_t elementEnd;
    uint32_t mode;
    uint32_t materialId;
} sl_mesh;

/* SL_Shader.hpp:41-64: SL_VertexShader, with the host pointer replaced by a program id */
typedef struct sl_vertex_shader {
    uint8_t  numVaryings;
    uint8_t  cullMode;
    uint32_t programId;
} sl_vertex_shader;

/* SL_Shader.hpp:74-98: SL_FragmentShader, with the host pointer replaced by a program id */
typedef struct sl_fragment_shader {
    uint8_t  numVaryings;
    uint8_t  numOutputs;
    uint8_t  blend;
    uint8_t  depthTest;
    uint8_t  depthMask;
    uint32_t programId;
} sl_fragment_shader;

/* Per-frame counters.  Not in the reference (it has no counters); they carry
 * the quantities the parity tests and the benchmark report. */
typedef struct sl_frame_stats {
    uint64_t triangles_submitted;   /* triangles fetched by the vertex stage */
    uint64_t triangles_binned;      /* setup records that reached binning (after cull/clip/bbox reject) */
    uint64_t tile_pairs;            /* (tile, record) bin entries */
    uint64_t fragments_tested;      /* covered samples that were depth tested */
    uint64_t fragments_shaded;      /* samples that passed the depth test and ran the fragment program */
    uint64_t frames;                /* batched frames executed */
    double   ms_vertex;             /* vertex + setup stage (device time) */
    double   ms_bin;                /* tile binning (device time) */
    double   ms_raster;             /* raster + fragment stage (device time) */
    double   ms_total;              /* whole batched frame (device time) */
    double   ms_setup_kernel;       /* setup's counting pass alone (part of ms_vertex) */
    double   ms_clip_kernel;        /* the clip kernel alone (part of ms_vertex) */
    uint64_t setup_launches;        /* setup kernel launches (re-runs after a capacity grow included) */
    uint64_t raster_launches;       /* raster kernel launches */
    double   ms_backward;           /* backward kernels (device time, counted at the next synchronisation) */
    uint64_t backward_launches;     /* backward passes executed (counted likewise) */
    uint64_t visibility_rebuilds;   /* framebuffers redrawn because a shared visibility attachment held another's draws */
} sl_frame_stats;

typedef struct sl_context sl_context;

/* ---- context (SL_Context.hpp:59-306) ------------------------------------ */
int         sl_context_create(sl_context** outCtx);            /* SL_Context::SL_Context(), SL_Context.cpp */
void        sl_context_destroy(sl_context* ctx);               /* SL_Context::~SL_Context / terminate(), SL_Context.cpp */
const char* sl_last_error(const sl_context* ctx);
unsigned    sl_num_threads(const sl_context* ctx);             /* SL_Context::num_threads(), SL_Context.cpp:1135 */
unsigned    sl_set_num_threads(sl_context* ctx, unsigned n);   /* SL_Context::num_threads(unsigned), SL_Context.cpp:1145 */
int         sl_set_viewport(sl_context* ctx, int32_t x, int32_t y, int32_t w, int32_t h); /* SL_ViewportState::viewport, SL_ViewportState.hpp:150 */
int         sl_set_scissor(sl_context* ctx, int32_t x, int32_t y, int32_t w, int32_t h);  /* SL_ViewportState::scissor, SL_ViewportState.hpp:175 */

/* Raster tile configuration (no counterpart in the reference, whose raster
 * splits scanlines across CPU threads, SL_ScanlineBounds.hpp:31-50).
 * Configurations 0..11 select the screen tile each GPU workgroup owns (8x8
 * to 64x64) and how many column segments a tile row is split into; see
 * sl_get_raster_config.  Configuration 11 (64x64) falls back to 7 (32x32)
 * for a batch with an RGBA32F colour target or a visibility attachment.
 * sl_get_raster_config(ctx, -1, ...) reports the current one and returns its
 * index; results do not depend on the configuration. */
int         sl_set_raster_config(sl_context* ctx, int config);
/* Heavy tiles (no counterpart in the reference): when a batch's draws are
 * all unblended with one depth function on an R_FLOAT depth target (depth
 * writes on with LESS / LESS_EQUAL / GREATER / GREATER_EQUAL, or off), a tile
 * with more than 4 * partRecords records is split into parts of about
 * partRecords records (at most 1024 parts) rasterised by separate workgroups
 * and merged exactly.  By default batches with point or line draws split
 * with parts of 256 records; sl_set_raster_split(n > 0) splits every such
 * batch with parts of n records, 0 turns the split off.  Results never
 * depend on it. */
int         sl_set_raster_split(sl_context* ctx, uint32_t partRecords);
#define SL_RASTER_SPLIT_DEFAULT 0xFFFFFFFFu   /* sl_set_raster_split: back to the default policy */
/* Occupancy (no counterpart in the reference): caps the raster workgroups
 * resident on a CU at workgroupsPerCU (1..32; 0, the default: as many as the
 * kernel's registers and LDS allow) by giving each workgroup that share of
 * the CU's LDS.  Results never depend on it. */
int         sl_set_raster_occupancy(sl_context* ctx, uint32_t workgroupsPerCU);
/* Depth rejection (no counterpart in the reference): with enable non-zero,
 * batches of filled triangles without a visibility attachment use the
 * raster variant that skips, for each wave of a tile holding 64 or more
 * records, the records whose depth provably fails a LESS / LESS_EQUAL test
 * at every pixel the wave owns (R_FLOAT depth targets only; such records
 * only count their depth tests).  Off by default: it pays off where depth
 * complexity is high.  Results never depend on it. */
int         sl_set_raster_depth_reject(sl_context* ctx, int enable);
int         sl_get_raster_config(const sl_context* ctx, int config, int* tileW, int* tileH, int* segments);

/* ---- vertex buffers (SL_VertexBuffer.hpp / .cpp) ------------------------ */
size_t sl_create_vbo(sl_context* ctx);                                                      /* SL_Context::create_vbo */
int    sl_vbo_init(sl_context* ctx, size_t vbo, size_t numBytes, const void* data);         /* SL_VertexBuffer::init, SL_VertexBuffer.cpp:96 */
int    sl_vbo_assign(sl_context* ctx, size_t vbo, const void* data, size_t offset, size_t numBytes); /* SL_VertexBuffer::assign, SL_VertexBuffer.hpp */
int    sl_vbo_read(sl_context* ctx, size_t vbo, void* out, size_t offset, size_t numBytes);         /* SL_VertexBuffer::data, SL_VertexBuffer.hpp:45-47,95-110 (read back) */

/* ---- index buffers (SL_IndexBuffer.hpp / .cpp) -------------------------- */
size_t sl_create_ibo(sl_context* ctx);                                                      /* SL_Context::create_ibo */
int    sl_ibo_init(sl_context* ctx, size_t ibo, uint32_t numElements, int dataType, const void* data); /* SL_IndexBuffer::init, SL_IndexBuffer.cpp:138 */

/* ---- vertex arrays (SL_VertexArray.hpp / .cpp) -------------------------- */
size_t sl_create_vao(sl_context* ctx);                                                      /* SL_Context::create_vao, SL_Context.cpp:181 */
int    sl_vao_set_vertex_buffer(sl_context* ctx, size_t vao, size_t vbo);                   /* SL_VertexArray::set_vertex_buffer */
int    sl_vao_set_index_buffer(sl_context* ctx, size_t vao, size_t ibo);                    /* SL_VertexArray::set_index_buffer */
int    sl_vao_set_num_bindings(sl_context* ctx, size_t vao, size_t numBindings);            /* SL_VertexArray::set_num_bindings, SL_VertexArray.cpp:162 */
int    sl_vao_set_binding(sl_context* ctx, size_t vao, size_t bindId, ptrdiff_t offset,
                          ptrdiff_t stride, int dimens, int dataType);                     /* SL_VertexArray::set_binding, SL_VertexArray.cpp:205 */

/* ---- textures (SL_Texture.hpp / .cpp) ----------------------------------- */
size_t sl_create_texture(sl_context* ctx);                                                  /* SL_Context::create_texture, SL_Context.cpp:232 */
int    sl_texture_init(sl_context* ctx, size_t tex, int colorType, uint16_t w, uint16_t h, uint16_t d); /* SL_Texture::init, SL_Texture.cpp:276 */
int    sl_texture_set_texels(sl_context* ctx, size_t tex, const void* data, size_t numBytes);           /* SL_Texture::set_texels (ordered layout) */
int    sl_texture_read(sl_context* ctx, size_t tex, void* out, size_t numBytes);            /* SL_Texture::data() readback */

/* ---- framebuffers (SL_Framebuffer.hpp / .cpp) --------------------------- */
size_t sl_create_framebuffer(sl_context* ctx);                                              /* SL_Context::create_framebuffer */
int    sl_fbo_reserve_color_buffers(sl_context* ctx, size_t fbo, unsigned n);              /* SL_Framebuffer::reserve_color_buffers, SL_Framebuffer.cpp:823 */
int    sl_fbo_attach_color_buffer(sl_context* ctx, size_t fbo, unsigned index, size_t tex); /* SL_Framebuffer::attach_color_buffer, SL_Framebuffer.cpp:855 */
int    sl_fbo_attach_depth_buffer(sl_context* ctx, size_t fbo, size_t tex);                 /* SL_Framebuffer::attach_depth_buffer, SL_Framebuffer.cpp:890 */
int    sl_fbo_valid(sl_context* ctx, size_t fbo);                                           /* SL_Framebuffer::valid, SL_Framebuffer.cpp:930-990 */
/* Optional parity attachments (no counterpart in the reference): per pixel,
 * the 1-based submission number of the last triangle that passed the depth
 * test, and the number of depth-test passes.  Both must be R_32U textures of
 * the framebuffer's size; pass SL_INVALID_ID to detach. */
int    sl_fbo_attach_debug_buffers(sl_context* ctx, size_t fbo, size_t lastPrimTex, size_t passCountTex);

/* ---- uniform buffers (SL_UniformBuffer.hpp) ----------------------------- */
size_t sl_create_ubo(sl_context* ctx);                                                      /* SL_Context::create_ubo */
int    sl_ubo_assign(sl_context* ctx, size_t ubo, const void* data, size_t offset, size_t numBytes); /* SL_UniformBuffer::assign */

/* ---- shaders (SL_Context.cpp:488-604) ------------------------------------ */
size_t sl_create_shader(sl_context* ctx, const sl_vertex_shader* vs,
                        const sl_fragment_shader* fs, size_t uboId);   /* SL_Context::create_shader; uboId may be SL_INVALID_ID */

/* ---- drawing (SL_Context.cpp:714-740) ------------------------------------
 * Every render mode is drawn: filled and wireframe triangles
 * (SL_TriRasterizer.cpp:116-202, 806-923), lines (SL_LineProcessor.cpp,
 * SL_LineRasterizer.cpp) and points (SL_PointProcessor.cpp,
 * SL_PointRasterizer.cpp), indexed or not.  As in the reference, the render
 * mode of the first mesh of an sl_draw_multiple call picks the processor and
 * rasteriser for all of its meshes, and each mesh's own mode whether it reads
 * its index buffer.  A point or line texel is addressed as x + W*y: a snapped
 * x of W lands on texel 0 of row y + 1; texels at or past W*H are dropped.
 * Limits: visibility attachments are for filled triangles only; line draws
 * need a depth target of at most 8192 x 8192 (SL_ERR_UNSUPPORTED).
 * Targets: depth R_FLOAT or R_HALF; colour attachment 0 8-bit (R, RG, RGB,
 * RGBA) or float (R, RG, RGB, RGBA), written as assign_pixel /
 * assign_alpha_pixel do (SL_Framebuffer.cpp:42-80, 268-370).  Visibility
 * attachments need RGBA_FLOAT colour and R_FLOAT depth. */
int sl_draw(sl_context* ctx, const sl_mesh* m, size_t shaderId, size_t fboId);                        /* SL_Context::draw */
int sl_draw_multiple(sl_context* ctx, const sl_mesh* m, size_t numMeshes, size_t shaderId, size_t fboId); /* SL_Context::draw_multiple */
int sl_draw_instanced(sl_context* ctx, const sl_mesh* m, size_t numInstances, size_t shaderId, size_t fboId); /* SL_Context::draw_instanced */

/* ---- a draw that allocates nothing (no counterpart in the reference) ------
 * For callers that cannot allocate device memory (serving loops and
 * harnesses that forbid hipMalloc): sl_draw_static records the draw like
 * sl_draw and runs the pending batch at once, with every per-frame buffer --
 * setup records, varyings, bin entries, clip jobs, counters, tile ranges, the
 * draw table and the uniform snapshots -- laid out in one device array of
 * SL_STATIC_SCRATCH_BYTES compiled into the library (each GPU holds its own
 * copy until the process ends).  It calls no hipMalloc, hipFree or other
 * allocating routine and returns when the frame is complete.  A frame whose
 * buffers do not fit returns SL_ERR_CAPACITY before any pixel is touched.
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (from several host threads or contexts) race.  Results equal sl_draw's
 * followed by sl_finish. */
#define SL_STATIC_SCRATCH_BYTES ((size_t)64 << 20)
int sl_draw_static(sl_context* ctx, const sl_mesh* m, size_t shaderId, size_t fboId);

/* ---- clears (SL_Context.cpp:881-964) ------------------------------------- */
int sl_clear_color_buffer(sl_context* ctx, size_t fbo, unsigned attachment, const double rgba[4]);      /* SL_Context::clear_color_buffer */
int sl_clear_depth_buffer(sl_context* ctx, size_t fbo, double depth);                                  /* SL_Context::clear_depth_buffer */
int sl_clear_framebuffer(sl_context* ctx, size_t fbo, unsigned attachment, const double rgba[4], double depth); /* SL_Context::clear_framebuffer */

/* ---- blit (SL_Context.cpp:747-875, SL_BlitProcessor.cpp:620-665) --------
 * Nearest-neighbour copy of a texture region into another texture with
 * 16.16 fixed-point scaling, the source rows read bottom-up, and the colour
 * conversion of the reference's blit operators.  Sources: RGBA_8U,
 * RGBA_FLOAT, R_FLOAT (written as (0, 0, r, 1), SL_BlitProcessor.cpp:303-306);
 * destinations: RGBA_8U, RGBA_FLOAT.  Where the reference would read or
 * write outside a texture (a destination region past its height, a source
 * coordinate past the source), those pixels are skipped. */
int sl_blit(sl_context* ctx, size_t outTex, size_t inTex);                       /* SL_Context::blit(out, in) */
int sl_blit_region(sl_context* ctx, size_t outTex, size_t inTex,
                   uint16_t srcX0, uint16_t srcY0, uint16_t srcX1, uint16_t srcY1,
                   uint16_t dstX0, uint16_t dstY0, uint16_t dstX1, uint16_t dstY1);  /* SL_Context::blit(out, in, src/dst rects) */

/* ---- HIP-graph frames (no reference counterpart) -------------------------
 * sl_frame_graph_capture executes the pending batch (the clears and draws
 * recorded since the last synchronisation point) once, then captures the same
 * frame -- uniform snapshots included -- into a HIP graph with private
 * buffers sized from that run.  sl_frame_graph_launch replays it on the
 * context's stream without host round trips; sl_finish waits and adds the
 * replays to the counters (SL_ERR_CAPACITY if a replay outgrew the captured
 * capacity).  One graph per context; capturing again replaces it. */
int sl_frame_graph_capture(sl_context* ctx);
int sl_frame_graph_launch(sl_context* ctx);
int sl_frame_graph_release(sl_context* ctx);

/* ---- differentiable draws (no reference counterpart) ---------------------
 * Fixed-visibility adjoint of the draw path for the programs whose output is
 * smooth in their inputs: SL_PROGRAM_COLOR_TEXTURED (gradients to the vertex
 * colours and to the RGBA32F texels) and SL_PROGRAM_PASSTHROUGH_GOURAUD
 * (gradients to the vertex colours).
 *
 * sl_framebuffer_attach_visibility: an RGBA32F texture of the framebuffer's
 *   size.  While attached, every fragment that passes the depth test writes
 *   {seq, b0, b1, b2}: seq = the 1-based submitted-triangle number as the bit
 *   pattern of the red channel (0 = no fragment), b = the fragment's
 *   perspective-correct weights of the source triangle's three vertices,
 *   carried through clipping like a varying.  Needs numVaryings <= 3 in the
 *   shaders drawn (the weights use the next varying slot).  Clears write 0.
 *   texId = SL_NO_RESOURCE detaches.  Several framebuffers may share one
 *   attachment (visibility recomputed instead of kept): when it holds another
 *   framebuffer's draws, the backward passes of this one (smoothing and
 *   fixed-visibility) first draw its saved draws again from their depth clear
 *   -- the same visibility, depth and colour bit for bit -- and count it in
 *   sl_frame_stats::visibility_rebuilds.  Draws that do not start from a depth
 *   clear return SL_ERR_STATE there, blended draws SL_ERR_UNSUPPORTED.
 * sl_vbo_attach_gradient: a VBO receiving dL/d(colour attribute) as one
 *   float4 per vertex id (binding 1 of the draws that read the VBO).
 * sl_texture_attach_gradient: an RGBA32F texture of the same size receiving
 *   dL/d(texel); bilinear taps at x = W / y = H land where the forward read
 *   them (the next row's first texel / the padding).
 * sl_gradients_zero: zeroes every attached gradient buffer (position
 *   gradients of the smoothed-coverage mode included).
 * sl_loss_l2_grad: grad = scale * 2 * (color - target) for two RGBA32F
 *   textures of one size (the gradient of scale * sum((color - target)^2)).
 * sl_framebuffer_backward: adds the gradient of the framebuffer's last
 *   executed batch, given dL/d(colour attachment 0) as an RGBA32F texture,
 *   into the attached gradient buffers.  The batch must have run with a
 *   visibility attachment, blending off, and resources unchanged since
 *   (re-initialising a VBO, IBO or texture drops the saved batch).  Like
 *   sl_loss_l2_grad and sl_gradients_zero it is queued on the context's
 *   stream; readbacks and sl_finish wait for it. */
int sl_framebuffer_attach_visibility(sl_context* ctx, size_t fboId, size_t texId);
int sl_vbo_attach_gradient(sl_context* ctx, size_t vboId, size_t gradVboId);
int sl_texture_attach_gradient(sl_context* ctx, size_t texId, size_t gradTexId);
int sl_gradients_zero(sl_context* ctx);
int sl_loss_l2_grad(sl_context* ctx, size_t colorTex, size_t targetTex, size_t gradImageTex, float scale);
int sl_framebuffer_backward(sl_context* ctx, size_t fboId, size_t gradImageTex);

/* ---- smoothed span coverage (no reference counterpart): gradients to the
 *      screen positions of the vertices ---------------------------------
 * The hard coverage rule of SL_ScanlineBounds::step (SL_ScanlineBounds.hpp:
 * 310-322) puts pixel x of row y inside a triangle when its span ends satisfy
 * e_min < x + 1 <= e_max.  The smoothed image is the colour attachment with
 * every horizontal coverage boundary smeared to the right: where the
 * visibility winners of pixels x-1 and x differ and a span end e of one of
 * them lies in [x, x + 1) (the right end of the left pixel's triangle, else
 * the left end of the right pixel's), out(x) = (1 - w) C(x) + w C(x-1) with
 * w = exp(-(x + 1 - e) / sigma).  w is continuous in e and tends to 0 with
 * sigma for every such e, so sigma = 0 returns the colour attachment bit for
 * bit.  Boundaries of clipped triangles are not smoothed.
 *
 * sl_framebuffer_attach_positions: buffers of float2 per (saved draw,
 *   vertex id) -- draw k of the framebuffer's saved batch (submission order
 *   since the last depth clear) at float2 offset k * numVertices: gradVbo
 *   receives dL/d(screen x, y) of the vertex, offsetVbo (optional) adds a
 *   screen-space offset to the snapped vertex position inside the smoothing
 *   passes.  SL_INVALID_ID detaches either.
 * sl_smooth_coverage: writes the smoothed image of the framebuffer's last
 *   batch (visibility attached, RGBA_FLOAT colour attachment 0) to outTex
 *   (RGBA_FLOAT, framebuffer size, not the colour attachment).
 * sl_smooth_coverage_backward: given dL/d(smoothed image), writes dL/d(colour
 *   attachment) to gradImageTex (which sl_framebuffer_backward then takes) and
 *   adds dL/d(screen position) into the attached position gradients, with the
 *   visibility winners held fixed.  Queued on the context's stream. */
int sl_framebuffer_attach_positions(sl_context* ctx, size_t fboId, size_t gradVboId, size_t offsetVboId, uint32_t numVertices);
int sl_smooth_coverage(sl_context* ctx, size_t fboId, size_t outTexId, float sigma);
int sl_smooth_coverage_backward(sl_context* ctx, size_t fboId, size_t gradSmoothTexId, float sigma, size_t gradImageTexId);

/* ---- synchronisation and counters (no reference counterpart: the
 *      reference's draw is synchronous, SL_ProcessorPool.cpp:326) ---------- */
int sl_finish(sl_context* ctx);
int sl_get_stats(const sl_context* ctx, sl_frame_stats* out);
int sl_reset_stats(sl_context* ctx);

#ifdef __cplusplus
}
#endif

#endif /* SL_GPU_H */
