/*
 * SL_Context.hpp -- header-only C++ facade over include/sl_gpu.h.
 *
 * Keeps the shape of SoftLight's C++ API for the triangle draw path so that a
 * host program written against the reference (softlight/include/softlight/
 * SL_Context.hpp:59-306 and the resource headers) recompiles against the
 * MI355X library with its shader descriptions changed from host function
 * pointers to device program ids:
 *
 *   SL_Context ctx;
 *   size_t vboId = ctx.create_vbo();
 *   ctx.vbo(vboId).init(numBytes, data);                // SL_VertexBuffer::init
 *   SL_VertexArray& vao = ctx.vao(ctx.create_vao());   // set_vertex_buffer/set_binding
 *   ctx.texture(texId).init(SL_COLOR_RGBA_8U, 640, 480);
 *   ctx.framebuffer(fboId).attach_color_buffer(0, ctx.texture(texId).view());
 *   size_t shaderId = ctx.create_shader(vertShader, fragShader, uboId);
 *   ctx.clear_framebuffer(fboId, 0, {0, 0, 0, 0}, 1.0);
 *   ctx.draw(mesh, shaderId, fboId);                    // synchronous, like the reference
 *
 * `draw`, `draw_multiple` and `draw_instanced` return after the GPU finished
 * (the reference's draw returns after its worker pool drained,
 * SL_ProcessorPool.cpp:326); `draw_deferred` and `finish` expose the batched
 * path.  Handles are dense size_t indices as in the reference.
 */
#ifndef SOFTLIGHT_AMD_SL_CONTEXT_HPP
#define SOFTLIGHT_AMD_SL_CONTEXT_HPP

#include <array>
#include <cstddef>
#include <cstdint>
#include <deque>
#include <stdexcept>

#include "../sl_gpu.h"

typedef sl_mesh SL_Mesh;

/* SL_Shader.hpp:41-64 with the function pointer replaced by a device program id */
struct SL_VertexShader
{
    uint8_t  numVaryings;
    uint8_t  cullMode;        /* SL_CULL_* */
    uint32_t programId;       /* SL_PROGRAM_* */
};

/* SL_Shader.hpp:74-98 with the function pointer replaced by a device program id */
struct SL_FragmentShader
{
    uint8_t  numVaryings;
    uint8_t  numOutputs;
    uint8_t  blend;           /* SL_BLEND_* */
    uint8_t  depthTest;       /* SL_DEPTH_TEST_* */
    uint8_t  depthMask;       /* SL_DEPTH_MASK_* */
    uint32_t programId;       /* SL_PROGRAM_* */
};

/* SL_Texture.hpp:29-44: a view names its texture */
struct SL_TextureView
{
    size_t textureId;
};

class SL_VertexBuffer
{
  public:
    SL_VertexBuffer(sl_context* c, size_t id) : mCtx(c), mId(id) {}
    int init(size_t numBytes, const void* pData = nullptr) noexcept { return sl_vbo_init(mCtx, mId, numBytes, pData); }
    void assign(const void* pInputData, ptrdiff_t offset, std::size_t numBytes) noexcept
    {
        (void)sl_vbo_assign(mCtx, mId, pInputData, (size_t)offset, numBytes);
    }
  private:
    sl_context* mCtx;
    size_t mId;
};

class SL_IndexBuffer
{
  public:
    SL_IndexBuffer(sl_context* c, size_t id) : mCtx(c), mId(id) {}
    int init(uint32_t numElements, int type, const void* pData = nullptr) noexcept
    {
        return sl_ibo_init(mCtx, mId, numElements, type, pData);
    }
  private:
    sl_context* mCtx;
    size_t mId;
};

class SL_VertexArray
{
  public:
    SL_VertexArray(sl_context* c, size_t id) : mCtx(c), mId(id) {}
    void set_vertex_buffer(size_t vboId) noexcept { (void)sl_vao_set_vertex_buffer(mCtx, mId, vboId); }
    void set_index_buffer(size_t iboId) noexcept { (void)sl_vao_set_index_buffer(mCtx, mId, iboId); }
    int set_num_bindings(std::size_t n) noexcept { return sl_vao_set_num_bindings(mCtx, mId, n); }
    void set_binding(std::size_t bindId, ptrdiff_t offset, ptrdiff_t stride, int dimens, int type) noexcept
    {
        (void)sl_vao_set_binding(mCtx, mId, bindId, offset, stride, dimens, type);
    }
  private:
    sl_context* mCtx;
    size_t mId;
};

class SL_Texture
{
  public:
    SL_Texture(sl_context* c, size_t id) : mCtx(c), mId(id) {}
    int init(int type, uint16_t w, uint16_t h, uint16_t d = 1) noexcept
    {
        mW = w; mH = h; mD = d;
        return sl_texture_init(mCtx, mId, type, w, h, d);
    }
    int set_texels(const void* data, size_t numBytes) noexcept { return sl_texture_set_texels(mCtx, mId, data, numBytes); }
    int read(void* out, size_t numBytes) noexcept { return sl_texture_read(mCtx, mId, out, numBytes); }
    SL_TextureView view() const noexcept { return SL_TextureView{mId}; }
    uint16_t width() const noexcept { return mW; }
    uint16_t height() const noexcept { return mH; }
    uint16_t depth() const noexcept { return mD; }
  private:
    sl_context* mCtx;
    size_t mId;
    uint16_t mW = 0, mH = 0, mD = 0;
};

class SL_Framebuffer
{
  public:
    SL_Framebuffer(sl_context* c, size_t id) : mCtx(c), mId(id) {}
    int reserve_color_buffers(unsigned n) noexcept { return sl_fbo_reserve_color_buffers(mCtx, mId, n); }
    int attach_color_buffer(unsigned index, const SL_TextureView& t) noexcept { return sl_fbo_attach_color_buffer(mCtx, mId, index, t.textureId); }
    int attach_depth_buffer(const SL_TextureView& d) noexcept { return sl_fbo_attach_depth_buffer(mCtx, mId, d.textureId); }
    int valid() const noexcept { return sl_fbo_valid(mCtx, mId); }
  private:
    sl_context* mCtx;
    size_t mId;
};

class SL_UniformBuffer
{
  public:
    SL_UniformBuffer(sl_context* c, size_t id) : mCtx(c), mId(id) {}
    void assign(const void* pBytes, unsigned long long offset, unsigned long long numBytes) noexcept
    {
        (void)sl_ubo_assign(mCtx, mId, pBytes, (size_t)offset, (size_t)numBytes);
    }
    template <typename T>
    void assign(const T& value, unsigned long long offset = 0) noexcept
    {
        static_assert(sizeof(T) <= SL_MAX_UNIFORM_BUFFER_SIZE, "uniform block too large");
        assign(&value, offset, sizeof(T));
    }
  private:
    sl_context* mCtx;
    size_t mId;
};

class SL_Context
{
  public:
    SL_Context()
    {
        if (sl_context_create(&mCtx) != SL_OK) throw std::runtime_error("SL_Context: no usable GPU");
    }
    ~SL_Context() noexcept { sl_context_destroy(mCtx); }
    SL_Context(const SL_Context&) = delete;
    SL_Context& operator=(const SL_Context&) = delete;

    sl_context* native() noexcept { return mCtx; }

    std::size_t create_vbo() { mVbos.emplace_back(mCtx, sl_create_vbo(mCtx)); return mVbos.size() - 1; }
    SL_VertexBuffer& vbo(std::size_t i) { return mVbos[i]; }
    std::size_t create_ibo() { mIbos.emplace_back(mCtx, sl_create_ibo(mCtx)); return mIbos.size() - 1; }
    SL_IndexBuffer& ibo(std::size_t i) { return mIbos[i]; }
    std::size_t create_vao() { mVaos.emplace_back(mCtx, sl_create_vao(mCtx)); return mVaos.size() - 1; }
    SL_VertexArray& vao(std::size_t i) { return mVaos[i]; }
    std::size_t create_texture() { mTexs.emplace_back(mCtx, sl_create_texture(mCtx)); return mTexs.size() - 1; }
    SL_Texture& texture(std::size_t i) { return mTexs[i]; }
    std::size_t create_framebuffer() { mFbos.emplace_back(mCtx, sl_create_framebuffer(mCtx)); return mFbos.size() - 1; }
    SL_Framebuffer& framebuffer(std::size_t i) { return mFbos[i]; }
    std::size_t create_ubo() { mUbos.emplace_back(mCtx, sl_create_ubo(mCtx)); return mUbos.size() - 1; }
    SL_UniformBuffer& ubo(std::size_t i) { return mUbos[i]; }

    /* SL_Context.cpp:488-604: returns (size_t)-1 on validation failure */
    std::size_t create_shader(const SL_VertexShader& v, const SL_FragmentShader& f)
    {
        return create_shader(v, f, SL_INVALID_ID);
    }
    std::size_t create_shader(const SL_VertexShader& v, const SL_FragmentShader& f, std::size_t uniformIndex)
    {
        const sl_vertex_shader vs{v.numVaryings, v.cullMode, v.programId};
        const sl_fragment_shader fs{f.numVaryings, f.numOutputs, f.blend, f.depthTest, f.depthMask, f.programId};
        return sl_create_shader(mCtx, &vs, &fs, uniformIndex);
    }

    /* SL_Context.cpp:714-740 */
    void draw(const SL_Mesh& m, size_t shaderId, size_t fboId) noexcept
    {
        if (sl_draw(mCtx, &m, shaderId, fboId) == SL_OK) (void)sl_finish(mCtx);
    }
    void draw_multiple(const SL_Mesh* meshes, size_t numMeshes, size_t shaderId, size_t fboId) noexcept
    {
        if (sl_draw_multiple(mCtx, meshes, numMeshes, shaderId, fboId) == SL_OK) (void)sl_finish(mCtx);
    }
    void draw_instanced(const SL_Mesh& m, size_t numInstances, size_t shaderId, size_t fboId) noexcept
    {
        if (sl_draw_instanced(mCtx, &m, numInstances, shaderId, fboId) == SL_OK) (void)sl_finish(mCtx);
    }
    /* batched path: record now, execute at finish() or the next readback */
    int draw_deferred(const SL_Mesh& m, size_t shaderId, size_t fboId) noexcept { return sl_draw(mCtx, &m, shaderId, fboId); }
    int finish() noexcept { return sl_finish(mCtx); }

    /* SL_Context.cpp:881-964 */
    void clear_color_buffer(size_t fboId, unsigned attachmentId, const std::array<double, 4>& color) noexcept
    {
        (void)sl_clear_color_buffer(mCtx, fboId, attachmentId, color.data());
    }
    void clear_depth_buffer(size_t fboId, double depth) noexcept { (void)sl_clear_depth_buffer(mCtx, fboId, depth); }
    void clear_framebuffer(size_t fboId, unsigned attachmentId, const std::array<double, 4>& color, double depth) noexcept
    {
        (void)sl_clear_framebuffer(mCtx, fboId, attachmentId, color.data(), depth);
    }

    /* SL_Context.cpp:1135-1148 (no host worker pool on the GPU path) */
    unsigned num_threads() const noexcept { return sl_num_threads(mCtx); }
    unsigned num_threads(unsigned n) noexcept { return sl_set_num_threads(mCtx, n); }

  private:
    sl_context* mCtx = nullptr;
    std::deque<SL_VertexBuffer> mVbos;
    std::deque<SL_IndexBuffer> mIbos;
    std::deque<SL_VertexArray> mVaos;
    std::deque<SL_Texture> mTexs;
    std::deque<SL_Framebuffer> mFbos;
    std::deque<SL_UniformBuffer> mUbos;
};

#endif /* SOFTLIGHT_AMD_SL_CONTEXT_HPP */
