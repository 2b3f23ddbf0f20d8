// A host program written the way the reference's tests/sl_draw_test.cpp is,
// against the C++ facade of the MI355X library: one indexed triangle per
// 3 vertices, pass-through vertex program, Gouraud fragment program, RGBA8 +
// R32F targets.  Writes the colour and depth buffers to the files named on
// the command line so the test can compare them with the oracle.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "softlight_amd/SL_Context.hpp"

struct ColoredVertex
{
    float pos[3];
    uint8_t color[4];
};

int main(int argc, char** argv)
{
    if (argc < 4) { std::fprintf(stderr, "usage: %s vertices.bin color.out depth.out\n", argv[0]); return 2; }
    std::FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 3;
    std::vector<ColoredVertex> verts;
    ColoredVertex v;
    while (std::fread(&v, sizeof(v), 1, f) == 1) verts.push_back(v);
    std::fclose(f);
    const uint16_t W = 160, H = 120;

    SL_Context context;
    const size_t fboId = context.create_framebuffer();
    const size_t texId = context.create_texture();
    const size_t depthId = context.create_texture();
    const size_t vaoId = context.create_vao();
    const size_t vboId = context.create_vbo();

    const SL_VertexShader vertShader{1, SL_CULL_OFF, SL_PROGRAM_PASSTHROUGH_GOURAUD};
    const SL_FragmentShader fragShader{1, 1, SL_BLEND_OFF, SL_DEPTH_TEST_LESS_THAN, SL_DEPTH_MASK_ON, SL_PROGRAM_PASSTHROUGH_GOURAUD};
    const size_t shaderId = context.create_shader(vertShader, fragShader);
    if (shaderId == SL_INVALID_ID) return 4;

    if (context.vbo(vboId).init(verts.size() * sizeof(ColoredVertex), verts.data()) != 0) return 5;
    SL_VertexArray& vao = context.vao(vaoId);
    vao.set_vertex_buffer(vboId);
    if (vao.set_num_bindings(2) != 2) return 6;
    vao.set_binding(0, 0, sizeof(ColoredVertex), VERTEX_DIMENSION_3, VERTEX_DATA_FLOAT);
    vao.set_binding(1, 12, sizeof(ColoredVertex), VERTEX_DIMENSION_4, VERTEX_DATA_BYTE);

    if (context.texture(texId).init(SL_COLOR_RGBA_8U, W, H, 1) != 0) return 7;
    if (context.texture(depthId).init(SL_COLOR_R_FLOAT, W, H, 1) != 0) return 8;
    SL_Framebuffer& fbo = context.framebuffer(fboId);
    if (fbo.reserve_color_buffers(1) != 0) return 9;
    if (fbo.attach_color_buffer(0, context.texture(texId).view()) != 0) return 10;
    if (fbo.attach_depth_buffer(context.texture(depthId).view()) != 0) return 11;
    if (fbo.valid() != 0) return 12;

    context.clear_framebuffer(fboId, 0, {0.0, 0.0, 0.0, 0.0}, 1.0);
    SL_Mesh m{vaoId, 0, verts.size(), RENDER_MODE_TRIANGLES, 0};
    context.draw(m, shaderId, fboId);

    std::vector<uint8_t> color((size_t)W * H * 4);
    std::vector<float> depth((size_t)W * H);
    if (context.texture(texId).read(color.data(), color.size()) != 0) return 13;
    if (context.texture(depthId).read(depth.data(), depth.size() * 4) != 0) return 14;
    std::FILE* c = std::fopen(argv[2], "wb");
    std::FILE* d = std::fopen(argv[3], "wb");
    if (!c || !d) return 15;
    std::fwrite(color.data(), 1, color.size(), c);
    std::fwrite(depth.data(), 4, depth.size(), d);
    std::fclose(c);
    std::fclose(d);
    std::printf("facade draw ok: %zu triangles\n", verts.size() / 3);
    return 0;
}
