// A host program that draws a scene file the way the reference's scene demos
// do (load a file into a scene graph, move nodes, update the graph, draw
// every mesh node with its model matrix and material; softlight/tests/
// sl_instancing_test.cpp:233-280), through the C++ facade of the MI355X
// library.
//
//   scene_graph_draw dump   scene.obj nodes.txt outdir
//       load and build the graph without a GPU; write the graph, the model
//       matrices, the buffers and the draw list to outdir (built with
//       SL_SCENE_DUMP_ONLY this is the only mode, and the program does not
//       link the library)
//   scene_graph_draw render scene.obj nodes.txt W H color.out depth.out [frames]
//       the same on the GPU: import, update, submit the draw list into an
//       RGBA8 + R32F target, read both back; with frames > 1 also time the
//       per-frame host work and print it as one JSON line; the base uniform
//       block is written to color.out.uniforms
//
// nodes.txt, applied in order after loading (names are node names; '-' is
// the root):
//   xform  NAME px py pz sx sy sz qx qy qz qw     set position, scale, orientation
//   rotate NAME qx qy qz qw                       orientation = normalize(orientation * q)
//   inst   NEW PARENT SOURCE px py pz sx sy sz qx qy qz qw
//                                                 a mesh node drawing SOURCE's meshes
//   empty  NEW PARENT px py pz sx sy sz qx qy qz qw
//   mnode  NEW PARENT K MESH_1 .. MESH_K px py pz sx sy sz qx qy qz qw
//                                                 a mesh node drawing the graph's meshes MESH_i
//   update                                        update the graph now
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "softlight_amd/SL_SceneFileLoader.hpp"

namespace {

bool read_transform(std::istringstream& ls, SL_Transform& t)
{
    float p[3], s[3], q[4];
    if (!(ls >> p[0] >> p[1] >> p[2] >> s[0] >> s[1] >> s[2] >> q[0] >> q[1] >> q[2] >> q[3])) return false;
    t.position(p[0], p[1], p[2]);
    t.scaling(s[0], s[1], s[2]);
    t.orientation(SL_Quat{q[0], q[1], q[2], q[3]});
    return true;
}

size_t node_of(const SL_SceneGraph& g, const std::string& name)
{
    return name == "-" ? SCENE_NODE_ROOT_ID : g.find_node_id(name);
}

int apply_commands(SL_SceneGraph& g, const std::string& path)
{
    std::ifstream in(path);
    if (!in) { std::fprintf(stderr, "cannot open %s\n", path.c_str()); return 1; }
    std::string line;
    while (std::getline(in, line))
    {
        std::istringstream ls(line);
        std::string cmd;
        if (!(ls >> cmd) || cmd[0] == '#') continue;
        if (cmd == "update") { g.update(); continue; }
        std::string name;
        ls >> name;
        if (cmd == "xform" || cmd == "rotate")
        {
            const size_t id = g.find_node_id(name);
            if (id == SCENE_NODE_ROOT_ID) { std::fprintf(stderr, "no node %s\n", name.c_str()); return 1; }
            SL_Transform& t = g.mCurrentTransforms[id];
            if (cmd == "xform")
            {
                if (!read_transform(ls, t)) { std::fprintf(stderr, "bad line: %s\n", line.c_str()); return 1; }
            }
            else
            {
                float q[4];
                if (!(ls >> q[0] >> q[1] >> q[2] >> q[3])) { std::fprintf(stderr, "bad line: %s\n", line.c_str()); return 1; }
                t.rotate(SL_Quat{q[0], q[1], q[2], q[3]});
            }
        }
        else if (cmd == "inst" || cmd == "empty" || cmd == "mnode")
        {
            std::string parentName, source;
            ls >> parentName;
            if (cmd == "inst") ls >> source;
            std::vector<size_t> meshIds;
            if (cmd == "mnode")
            {
                size_t k = 0;
                ls >> k;
                for (size_t i = 0; i < k; ++i)
                {
                    size_t m = 0;
                    ls >> m;
                    meshIds.push_back(m);
                }
            }
            SL_Transform t;
            if (!read_transform(ls, t)) { std::fprintf(stderr, "bad line: %s\n", line.c_str()); return 1; }
            const size_t parent = node_of(g, parentName);
            if (parentName != "-" && parent == SCENE_NODE_ROOT_ID) { std::fprintf(stderr, "no node %s\n", parentName.c_str()); return 1; }
            size_t id;
            if (cmd == "empty")
                id = g.insert_empty_node(parent, name.c_str(), t);
            else if (cmd == "mnode")
                id = g.insert_mesh_node(parent, name.c_str(), meshIds.size(), meshIds.data(), t);
            else
            {
                const size_t src = g.find_node_id(source);
                if (src == SCENE_NODE_ROOT_ID || g.mNodes[src].type != NODE_TYPE_MESH)
                {
                    std::fprintf(stderr, "no mesh node %s\n", source.c_str());
                    return 1;
                }
                const std::vector<size_t> ids = g.mNodeMeshes[g.mNodes[src].dataId];
                id = g.insert_mesh_node(parent, name.c_str(), ids.size(), ids.data(), t);
            }
            if (id == SCENE_NODE_ROOT_ID) { std::fprintf(stderr, "insert failed: %s\n", line.c_str()); return 1; }
            /* a new node is dirty until the next update, like an edited one */
            g.mCurrentTransforms[id].set_dirty();
        }
        else
        {
            std::fprintf(stderr, "unknown command %s\n", cmd.c_str());
            return 1;
        }
    }
    g.update();
    return 0;
}

template <typename T>
bool write_raw(const std::string& path, const std::vector<T>& v)
{
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = v.empty() || std::fwrite(v.data(), sizeof(T), v.size(), f) == v.size();
    return std::fclose(f) == 0 && ok;
}

int dump(const SL_SceneFileData& d, SL_SceneGraph& g, const std::string& out)
{
    std::ofstream gr(out + "/graph.txt");
    for (size_t i = 0; i < g.mNodes.size(); ++i)
        gr << i << ' ' << (long long)(g.mNodeParentIds[i] == SCENE_NODE_ROOT_ID ? -1 : (long long)g.mNodeParentIds[i]) << ' '
           << (int)g.mNodes[i].type << ' ' << g.num_total_children(i) << ' ' << g.num_immediate_children(i) << ' '
           << g.mNodeNames[i] << '\n';
    std::ofstream ch(out + "/is_child.txt");
    for (size_t i = 0; i < g.mNodes.size(); ++i)
    {
        for (size_t j = 0; j < g.mNodes.size(); ++j) ch << (g.node_is_child(i, j) ? '1' : '0');
        ch << (g.node_is_child(i, SCENE_NODE_ROOT_ID) ? '1' : '0') << '\n';
    }
    std::vector<float> model;
    for (const SL_Mat4& m : g.mModelMatrices) model.insert(model.end(), m.m, m.m + 16);
    std::ofstream ms(out + "/meshes.txt");
    for (const SL_Mesh& m : g.mMeshes) ms << m.elementBegin << ' ' << m.elementEnd << ' ' << m.materialId << '\n';
    std::ofstream mt(out + "/materials.txt");
    mt << std::setprecision(9);
    for (const SL_Material& m : g.mMaterials)
    {
        for (float v : m.ambient) mt << v << ' ';
        for (float v : m.diffuse) mt << v << ' ';
        for (float v : m.specular) mt << v << ' ';
        mt << m.shininess << ' ' << (long long)(m.textures[SL_MATERIAL_TEXTURE_DIFFUSE] == SL_NO_RESOURCE ? -1 : (long long)m.textures[SL_MATERIAL_TEXTURE_DIFFUSE]) << '\n';
    }
    const SL_DrawList dl = sl_build_draw_list(g);
    std::ofstream dr(out + "/draws.txt");
    for (const SL_DrawItem& it : dl.items) dr << it.nodeId << ' ' << it.firstMesh << ' ' << it.numMeshes << ' ' << it.materialId << '\n';
    std::ofstream dm(out + "/draw_meshes.txt");
    for (const SL_Mesh& m : dl.meshes) dm << m.elementBegin << ' ' << m.elementEnd << ' ' << m.materialId << ' ' << m.mode << '\n';
    for (size_t i = 0; i < d.images.size(); ++i)
    {
        std::ofstream im(out + "/image" + std::to_string(i) + ".txt");
        im << d.images[i].w << ' ' << d.images[i].h << '\n';
        if (!write_raw(out + "/image" + std::to_string(i) + ".rgba", d.images[i].rgba)) return 1;
    }
    if (!write_raw(out + "/model.f32", model) || !write_raw(out + "/vertices.f32", d.vertices) ||
        !write_raw(out + "/indices.u32", d.indices))
        return 1;
    return 0;
}

double ms_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

#ifndef SL_SCENE_DUMP_ONLY
/* perspective(fovy, aspect, near, far), column-major, GL clip conventions */
void perspective(float fovyDeg, float aspect, float zn, float zf, float out[16])
{
    const double f = 1.0 / std::tan(fovyDeg * 3.14159265358979323846 / 360.0);
    std::memset(out, 0, 16 * sizeof(float));
    out[0] = (float)(f / aspect);
    out[5] = (float)f;
    out[10] = (float)((zf + zn) / (zn - zf));
    out[11] = -1.f;
    out[14] = (float)(2.0 * zf * zn / (zn - zf));
}

int render(const SL_SceneFileData& d, const std::string& cmds, int W, int H, const std::string& colorOut,
           const std::string& depthOut, int frames, double loadMs)
{
    SL_Context ctx;
    const size_t colorId = ctx.create_texture();
    const size_t depthId = ctx.create_texture();
    if (ctx.texture(colorId).init(SL_COLOR_RGBA_8U, (uint16_t)W, (uint16_t)H) != SL_OK ||
        ctx.texture(depthId).init(SL_COLOR_R_FLOAT, (uint16_t)W, (uint16_t)H) != SL_OK)
        return 4;
    const size_t fboId = ctx.create_framebuffer();
    SL_Framebuffer& fbo = ctx.framebuffer(fboId);
    if (fbo.reserve_color_buffers(1) != SL_OK || fbo.attach_color_buffer(0, SL_TextureView{colorId}) != SL_OK ||
        fbo.attach_depth_buffer(SL_TextureView{depthId}) != SL_OK || fbo.valid() != SL_OK)
        return 5;

    /* a white 4x4 texture for materials without a diffuse map */
    const size_t whiteId = ctx.create_texture();
    std::vector<uint8_t> white(4 * 4 * 4, 255);
    if (ctx.texture(whiteId).init(SL_COLOR_RGBA_8U, 4, 4) != SL_OK || ctx.texture(whiteId).set_texels(white.data(), white.size()) != SL_OK)
        return 6;

    const auto t0 = std::chrono::steady_clock::now();
    SL_SceneGraph g;
    if (sl_import_scene(ctx, d, g) == SCENE_NODE_ROOT_ID) return 7;
    const double importMs = ms_since(t0);
    if (apply_commands(g, cmds) != 0) return 8;

    const size_t uboId = ctx.create_ubo();
    const SL_VertexShader vs{2, SL_CULL_BACK_FACE, SL_PROGRAM_LAMBERT_TEXTURED};
    const SL_FragmentShader fs{2, 1, SL_BLEND_OFF, SL_DEPTH_TEST_LESS_THAN, SL_DEPTH_MASK_ON, SL_PROGRAM_LAMBERT_TEXTURED};
    const size_t shaderId = ctx.create_shader(vs, fs, uboId);
    if (shaderId == SL_INVALID_ID) return 9;

    sl_std_uniforms base{};
    perspective(60.f, (float)W / (float)H, 0.1f, 100.f, base.mvp);
    const float L[3] = {-0.4f, -1.0f, -0.6f};
    const float len = std::sqrt((L[0] * L[0] + L[1] * L[1]) + L[2] * L[2]);
    for (int i = 0; i < 3; ++i) base.light_dir[i] = L[i] / len;
    for (int i = 0; i < 4; ++i) base.tint[i] = 1.f;
    base.texture = (uint32_t)whiteId;
    base.instance_vbo = SL_NO_RESOURCE;
    {
        /* the base uniforms, for a program that redraws the frame elsewhere */
        std::vector<uint8_t> ub(sizeof(base));
        std::memcpy(ub.data(), &base, sizeof(base));
        if (!write_raw(colorOut + ".uniforms", ub)) return 14;
    }

    double buildMs = 0, submitMs = 0, frameMs = 0, bestMs = 1e30;
    size_t items = 0, tris = 0;
    for (int f = 0; f < frames; ++f)
    {
        const auto f0 = std::chrono::steady_clock::now();
        /* the per-frame host work: move one node, update, rebuild the list */
        if (frames > 1 && !g.mNodes.empty())
        {
            g.mCurrentTransforms[0].move(0.f, 0.f, 0.f);
            g.update();
        }
        const SL_DrawList dl = sl_build_draw_list(g);
        const double b = ms_since(f0);
        ctx.clear_framebuffer(fboId, 0, {0.0, 0.0, 0.0, 0.0}, 1.0);
        const auto s0 = std::chrono::steady_clock::now();
        if (sl_submit_draw_list(ctx, g, dl, shaderId, fboId, uboId, base) != SL_OK) return 10;
        const double s = ms_since(s0);
        if (ctx.finish() != SL_OK) return 11;
        const double fm = ms_since(f0);
        if (f > 0 || frames == 1) { buildMs += b; submitMs += s; frameMs += fm; bestMs = fm < bestMs ? fm : bestMs; }
        items = dl.items.size();
        tris = dl.num_triangles(g);
    }
    std::vector<uint8_t> color((size_t)W * H * 4);
    std::vector<float> depth((size_t)W * H);
    if (ctx.texture(colorId).read(color.data(), color.size()) != SL_OK ||
        ctx.texture(depthId).read(depth.data(), depth.size() * 4) != SL_OK)
        return 12;
    if (!write_raw(colorOut, color) || !write_raw(depthOut, depth)) return 13;
    if (frames > 1)
    {
        const int n = frames - 1;
        std::printf("{\"load_ms\": %.3f, \"import_ms\": %.3f, \"nodes\": %zu, \"draw_items\": %zu, \"triangles\": %zu, "
                    "\"build_ms\": %.4f, \"submit_ms\": %.4f, \"frame_ms\": %.4f, \"best_frame_ms\": %.4f, \"frames\": %d}\n",
                    loadMs, importMs, g.mNodes.size(), items, tris, buildMs / n, submitMs / n, frameMs / n, bestMs, n);
    }
    return 0;
}
#endif

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 5)
    {
        std::fprintf(stderr, "usage: %s dump scene.obj nodes.txt outdir\n"
                             "       %s render scene.obj nodes.txt W H color.out depth.out [frames]\n", argv[0], argv[0]);
        return 2;
    }
    const std::string mode = argv[1];
    const auto t0 = std::chrono::steady_clock::now();
    SL_SceneFileLoader loader;
    const int r = loader.load(argv[2]);
    if (r != SL_SCENE_LOAD_OK)
    {
        std::fprintf(stderr, "load failed (%d): %s\n", r, loader.error().c_str());
        return 3;
    }
    const double loadMs = ms_since(t0);
    if (mode == "dump")
    {
        SL_SceneGraph g;
        std::vector<uint32_t> texIds;
        for (size_t i = 0; i < loader.data().images.size(); ++i) texIds.push_back((uint32_t)i);
        sl_build_scene_graph(loader.data(), g, 0, texIds);
        if (apply_commands(g, argv[3]) != 0) return 8;
        return dump(loader.data(), g, argv[4]);
    }
#ifndef SL_SCENE_DUMP_ONLY
    if (mode == "render" && argc >= 8)
    {
        const int frames = argc > 8 ? std::atoi(argv[8]) : 1;
        return render(loader.data(), argv[3], std::atoi(argv[4]), std::atoi(argv[5]), argv[6], argv[7],
                      frames < 1 ? 1 : frames, loadMs);
    }
#endif
    std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
    return 2;
}
