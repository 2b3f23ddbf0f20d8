/*
 * SL_SceneFileLoader.hpp -- scene files into a scene graph and per-frame
 * draw submission, over the C++ facade.
 *
 * The reference loads scenes through Assimp (softlight/src/
 * SL_SceneFileLoader.cpp:240-330) into one vertex buffer and one index
 * buffer per vertex layout, a mesh per (object, material) pair, a material
 * table, and a node tree inserted parent-first.  Assimp is not part of this
 * library; the loader here reads Wavefront OBJ (+ MTL, + binary PPM diffuse
 * maps) itself and produces the same shape of result:
 *
 *   - one interleaved vertex buffer, 32 bytes a vertex: position f32x3,
 *     uv f32x2, normal f32x3 -- the reference's standard vertex in its
 *     attribute order (POSITION, TEXTURE, NORMAL; SL_Geometry.hpp:75-131);
 *   - one index buffer whose element type is the narrowest that holds the
 *     index count (sl_required_index_type, as the reference picks it,
 *     SL_SceneFileLoader.cpp:432);
 *   - one SL_Mesh per run of faces of one object with one material,
 *     RENDER_MODE_INDEXED_TRIANGLES over [elementBegin, elementEnd);
 *   - nodes: an empty root node named after the file, and one mesh node per
 *     OBJ object ('o' / 'g') as its child, identity transforms.
 *
 * Vertices are de-duplicated on their (v, vt, vn) index triple, numbered in
 * first-use order.  Polygons are split into a fan around their first vertex.
 * A face corner without 'vt' gets uv (0, 0); a face without 'vn' gets its
 * geometric normal normalize(cross(p1 - p0, p2 - p0)) on corners that have
 * none (a corner that appears with and without a normal is two vertices).
 *
 * Loading is host-only (`load`); `sl_import_scene` uploads the result into a
 * context and fills a scene graph; `sl_submit_draw_list` records a frame.
 */
#ifndef SOFTLIGHT_AMD_SL_SCENE_FILE_LOADER_HPP
#define SOFTLIGHT_AMD_SL_SCENE_FILE_LOADER_HPP

#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

#include "SL_Context.hpp"
#include "SL_SceneGraph.hpp"

/* load() results (the reference returns bool; the codes say which step failed) */
enum SL_SceneLoadStatus
{
    SL_SCENE_LOAD_OK = 0,
    SL_SCENE_LOAD_NO_FILE = -1,
    SL_SCENE_LOAD_BAD_SYNTAX = -2,
    SL_SCENE_LOAD_BAD_INDEX = -3,
    SL_SCENE_LOAD_BAD_MATERIAL = -4,
    SL_SCENE_LOAD_BAD_IMAGE = -5,
    SL_SCENE_LOAD_EMPTY = -6
};

/* SL_IndexBuffer / SL_SceneFileLoader.cpp:432: the index type for a count */
inline int sl_required_index_type(size_t numIndices) noexcept
{
    return numIndices <= 0xFFu ? VERTEX_DATA_BYTE : numIndices <= 0xFFFFu ? VERTEX_DATA_SHORT : VERTEX_DATA_INT;
}

struct SL_SceneImage
{
    uint16_t w = 0, h = 0;
    std::vector<uint8_t> rgba;                   /* w*h RGBA8, row 0 first */
};

/* Everything a file holds, on the host. */
struct SL_SceneFileData
{
    static constexpr uint32_t kStride = 32;      /* position 0, uv 12, normal 20 */
    std::vector<float> vertices;                 /* 8 floats a vertex */
    std::vector<uint32_t> indices;
    struct Mesh { uint32_t begin, end, material; };
    std::vector<Mesh> meshes;
    struct Material { SL_Material m; int diffuseImage; };   /* image index or -1 */
    std::vector<Material> materials;
    std::vector<std::string> materialNames;
    std::vector<SL_SceneImage> images;
    struct Node { std::string name; size_t parent; std::vector<size_t> meshes; };
    std::vector<Node> nodes;
    size_t num_vertices() const noexcept { return vertices.size() / 8; }
};

class SL_SceneFileLoader
{
  public:
    const SL_SceneFileData& data() const noexcept { return mData; }
    const std::string& error() const noexcept { return mError; }

    int load(const std::string& path)
    {
        mData = SL_SceneFileData{};
        mError.clear();
        mDefaultMaterial = -1;   // a material index of an earlier file means nothing in this one
        std::ifstream in(path);
        if (!in) return fail(SL_SCENE_LOAD_NO_FILE, "cannot open " + path);
        const std::string dir = dir_of(path);
        std::vector<float> P, T, N;              /* 3, 2, 3 floats each */
        std::map<std::tuple<long, long, long, uint32_t, uint32_t, uint32_t>, uint32_t> vmap;
        long curMat = -1;
        int curObj = -1;
        std::string line;
        size_t lineNo = 0;
        const std::string base = base_name(path);
        mData.nodes.push_back(SL_SceneFileData::Node{base, SCENE_NODE_ROOT_ID, {}});
        auto open_mesh = [&]() {
            if (curObj < 0)
            {
                mData.nodes.push_back(SL_SceneFileData::Node{base + ":default", 0, {}});
                curObj = (int)mData.nodes.size() - 1;
            }
            std::vector<size_t>& ms = mData.nodes[(size_t)curObj].meshes;
            const uint32_t mat = material_for(curMat);
            const uint32_t at = (uint32_t)mData.indices.size();
            if (!ms.empty())
            {
                SL_SceneFileData::Mesh& last = mData.meshes[ms.back()];
                if (last.material == mat && last.end == at) return;   /* still the open run */
            }
            mData.meshes.push_back(SL_SceneFileData::Mesh{at, at, mat});
            ms.push_back(mData.meshes.size() - 1);
        };
        while (std::getline(in, line))
        {
            ++lineNo;
            if (!line.empty() && line.back() == '\r') line.pop_back();
            std::istringstream ls(line);
            std::string tag;
            if (!(ls >> tag) || tag[0] == '#') continue;
            if (tag == "v" || tag == "vn")
            {
                float a, b, c;
                if (!(ls >> a >> b >> c)) return fail(SL_SCENE_LOAD_BAD_SYNTAX, at_line(lineNo, tag));
                std::vector<float>& dst = tag == "v" ? P : N;
                dst.push_back(a); dst.push_back(b); dst.push_back(c);
            }
            else if (tag == "vt")
            {
                float a, b;
                if (!(ls >> a >> b)) return fail(SL_SCENE_LOAD_BAD_SYNTAX, at_line(lineNo, tag));
                T.push_back(a); T.push_back(b);
            }
            else if (tag == "o" || tag == "g")
            {
                std::string name;
                std::getline(ls >> std::ws, name);
                mData.nodes.push_back(SL_SceneFileData::Node{name.empty() ? base + ":" + tag : name, 0, {}});
                curObj = (int)mData.nodes.size() - 1;
            }
            else if (tag == "mtllib")
            {
                std::string file;
                std::getline(ls >> std::ws, file);
                const int r = load_mtl(dir + file);
                if (r != SL_SCENE_LOAD_OK) return r;
            }
            else if (tag == "usemtl")
            {
                std::string name;
                std::getline(ls >> std::ws, name);
                curMat = -1;
                for (size_t i = 0; i < mData.materialNames.size(); ++i)
                    if (mData.materialNames[i] == name) curMat = (long)i;
                if (curMat < 0) return fail(SL_SCENE_LOAD_BAD_MATERIAL, at_line(lineNo, "unknown material " + name));
            }
            else if (tag == "f")
            {
                std::vector<std::tuple<long, long, long>> corners;
                std::string c;
                while (ls >> c)
                {
                    long vi = 0, ti = 0, ni = 0;
                    if (!parse_corner(c, vi, ti, ni)) return fail(SL_SCENE_LOAD_BAD_SYNTAX, at_line(lineNo, c));
                    if (!resolve(vi, P.size() / 3, false) || !resolve(ti, T.size() / 2, true) ||
                        !resolve(ni, N.size() / 3, true))
                        return fail(SL_SCENE_LOAD_BAD_INDEX, at_line(lineNo, c));
                    corners.emplace_back(vi, ti, ni);
                }
                if (corners.size() < 3) return fail(SL_SCENE_LOAD_BAD_SYNTAX, at_line(lineNo, "face"));
                open_mesh();
                for (size_t k = 1; k + 1 < corners.size(); ++k)
                {
                    const std::tuple<long, long, long> tri[3] = {corners[0], corners[k], corners[k + 1]};
                    float fn[3];
                    face_normal(P, tri, fn);
                    for (const auto& t : tri)
                    {
                        const long vi = std::get<0>(t), ti = std::get<1>(t), ni = std::get<2>(t);
                        /* a corner without a normal takes this face's, so the face normal is part of its key */
                        uint32_t nb[3] = {0, 0, 0};
                        if (ni < 0) std::memcpy(nb, fn, 12);
                        const auto key = std::make_tuple(vi, ti, ni, nb[0], nb[1], nb[2]);
                        auto it = vmap.find(key);
                        uint32_t id;
                        if (it != vmap.end())
                            id = it->second;
                        else
                        {
                            id = (uint32_t)mData.num_vertices();
                            const float* p = &P[3 * (size_t)vi];
                            const float uv[2] = {ti >= 0 ? T[2 * (size_t)ti] : 0.f, ti >= 0 ? T[2 * (size_t)ti + 1] : 0.f};
                            const float* n = ni >= 0 ? &N[3 * (size_t)ni] : fn;
                            mData.vertices.insert(mData.vertices.end(), {p[0], p[1], p[2], uv[0], uv[1], n[0], n[1], n[2]});
                            vmap.emplace(key, id);
                        }
                        mData.indices.push_back(id);
                    }
                }
                mData.meshes[mData.nodes[(size_t)curObj].meshes.back()].end = (uint32_t)mData.indices.size();
            }
            /* s, l, p and other statements are ignored */
        }
        /* drop object nodes that received no faces */
        std::vector<SL_SceneFileData::Node> kept;
        kept.push_back(mData.nodes[0]);
        for (size_t i = 1; i < mData.nodes.size(); ++i)
            if (!mData.nodes[i].meshes.empty()) kept.push_back(mData.nodes[i]);
        mData.nodes.swap(kept);
        if (mData.indices.empty()) return fail(SL_SCENE_LOAD_EMPTY, "no faces in " + path);
        return SL_SCENE_LOAD_OK;
    }

  private:
    SL_SceneFileData mData;
    std::string mError;
    long mDefaultMaterial = -1;

    int fail(int code, const std::string& msg) { mError = msg; return code; }
    static std::string at_line(size_t n, const std::string& what) { return "line " + std::to_string(n) + ": " + what; }
    static std::string dir_of(const std::string& p)
    {
        const size_t s = p.find_last_of('/');
        return s == std::string::npos ? std::string() : p.substr(0, s + 1);
    }
    static std::string base_name(const std::string& p)
    {
        const size_t s = p.find_last_of('/');
        return s == std::string::npos ? p : p.substr(s + 1);
    }

    /* OBJ indices are 1-based; negative ones count back from the last element.
     * 0 means absent (-1 after resolving) where allowed. */
    static bool resolve(long& i, size_t count, bool optional)
    {
        if (i == 0) { i = -1; return optional; }
        const long r = i > 0 ? i - 1 : (long)count + i;
        if (r < 0 || (size_t)r >= count) return false;
        i = r;
        return true;
    }

    static bool parse_corner(const std::string& c, long& v, long& t, long& n)
    {
        const char* s = c.c_str();
        char* e = nullptr;
        v = std::strtol(s, &e, 10);
        if (e == s) return false;
        if (*e == 0) return true;
        if (*e != '/') return false;
        s = e + 1;
        if (*s != '/')
        {
            t = std::strtol(s, &e, 10);
            if (e == s) return false;
            if (*e == 0) return true;
            if (*e != '/') return false;
            s = e;
        }
        s += 1;
        n = std::strtol(s, &e, 10);
        return e != s && *e == 0;
    }

    static void face_normal(const std::vector<float>& P, const std::tuple<long, long, long> tri[3], float out[3])
    {
        const float* a = &P[3 * (size_t)std::get<0>(tri[0])];
        const float* b = &P[3 * (size_t)std::get<0>(tri[1])];
        const float* c = &P[3 * (size_t)std::get<0>(tri[2])];
        const float u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        const float v[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        const float n[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
        const float len = std::sqrt((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2]);
        const float inv = len > 0.f ? 1.f / len : 0.f;
        for (int i = 0; i < 3; ++i) out[i] = n[i] * inv;
    }
    uint32_t material_for(long m)
    {
        if (m >= 0) return (uint32_t)m;
        if (mDefaultMaterial < 0)
        {
            mData.materials.push_back(SL_SceneFileData::Material{sl_default_material(), -1});
            mData.materialNames.push_back("");
            mDefaultMaterial = (long)mData.materials.size() - 1;
        }
        return (uint32_t)mDefaultMaterial;
    }

    /* MTL: newmtl, Ka, Kd, Ks, Ns, d (alpha of the diffuse colour), map_Kd */
    int load_mtl(const std::string& path)
    {
        std::ifstream in(path);
        if (!in) return fail(SL_SCENE_LOAD_BAD_MATERIAL, "cannot open " + path);
        const std::string dir = dir_of(path);
        std::string line;
        SL_SceneFileData::Material* cur = nullptr;
        while (std::getline(in, line))
        {
            if (!line.empty() && line.back() == '\r') line.pop_back();
            std::istringstream ls(line);
            std::string tag;
            if (!(ls >> tag) || tag[0] == '#') continue;
            if (tag == "newmtl")
            {
                std::string name;
                std::getline(ls >> std::ws, name);
                mData.materials.push_back(SL_SceneFileData::Material{sl_default_material(), -1});
                mData.materialNames.push_back(name);
                cur = &mData.materials.back();
                continue;
            }
            if (!cur) return fail(SL_SCENE_LOAD_BAD_MATERIAL, path + ": statement before newmtl");
            if (tag == "Ka" || tag == "Kd" || tag == "Ks")
            {
                float* dst = tag == "Ka" ? cur->m.ambient : tag == "Kd" ? cur->m.diffuse : cur->m.specular;
                if (!(ls >> dst[0] >> dst[1] >> dst[2])) return fail(SL_SCENE_LOAD_BAD_MATERIAL, path + ": " + line);
            }
            else if (tag == "Ns")
            {
                if (!(ls >> cur->m.shininess)) return fail(SL_SCENE_LOAD_BAD_MATERIAL, path + ": " + line);
            }
            else if (tag == "d")
            {
                if (!(ls >> cur->m.diffuse[3])) return fail(SL_SCENE_LOAD_BAD_MATERIAL, path + ": " + line);
            }
            else if (tag == "map_Kd")
            {
                std::string file;
                std::getline(ls >> std::ws, file);
                SL_SceneImage img;
                if (!load_ppm(dir + file, img)) return fail(SL_SCENE_LOAD_BAD_IMAGE, "cannot read image " + dir + file);
                mData.images.push_back(std::move(img));
                cur->diffuseImage = (int)mData.images.size() - 1;
            }
        }
        return SL_SCENE_LOAD_OK;
    }

    /* binary PPM (P6, maxval 255) -> RGBA8 with alpha 255 */
    static bool load_ppm(const std::string& path, SL_SceneImage& img)
    {
        std::ifstream in(path, std::ios::binary);
        if (!in) return false;
        std::string magic;
        long w = 0, h = 0, mx = 0;
        auto skip = [&]() {
            for (;;)
            {
                int ch = in.peek();
                if (ch == '#') { std::string c; std::getline(in, c); }
                else if (ch != EOF && std::isspace(ch)) in.get();
                else break;
            }
        };
        in >> magic;
        skip(); in >> w; skip(); in >> h; skip(); in >> mx;
        if (!in || magic != "P6" || w <= 0 || h <= 0 || w > 65535 || h > 65535 || mx != 255) return false;
        in.get();
        std::vector<uint8_t> rgb((size_t)w * (size_t)h * 3);
        if (!in.read(reinterpret_cast<char*>(rgb.data()), (std::streamsize)rgb.size())) return false;
        img.w = (uint16_t)w;
        img.h = (uint16_t)h;
        img.rgba.resize((size_t)w * (size_t)h * 4);
        for (size_t i = 0, n = (size_t)w * (size_t)h; i < n; ++i)
        {
            img.rgba[4 * i + 0] = rgb[3 * i + 0];
            img.rgba[4 * i + 1] = rgb[3 * i + 1];
            img.rgba[4 * i + 2] = rgb[3 * i + 2];
            img.rgba[4 * i + 3] = 255;
        }
        return true;
    }
};

/*-----------------------------------------------------------------------------
 * The graph part of an import, without a context: materials and meshes are
 * appended to the graph's tables (meshes drawing through vertex array
 * `vaoId`, image i bound as texture texIds[i]) and the file's nodes are
 * appended under `parentId`, parent-first.  Returns the file's root node.
-----------------------------------------------------------------------------*/
inline size_t sl_build_scene_graph(const SL_SceneFileData& d, SL_SceneGraph& g, size_t vaoId,
                                   const std::vector<uint32_t>& texIds, size_t parentId = SCENE_NODE_ROOT_ID)
{
    const uint32_t matBase = (uint32_t)g.mMaterials.size();
    for (const SL_SceneFileData::Material& m : d.materials)
    {
        SL_Material mm = m.m;
        if (m.diffuseImage >= 0 && (size_t)m.diffuseImage < texIds.size())
            mm.textures[SL_MATERIAL_TEXTURE_DIFFUSE] = texIds[(size_t)m.diffuseImage];
        g.insert_material(mm);
    }
    const size_t meshBase = g.mMeshes.size();
    for (const SL_SceneFileData::Mesh& m : d.meshes)
        g.insert_mesh(SL_Mesh{vaoId, m.begin, m.end, RENDER_MODE_INDEXED_TRIANGLES, matBase + m.material});

    std::vector<size_t> nodeIds(d.nodes.size(), SCENE_NODE_ROOT_ID);
    for (size_t i = 0; i < d.nodes.size(); ++i)
    {
        const SL_SceneFileData::Node& n = d.nodes[i];
        const size_t parent = n.parent == SCENE_NODE_ROOT_ID ? parentId : nodeIds[n.parent];
        const SL_Transform t;
        if (n.meshes.empty())
            nodeIds[i] = g.insert_empty_node(parent, n.name.c_str(), t);
        else
        {
            std::vector<size_t> ids;
            for (size_t m : n.meshes) ids.push_back(meshBase + m);
            nodeIds[i] = g.insert_mesh_node(parent, n.name.c_str(), ids.size(), ids.data(), t);
        }
    }
    return nodeIds.empty() ? SCENE_NODE_ROOT_ID : nodeIds[0];
}

/*-----------------------------------------------------------------------------
 * Upload into a context (SL_SceneGraph::import + the loader's buffer
 * creation, SL_SceneFileLoader.cpp:430-560): one vbo, one ibo and one vao
 * for the file, a texture per image, materials and meshes appended to the
 * graph's tables, the nodes appended under `parentId`.  Returns the id of
 * the file's root node, or SCENE_NODE_ROOT_ID on a context error.
-----------------------------------------------------------------------------*/
inline size_t sl_import_scene(SL_Context& ctx, const SL_SceneFileData& d, SL_SceneGraph& g,
                              size_t parentId = SCENE_NODE_ROOT_ID)
{
    const size_t vbo = ctx.create_vbo();
    if (ctx.vbo(vbo).init(d.vertices.size() * sizeof(float), d.vertices.data()) != SL_OK) return SCENE_NODE_ROOT_ID;
    const size_t ibo = ctx.create_ibo();
    const int itype = sl_required_index_type(d.indices.size());
    std::vector<uint8_t> ib(d.indices.size() * (itype == VERTEX_DATA_BYTE ? 1 : itype == VERTEX_DATA_SHORT ? 2 : 4));
    for (size_t i = 0; i < d.indices.size(); ++i)
    {
        const uint32_t v = d.indices[i];
        if (itype == VERTEX_DATA_BYTE) ib[i] = (uint8_t)v;
        else if (itype == VERTEX_DATA_SHORT) { const uint16_t s = (uint16_t)v; std::memcpy(&ib[2 * i], &s, 2); }
        else std::memcpy(&ib[4 * i], &v, 4);
    }
    if (ctx.ibo(ibo).init((uint32_t)d.indices.size(), itype, ib.data()) != SL_OK) return SCENE_NODE_ROOT_ID;
    const size_t vao = ctx.create_vao();
    SL_VertexArray& va = ctx.vao(vao);
    va.set_vertex_buffer(vbo);
    va.set_index_buffer(ibo);
    /* bindings as the built-in programs read them: 0 position, 1 normal, 2 uv */
    if (va.set_num_bindings(3) != 3) return SCENE_NODE_ROOT_ID;     /* returns the binding count */
    va.set_binding(0, 0, SL_SceneFileData::kStride, VERTEX_DIMENSION_3, VERTEX_DATA_FLOAT);
    va.set_binding(1, 20, SL_SceneFileData::kStride, VERTEX_DIMENSION_3, VERTEX_DATA_FLOAT);
    va.set_binding(2, 12, SL_SceneFileData::kStride, VERTEX_DIMENSION_2, VERTEX_DATA_FLOAT);

    std::vector<uint32_t> texIds;
    for (const SL_SceneImage& img : d.images)
    {
        const size_t t = ctx.create_texture();
        if (ctx.texture(t).init(SL_COLOR_RGBA_8U, img.w, img.h, 1) != SL_OK ||
            ctx.texture(t).set_texels(img.rgba.data(), img.rgba.size()) != SL_OK)
            return SCENE_NODE_ROOT_ID;
        texIds.push_back((uint32_t)t);
    }
    return sl_build_scene_graph(d, g, vao, texIds, parentId);
}

/*-----------------------------------------------------------------------------
 * One frame of a draw list for the built-in Lambert programs: per item the
 * uniforms are `base` with the node's model matrix, the material's ambient
 * and diffuse colours (diffuse alpha into tint.w) and its diffuse texture
 * (base.texture when it has none); they are assigned to `uboId`, which the
 * shader reads, and the item is recorded as one draw_multiple.  The context
 * runs the frame at the next finish() or readback.  Returns the first
 * error code, or SL_OK.
-----------------------------------------------------------------------------*/
inline int sl_submit_draw_list(SL_Context& ctx, const SL_SceneGraph& g, const SL_DrawList& dl, size_t shaderId,
                               size_t fboId, size_t uboId, const sl_std_uniforms& base)
{
    for (const SL_DrawItem& it : dl.items)
    {
        sl_std_uniforms u = base;
        std::memcpy(u.model, g.mModelMatrices[it.nodeId].m, sizeof(u.model));
        const SL_Material& m = g.mMaterials[it.materialId];
        for (int i = 0; i < 3; ++i) { u.ambient[i] = m.ambient[i]; u.diffuse[i] = m.diffuse[i]; }
        u.tint[3] = base.tint[3] * m.diffuse[3];
        if (m.textures[SL_MATERIAL_TEXTURE_DIFFUSE] != SL_NO_RESOURCE) u.texture = m.textures[SL_MATERIAL_TEXTURE_DIFFUSE];
        ctx.ubo(uboId).assign(u);
        const int r = sl_draw_multiple(ctx.native(), &dl.meshes[it.firstMesh], it.numMeshes, shaderId, fboId);
        if (r != SL_OK) return r;
    }
    return SL_OK;
}

#endif /* SOFTLIGHT_AMD_SL_SCENE_FILE_LOADER_HPP */
