/*
 * SL_SceneGraph.hpp -- host-side scene graph for the MI355X library: node
 * hierarchy, transforms, meshes and materials, and the draw list a frame is
 * submitted from.
 *
 * Keeps the reference's data layout (softlight/include/softlight/
 * SL_SceneGraph.hpp:215-420): parallel arrays indexed by node id
 * (mNodeParentIds, mNodes, mNodeNames, mBaseTransforms, mCurrentTransforms,
 * mModelMatrices), mesh-node data indexed by SL_SceneNode::dataId
 * (mNumNodeMeshes, mNodeMeshes), and the shared mesh and material tables
 * (mMeshes, mMaterials).  A parent always has a lower index than its
 * children and a node's descendants follow it contiguously; insertion
 * appends, so callers insert a subtree depth-first (what a file loader does).
 *
 * What the reference does per frame (softlight/tests/sl_instancing_test.cpp:
 * 233-280: for every mesh node, for every sub-mesh, set the uniforms, draw)
 * is split here into `sl_build_draw_list`, which walks the graph once and
 * folds consecutive sub-meshes of a node that share a material into one
 * item, and `sl_submit_draw_list` (SL_SceneFileLoader.hpp), which records
 * every item as a deferred draw on the context's stream -- the uniforms are
 * snapshotted per draw -- and synchronises once per frame.
 *
 * Not kept: cameras, bones, animation, reparent/copy/delete of nodes.
 */
#ifndef SOFTLIGHT_AMD_SL_SCENE_GRAPH_HPP
#define SOFTLIGHT_AMD_SL_SCENE_GRAPH_HPP

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../sl_gpu.h"

typedef sl_mesh SL_Mesh;

/*-----------------------------------------------------------------------------
 * Column-major 4x4 matrices and quaternions.  The arithmetic order is fixed
 * (no contraction) so every build of a host program computes the same
 * model matrices: c[j][i] = ((a[0][i]*b[j][0] + a[1][i]*b[j][1])
 * + a[2][i]*b[j][2]) + a[3][i]*b[j][3], i.e. (a*b)*v == a*(b*v).
-----------------------------------------------------------------------------*/
struct SL_Mat4
{
    float m[16];                                 /* m[4*col + row] */
    float& at(int col, int row) { return m[4 * col + row]; }
    float at(int col, int row) const { return m[4 * col + row]; }
};

inline SL_Mat4 sl_mat4_identity() noexcept
{
    SL_Mat4 r{};
    r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.f;
    return r;
}

#pragma STDC FP_CONTRACT OFF
inline SL_Mat4 sl_mat4_mul(const SL_Mat4& a, const SL_Mat4& b) noexcept
{
    SL_Mat4 c;
    for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 4; ++i)
        {
            float s = a.at(0, i) * b.at(j, 0);
            s = s + a.at(1, i) * b.at(j, 1);
            s = s + a.at(2, i) * b.at(j, 2);
            s = s + a.at(3, i) * b.at(j, 3);
            c.at(j, i) = s;
        }
    return c;
}

struct SL_Quat { float x, y, z, w; };

inline SL_Quat sl_quat_identity() noexcept { return SL_Quat{0.f, 0.f, 0.f, 1.f}; }

/* Hamilton product: rotating by (a * b) rotates by b first, then a. */
inline SL_Quat sl_quat_mul(const SL_Quat& a, const SL_Quat& b) noexcept
{
    return SL_Quat{
        ((a.w * b.x + a.x * b.w) + a.y * b.z) - a.z * b.y,
        ((a.w * b.y - a.x * b.z) + a.y * b.w) + a.z * b.x,
        ((a.w * b.z + a.x * b.y) - a.y * b.x) + a.z * b.w,
        ((a.w * b.w - a.x * b.x) - a.y * b.y) - a.z * b.z};
}

inline SL_Quat sl_quat_normalize(const SL_Quat& q) noexcept
{
    const float inv = 1.f / std::sqrt(((q.x * q.x + q.y * q.y) + q.z * q.z) + q.w * q.w);
    return SL_Quat{q.x * inv, q.y * inv, q.z * inv, q.w * inv};
}

/* Unit quaternion for a rotation of `angle` radians about `axis` (normalised here). */
inline SL_Quat sl_quat_from_axis_angle(float ax, float ay, float az, float angle) noexcept
{
    const float len = std::sqrt((ax * ax + ay * ay) + az * az);
    const float s = std::sin(0.5f * angle) / len;
    return SL_Quat{ax * s, ay * s, az * s, std::cos(0.5f * angle)};
}

/* Rotation matrix of a unit quaternion (the standard form, e.g. Shoemake,
 * "Animating rotation with quaternion curves", SIGGRAPH 1985). */
inline SL_Mat4 sl_quat_to_mat4(const SL_Quat& q) noexcept
{
    const float xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z;
    const float xy = q.x * q.y, xz = q.x * q.z, yz = q.y * q.z;
    const float wx = q.w * q.x, wy = q.w * q.y, wz = q.w * q.z;
    SL_Mat4 r = sl_mat4_identity();
    r.at(0, 0) = 1.f - 2.f * (yy + zz);
    r.at(0, 1) = 2.f * (xy + wz);
    r.at(0, 2) = 2.f * (xz - wy);
    r.at(1, 0) = 2.f * (xy - wz);
    r.at(1, 1) = 1.f - 2.f * (xx + zz);
    r.at(1, 2) = 2.f * (yz + wx);
    r.at(2, 0) = 2.f * (xz + wy);
    r.at(2, 1) = 2.f * (yz - wx);
    r.at(2, 2) = 1.f - 2.f * (xx + yy);
    return r;
}
#pragma STDC FP_CONTRACT ON

/*-----------------------------------------------------------------------------
 * SL_Transform (softlight/include/softlight/SL_Transform.hpp, model type):
 * position, per-axis scale and orientation, and the cached matrix they give.
 * apply_transform(true) builds (T*S)*R -- rotate, then scale, then translate
 * (SL_Transform.cpp:319-327, get_rst_matrix :408-416); false builds
 * R*(T*S) (get_str_matrix :423-431).  apply_pre_transform(P) is
 * P * apply_transform() (:346-350): a child's matrix is its parent's matrix
 * times its own.
-----------------------------------------------------------------------------*/
class SL_Transform
{
  public:
    float mPosition[3] = {0.f, 0.f, 0.f};
    float mScaling[3] = {1.f, 1.f, 1.f};
    SL_Quat mOrientation = sl_quat_identity();
    SL_Mat4 mModelMat = sl_mat4_identity();
    bool mDirty = false;

    bool is_dirty() const noexcept { return mDirty; }
    void set_dirty() noexcept { mDirty = true; }
    void set_clean() noexcept { mDirty = false; }

    /* setters of SL_Transform.cpp:198-310; each marks the transform dirty */
    void position(float x, float y, float z) noexcept { mPosition[0] = x; mPosition[1] = y; mPosition[2] = z; mDirty = true; }
    void move(float dx, float dy, float dz) noexcept { mPosition[0] += dx; mPosition[1] += dy; mPosition[2] += dz; mDirty = true; }
    void scaling(float x, float y, float z) noexcept { mScaling[0] = x; mScaling[1] = y; mScaling[2] = z; mDirty = true; }
    void scale(float x, float y, float z) noexcept { mScaling[0] *= x; mScaling[1] *= y; mScaling[2] *= z; mDirty = true; }
    void orientation(const SL_Quat& q) noexcept { mOrientation = q; mDirty = true; }
    /* SL_Transform::rotate(quat) (SL_Transform.cpp:266-270): orientation = normalize(orientation * delta) */
    void rotate(const SL_Quat& delta) noexcept { mOrientation = sl_quat_normalize(sl_quat_mul(mOrientation, delta)); mDirty = true; }

    SL_Mat4 get_rst_matrix() const noexcept { return sl_mat4_mul(translate_scale(), sl_quat_to_mat4(mOrientation)); }
    SL_Mat4 get_str_matrix() const noexcept { return sl_mat4_mul(sl_quat_to_mat4(mOrientation), translate_scale()); }

    void apply_transform(bool useRST = true) noexcept
    {
        mModelMat = useRST ? get_rst_matrix() : get_str_matrix();
        set_clean();
    }
    void apply_pre_transform(const SL_Mat4& parent, bool useRST = true) noexcept
    {
        apply_transform(useRST);
        mModelMat = sl_mat4_mul(parent, mModelMat);
    }
    const SL_Mat4& transform() const noexcept { return mModelMat; }

  private:
    SL_Mat4 translate_scale() const noexcept
    {
        SL_Mat4 r = sl_mat4_identity();
        r.at(0, 0) = mScaling[0];
        r.at(1, 1) = mScaling[1];
        r.at(2, 2) = mScaling[2];
        r.at(3, 0) = mPosition[0];
        r.at(3, 1) = mPosition[1];
        r.at(3, 2) = mPosition[2];
        return r;
    }
};

/*-----------------------------------------------------------------------------
 * Nodes and materials (SL_SceneNode.hpp:12-62, SL_Material.hpp:19-63)
-----------------------------------------------------------------------------*/
enum SL_SceneNodeProp : size_t { SCENE_NODE_ROOT_ID = (size_t)-1 };

enum SL_SceneNodeType : size_t
{
    NODE_TYPE_EMPTY,
    NODE_TYPE_MESH,
    NODE_TYPE_CAMERA,
    NODE_TYPE_BONE
};

struct SL_SceneNode
{
    SL_SceneNodeType type;
    size_t dataId;
};

enum SL_MaterialProperty : uint32_t
{
    SL_MATERIAL_TEXTURE_AMBIENT = 0,
    SL_MATERIAL_TEXTURE_DIFFUSE = 1,
    SL_MATERIAL_TEXTURE_NORMAL = 2,
    SL_MATERIAL_TEXTURE_HEIGHT = 3,
    SL_MATERIAL_TEXTURE_SPECULAR = 4,
    SL_MATERIAL_TEXTURE_OPACITY = 5,
    SL_MATERIAL_TEXTURE_MISC0 = 6,
    SL_MATERIAL_TEXTURE_MISC1 = 7,
    SL_MATERIAL_MAX_TEXTURES = 8
};

/* Texture slots hold context texture handles (the reference holds
 * SL_Texture pointers); SL_NO_RESOURCE marks an empty slot. */
struct SL_Material
{
    uint32_t textures[SL_MATERIAL_MAX_TEXTURES];
    float ambient[4];
    float diffuse[4];
    float specular[4];
    float shininess;
};

inline SL_Material sl_default_material() noexcept
{
    SL_Material m{};
    for (uint32_t& t : m.textures) t = SL_NO_RESOURCE;
    for (int i = 0; i < 4; ++i) { m.ambient[i] = 0.f; m.diffuse[i] = 1.f; m.specular[i] = 0.f; }
    m.ambient[3] = 1.f;
    m.shininess = 0.f;
    return m;
}

/*-----------------------------------------------------------------------------
 * The graph
-----------------------------------------------------------------------------*/
class SL_SceneGraph
{
  public:
    std::vector<size_t> mNodeParentIds;
    std::vector<SL_SceneNode> mNodes;
    std::vector<std::string> mNodeNames;
    std::vector<SL_Mat4> mBaseTransforms;
    std::vector<SL_Transform> mCurrentTransforms;
    std::vector<SL_Mat4> mModelMatrices;
    std::vector<size_t> mNumNodeMeshes;
    std::vector<std::vector<size_t>> mNodeMeshes;
    std::vector<SL_Mesh> mMeshes;
    std::vector<SL_Material> mMaterials;

    /* SL_SceneGraph::clear_node_data */
    void clear_node_data() noexcept
    {
        mNodeParentIds.clear(); mNodes.clear(); mNodeNames.clear(); mBaseTransforms.clear();
        mCurrentTransforms.clear(); mModelMatrices.clear(); mNumNodeMeshes.clear(); mNodeMeshes.clear();
    }
    /* SL_SceneGraph::terminate */
    void terminate() noexcept { clear_node_data(); mMeshes.clear(); mMaterials.clear(); }

    size_t insert_mesh(const SL_Mesh& m) { mMeshes.push_back(m); return mMeshes.size() - 1; }
    size_t insert_material(const SL_Material& m) { mMaterials.push_back(m); return mMaterials.size() - 1; }

    /* SL_SceneGraph::insert_empty_node (SL_SceneGraph.cpp:1180-1205): appended,
     * its base, current and model matrices all the transform's matrix.
     * Returns SCENE_NODE_ROOT_ID for a parent id that does not exist. */
    size_t insert_empty_node(size_t parentId, const char* name, const SL_Transform& t)
    {
        if (parentId != SCENE_NODE_ROOT_ID && parentId >= mNodes.size()) return SCENE_NODE_ROOT_ID;
        mNodeParentIds.push_back(parentId);
        mNodes.push_back(SL_SceneNode{NODE_TYPE_EMPTY, ~(size_t)0});
        mNodeNames.emplace_back(name ? name : "");
        mBaseTransforms.push_back(t.transform());
        mCurrentTransforms.push_back(t);
        mModelMatrices.push_back(t.transform());
        return mNodes.size() - 1;
    }

    /* SL_SceneGraph::insert_mesh_node (SL_SceneGraph.cpp:1230-1260) */
    size_t insert_mesh_node(size_t parentId, const char* name, size_t numSubMeshes, const size_t* subMeshIds,
                            const SL_Transform& t)
    {
        if (numSubMeshes == 0) return SCENE_NODE_ROOT_ID;
        for (size_t i = 0; i < numSubMeshes; ++i)
            if (subMeshIds[i] >= mMeshes.size()) return SCENE_NODE_ROOT_ID;
        const size_t nodeId = insert_empty_node(parentId, name, t);
        if (nodeId == SCENE_NODE_ROOT_ID) return nodeId;
        mNodes[nodeId].type = NODE_TYPE_MESH;
        mNodes[nodeId].dataId = mNodeMeshes.size();
        mNodeMeshes.emplace_back(subMeshIds, subMeshIds + numSubMeshes);
        mNumNodeMeshes.push_back(numSubMeshes);
        return nodeId;
    }

    /* SL_SceneGraph::find_node_id (SL_SceneGraph.cpp:925-938): the last node of that name */
    size_t find_node_id(const std::string& name) const noexcept
    {
        for (size_t i = mNodeNames.size(); i-- > 0;)
            if (mNodeNames[i] == name) return i;
        return SCENE_NODE_ROOT_ID;
    }

    /* SL_SceneGraph::num_total_children (SL_SceneGraph.cpp:945-969) */
    size_t num_total_children(size_t nodeIndex) const noexcept
    {
        if (nodeIndex == SCENE_NODE_ROOT_ID) return mNodes.size();
        size_t n = 0;
        for (size_t c = nodeIndex + 1; c < mNodeParentIds.size(); ++c)
        {
            const size_t p = mNodeParentIds[c];
            if (p < nodeIndex || p == SCENE_NODE_ROOT_ID) break;
            ++n;
        }
        return n;
    }

    /* SL_SceneGraph::num_immediate_children (SL_SceneGraph.cpp:976-1004) */
    size_t num_immediate_children(size_t nodeIndex) const noexcept
    {
        if (nodeIndex == SCENE_NODE_ROOT_ID) return mNodes.size();
        size_t n = 0;
        for (size_t c = nodeIndex + 1; c < mNodeParentIds.size(); ++c)
        {
            const size_t p = mNodeParentIds[c];
            if (p == SCENE_NODE_ROOT_ID || p < nodeIndex) break;
            n += (p == nodeIndex);
        }
        return n;
    }

    /* SL_SceneGraph::node_is_child (SL_SceneGraph.cpp:1011-1045): whether
     * parentId is an ancestor of nodeIndex.  As in the reference, the
     * `parentId >= nodeIndex` rejection comes first, so the root id
     * (the largest size_t) is never reported as an ancestor. */
    bool node_is_child(size_t nodeIndex, size_t parentId) const noexcept
    {
        if (nodeIndex == SCENE_NODE_ROOT_ID || parentId >= nodeIndex || nodeIndex >= mNodes.size()) return false;
        for (size_t p = mNodeParentIds[nodeIndex]; p != SCENE_NODE_ROOT_ID; p = mNodeParentIds[p])
        {
            if (p == parentId) return true;
            if (p < parentId) return false;
        }
        return false;
    }

    /* SL_SceneGraph::update (SL_SceneGraph.cpp:259-299): a dirty node and
     * every node after it up to the first whose parent precedes it are
     * recomputed parent-first; a node's model matrix is its transform's. */
    void update() noexcept
    {
        const size_t n = mCurrentTransforms.size();
        for (size_t i = 0; i < n; ++i)
        {
            if (!mCurrentTransforms[i].is_dirty()) continue;
            update_node_transform(i);
            for (size_t j = i + 1; j < n; ++j)
            {
                if (mNodeParentIds[j] < i)
                {
                    i = j - 1;
                    break;
                }
                update_node_transform(j);
            }
        }
    }

  private:
    /* SL_SceneGraph::update_node_transform (SL_SceneGraph.cpp:223-252), without bones */
    void update_node_transform(size_t id) noexcept
    {
        SL_Transform& t = mCurrentTransforms[id];
        const size_t p = mNodeParentIds[id];
        if (p != SCENE_NODE_ROOT_ID)
            t.apply_pre_transform(mCurrentTransforms[p].transform());
        else
            t.apply_transform();
        mModelMatrices[id] = t.transform();
    }
};

/*-----------------------------------------------------------------------------
 * Draw lists
-----------------------------------------------------------------------------*/
/* One draw of a frame: `numMeshes` meshes starting at `firstMesh` in the
 * list's mesh array, all of node `nodeId` and material `materialId`. */
struct SL_DrawItem
{
    size_t nodeId;
    size_t firstMesh;
    size_t numMeshes;
    uint32_t materialId;
};

struct SL_DrawList
{
    std::vector<SL_Mesh> meshes;                 /* the meshes, in submission order */
    std::vector<SL_DrawItem> items;
    size_t num_triangles(const SL_SceneGraph&) const noexcept
    {
        size_t t = 0;
        for (const SL_Mesh& m : meshes) t += (m.elementEnd - m.elementBegin) / 3;
        return t;
    }
};

/* Walks the graph in node order -- the order the reference's render loop
 * draws in -- and emits each mesh node's sub-meshes, folding a run of
 * consecutive sub-meshes with one material and one render mode into a
 * single item (a draw_multiple call).  Call update() first. */
inline SL_DrawList sl_build_draw_list(const SL_SceneGraph& g)
{
    SL_DrawList out;
    for (size_t nodeId = 0; nodeId < g.mNodes.size(); ++nodeId)
    {
        const SL_SceneNode& n = g.mNodes[nodeId];
        if (n.type != NODE_TYPE_MESH) continue;
        const std::vector<size_t>& ids = g.mNodeMeshes[n.dataId];
        for (size_t k = 0; k < ids.size(); ++k)
        {
            const SL_Mesh& m = g.mMeshes[ids[k]];
            SL_DrawItem* last = out.items.empty() ? nullptr : &out.items.back();
            if (last && last->nodeId == nodeId && last->materialId == m.materialId &&
                out.meshes.back().mode == m.mode)
                ++last->numMeshes;
            else
                out.items.push_back(SL_DrawItem{nodeId, out.meshes.size(), 1, m.materialId});
            out.meshes.push_back(m);
        }
    }
    return out;
}

#endif /* SOFTLIGHT_AMD_SL_SCENE_GRAPH_HPP */
