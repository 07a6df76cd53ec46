// Host-side scene assembly: meshes + instances -> BVH2 traversal arrays.
//
// Follows Scene::build_mesh_accels / build_top_level_accel (render/scene.cpp:37-117):
//   - global triangle ids: tri_layout[m] = number of triangles of meshes 0..m-1 (scene.cpp:95-96)
//   - global texcoords and indices; indices i0..i2 offset by the mesh's first vertex, the
//     material (4th index) not offset (scene.cpp:75-93)
//   - mesh BVHs appended one after the other, layout[m] = index of mesh m's root (scene.cpp:52)
//   - top level built over instance bounds (gpu_adapter.cpp:164-185) and stored after the mesh
//     nodes, root = number of mesh nodes (scene.cpp:168-175)
//   - opacity masks: a buffer that starts with four opaque bytes; materials without a mask
//     use the 1x1 texel at offset 0 (core/mask.h:22-45)
#include <algorithm>
#include <atomic>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "bvh_build.h"
#include "bvh_cache.h"
#include "imba_hip.h"
#include "layout.h"
#include "status.h"

namespace imba {

struct MeshData {
    std::vector<float> vertices;     // 4 per vertex
    std::vector<uint32_t> indices;   // 4 per triangle
    std::vector<float> texcoords;    // 2 per vertex
    BBox bbox;
    TriSource source() const {
        TriSource s;
        s.vertices = vertices.data();
        s.indices = indices.data();
        s.tri_count = (int)(indices.size() / 4);
        return s;
    }
};

struct MaskData {
    int width = 0, height = 0;
    std::vector<uint8_t> texels;
};

struct MeshBvh {
    std::vector<ImbaNode> nodes;
    std::vector<ImbaVec4> tris;
    bool single_leaf = false;
};

// Loads mesh m's tree from its cache file (links relative to the mesh, global triangle ids).
bool load_cached(const std::string& path, int type, int tri_base, std::vector<char>& nodes, std::vector<ImbaVec4>& tris) {
    if (path.empty() || !cache_has_block(path.c_str(), type)) return false;
    return cache_load(path.c_str(), type, 0, 0, tri_base, nodes, tris) == IMBA_OK;
}

void build_mesh_bvh(const MeshData& mesh, int builder, int tri_base, MeshBvh& out) {
    const TriSource src = mesh.source();
    Bvh2MeshWriter w(out.nodes, out.tris, src, tri_base);
    auto on_node = [&](const BBox& p, int c, const BBox* b) { w.node(p, c, b); };
    auto on_leaf = [&](const BBox& bb, int c, const int* ids) { w.leaf(bb, c, ids); };
    if (builder == IMBA_BUILDER_SBVH) {
        // GpuMeshAdapter: SplitBvhBuilder<2>, leaf threshold 2, alpha 1e-5 (gpu_adapter.cpp:43,60)
        SplitBvhBuilder<2, GpuMeshCost> b;
        b.build(src, on_node, on_leaf, 2, 1e-5f);
    } else {
        // Binned SAH over the mesh (fast_bvh_builder.h:28-40) with the GPU mesh costs.
        BinnedSahBuilder<2, GpuMeshCost> b;
        b.build_mesh(src, on_node, on_leaf);
    }
    out.single_leaf = w.single_leaf;
}

struct MeshBvh4 {
    std::vector<ImbaNode4> nodes;
    std::vector<ImbaVec4> tris;
};

void build_mesh_bvh4(const MeshData& mesh, int builder, int tri_base, MeshBvh4& out) {
    const TriSource src = mesh.source();
    Bvh4MeshWriter w(out.nodes, out.tris, src, tri_base);
    auto on_node = [&](const BBox& p, int c, const BBox* b) { w.node(p, c, b); };
    auto on_leaf = [&](const BBox& bb, int c, const int* ids) { w.leaf(bb, c, ids); };
    if (builder == IMBA_BUILDER_SBVH) {
        // CpuMeshAdapter: SplitBvhBuilder<4>, leaf threshold 2, alpha 1e-4 (cpu_adapter.cpp:37,54)
        SplitBvhBuilder<4, CpuCost> b;
        b.build(src, on_node, on_leaf, 2, 1e-4f);
    } else {
        BinnedSahBuilder<4, CpuCost> b;
        b.build_mesh(src, on_node, on_leaf);
    }
}

// Runs fn(m) for m in [0, n) on up to `threads` threads.
template <typename Fn>
void parallel_over(int n, int threads, Fn&& fn) {
    const int nt = std::max(1, std::min(threads, n));
    std::atomic<int> next(0);
    auto work = [&]() {
        for (int m = next++; m < n; m = next++) fn(m);
    };
    if (nt == 1) {
        work();
        return;
    }
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; t++) pool.emplace_back(work);
    for (auto& th : pool) th.join();
}

// World-space instance boxes and centres from the mesh boxes (gpu_adapter.cpp:164-179,
// cpu_adapter.cpp:227-238): centre = M * c, extent = |M| * e.
void instance_bounds(const std::vector<MeshData>& meshes, const std::vector<InstanceDesc>& instances,
                     std::vector<BBox>& bounds, std::vector<float3>& centres) {
    const int ni = (int)instances.size();
    bounds.resize(ni);
    centres.resize(ni);
    for (int i = 0; i < ni; i++) {
        const InstanceDesc& d = instances[i];
        const BBox bb = meshes[d.mesh].bbox;
        const float4 c4((bb.max + bb.min) * 0.5f, 1.0f);
        const float4 e4((bb.max - bb.min) * 0.5f, 0.0f);
        float3 c, e;
        for (int r = 0; r < 3; r++) {
            const float4 row(d.mat[r][0], d.mat[r][1], d.mat[r][2], d.mat[r][3]);
            const float4 arow(std::fabs(row.x), std::fabs(row.y), std::fabs(row.z), std::fabs(row.w));
            c[r] = dot4(row, c4);
            e[r] = dot4(arow, e4);
        }
        centres[i] = c;
        bounds[i].min = c - e;
        bounds[i].max = c + e;
    }
}

}  // namespace imba

using namespace imba;

struct ImbaSceneBuild {
    std::vector<MeshData> meshes;
    std::vector<InstanceDesc> instances;
    std::vector<MaskData> masks;  // by material id

    // outputs
    bool built = false;
    int root = 0;
    std::vector<ImbaNode> nodes;
    std::vector<ImbaInstanceNode> inst_nodes;
    std::vector<ImbaVec4> tris;
    std::vector<int> indices;
    std::vector<ImbaVec2> texcoords;
    std::vector<ImbaTransparencyMask> mask_descs;
    std::vector<uint8_t> mask_buffer;
    std::vector<float> geom_normals;  // 3 per global triangle (object space)
    std::vector<int> tri_materials;
    std::vector<int> tri_layout;      // first global triangle id of each mesh
    std::vector<std::string> cache;   // per mesh: cache file or ""
    int cache_loaded = 0, cache_built = 0;

    // CPU layout (imba_build_cpu_layout)
    bool cpu_built = false;
    int cpu_root = 0;
    std::vector<ImbaNode4> cpu_nodes;
    std::vector<ImbaInstanceNode> cpu_inst_nodes;
    std::vector<ImbaVec4> cpu_tris;
};

extern "C" {

int imba_build_create(ImbaSceneBuild** out) {
    IMBA_CHECK_ARG(out, "out is NULL");
    *out = new ImbaSceneBuild();
    return IMBA_OK;
}

int imba_build_release(ImbaSceneBuild* b) {
    delete b;
    return IMBA_OK;
}

int imba_build_add_mesh(ImbaSceneBuild* b, const ImbaMeshDesc* m, int* mesh_id) {
    IMBA_CHECK_ARG(b && m, "NULL argument");
    IMBA_CHECK_ARG(m->vertices && m->indices, "mesh arrays are NULL");
    IMBA_CHECK_ARG(m->vertex_count > 0 && m->tri_count > 0, "empty mesh");
    for (int i = 0; i < 4 * m->tri_count; i++) {
        if ((i & 3) != 3 && m->indices[i] >= (uint32_t)m->vertex_count)
            return fail(IMBA_ERR_INVALID_ARG, "triangle %d references vertex %u of %d", i / 4, m->indices[i], m->vertex_count);
    }
    MeshData d;
    d.vertices.assign(m->vertices, m->vertices + 4 * (size_t)m->vertex_count);
    d.indices.assign(m->indices, m->indices + 4 * (size_t)m->tri_count);
    if (m->texcoords) d.texcoords.assign(m->texcoords, m->texcoords + 2 * (size_t)m->vertex_count);
    else d.texcoords.assign(2 * (size_t)m->vertex_count, 0.0f);
    d.bbox = BBox::empty();  // Mesh::compute_bounding_box (core/mesh.cpp:29-34)
    for (int i = 0; i < m->vertex_count; i++)
        d.bbox.extend(float3(d.vertices[4 * i], d.vertices[4 * i + 1], d.vertices[4 * i + 2]));
    b->meshes.push_back(std::move(d));
    if (mesh_id) *mesh_id = (int)b->meshes.size() - 1;
    b->built = false;
    return IMBA_OK;
}

int imba_build_add_instance(ImbaSceneBuild* b, int mesh_id, const float* mat16, int* inst_id) {
    IMBA_CHECK_ARG(b && mat16, "NULL argument");
    IMBA_CHECK_ARG(mesh_id >= 0 && mesh_id < (int)b->meshes.size(), "mesh_id out of range");
    InstanceDesc d;
    d.mesh = mesh_id;
    Mat4 m;
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) m.m[r][c] = mat16[4 * r + c];
    const Mat4 inv = invert(m);  // Mesh::Instance (core/mesh.h:19-26)
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 4; c++) { d.mat[r][c] = m.m[r][c]; d.inv_mat[r][c] = inv.m[r][c]; }
    b->instances.push_back(d);
    if (inst_id) *inst_id = (int)b->instances.size() - 1;
    b->built = false;
    return IMBA_OK;
}

int imba_build_set_mask(ImbaSceneBuild* b, int material, int width, int height, const uint8_t* texels) {
    IMBA_CHECK_ARG(b && texels, "NULL argument");
    IMBA_CHECK_ARG(material >= 0 && width > 0 && height > 0, "bad mask size");
    if ((int)b->masks.size() <= material) b->masks.resize(material + 1);
    b->masks[material].width = width;
    b->masks[material].height = height;
    b->masks[material].texels.assign(texels, texels + (size_t)width * height);
    b->built = false;
    return IMBA_OK;
}

int imba_build_run(ImbaSceneBuild* b, int builder, int threads) {
    IMBA_CHECK_ARG(b, "NULL argument");
    IMBA_CHECK_ARG(builder == IMBA_BUILDER_SBVH || builder == IMBA_BUILDER_BINNED_SAH, "unknown builder");
    IMBA_CHECK_ARG(!b->meshes.empty(), "no meshes");
    if (b->instances.empty()) {
        // No instances given: one identity instance per mesh (frontend/build_scene.cpp:539-543).
        const float id[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
        for (int m = 0; m < (int)b->meshes.size(); m++) imba_build_add_instance(b, m, id, nullptr);
    }
    const int nm = (int)b->meshes.size();

    // Global triangle ids, texcoords and indices.
    std::vector<int> tri_layout(nm);
    b->texcoords.clear();
    b->indices.clear();
    b->geom_normals.clear();
    b->tri_materials.clear();
    int max_material = 0;
    {
        int tri_offset = 0;
        for (int m = 0; m < nm; m++) {
            const MeshData& mesh = b->meshes[m];
            const int voff = (int)b->texcoords.size();
            const int nv = (int)(mesh.vertices.size() / 4);
            for (int i = 0; i < nv; i++) b->texcoords.push_back({mesh.texcoords[2 * i], mesh.texcoords[2 * i + 1]});
            for (size_t i = 0; i < mesh.indices.size(); i++)
                b->indices.push_back((int)mesh.indices[i] + ((i % 4 == 3) ? 0 : voff));
            const TriSource src = mesh.source();
            for (int t = 0; t < src.tri_count; t++) {
                const Tri tri = src.triangle(t);
                const float3 n = normalize(cross(tri.v1 - tri.v0, tri.v2 - tri.v0));
                b->geom_normals.push_back(n.x);
                b->geom_normals.push_back(n.y);
                b->geom_normals.push_back(n.z);
                const int mat = (int)mesh.indices[4 * t + 3];
                b->tri_materials.push_back(mat);
                max_material = std::max(max_material, mat);
            }
            tri_layout[m] = tri_offset;
            tri_offset += src.tri_count;
        }
    }
    b->tri_layout = tri_layout;

    // Mesh BVHs, in parallel over meshes; each one is built with mesh-local links.
    std::vector<MeshBvh> per_mesh(nm);
    std::atomic<int> loaded(0);
    b->cache.resize(nm);
    parallel_over(nm, threads, [&](int m) {
        std::vector<char> raw;
        if (load_cached(b->cache[m], IMBA_CACHE_BVH2, tri_layout[m], raw, per_mesh[m].tris)) {
            per_mesh[m].nodes.resize(raw.size() / sizeof(ImbaNode));
            std::memcpy(per_mesh[m].nodes.data(), raw.data(), raw.size());
            per_mesh[m].single_leaf = !per_mesh[m].nodes.empty() && per_mesh[m].nodes[0].right == IMBA_DUMMY_CHILD;
            loaded++;
            return;
        }
        build_mesh_bvh(b->meshes[m], builder, tri_layout[m], per_mesh[m]);
        if (!b->cache[m].empty())
            cache_store(b->cache[m].c_str(), IMBA_CACHE_BVH2, per_mesh[m].nodes.data(), (int)per_mesh[m].nodes.size(),
                        per_mesh[m].tris.data(), (int)per_mesh[m].tris.size(), 0, 0, tri_layout[m]);
    });
    b->cache_loaded = loaded;
    b->cache_built = nm - loaded;

    // Concatenate with global links (what one writer appending all meshes would produce).
    b->nodes.clear();
    b->tris.clear();
    std::vector<int> layout(nm);
    for (int m = 0; m < nm; m++) {
        const int node_base = (int)b->nodes.size();
        const int tri_base = (int)b->tris.size();
        layout[m] = node_base;
        MeshBvh& mb = per_mesh[m];
        for (size_t i = 0; i < mb.nodes.size(); i++) {
            ImbaNode n = mb.nodes[i];
            const bool dummy = mb.single_leaf && i == 0;
            auto fix = [&](int c) { return c >= 0 ? c + node_base : ~(~c + tri_base); };
            n.left = fix(n.left);
            if (!dummy) n.right = fix(n.right);
            b->nodes.push_back(n);
        }
        b->tris.insert(b->tris.end(), mb.tris.begin(), mb.tris.end());
        std::vector<ImbaNode>().swap(mb.nodes);
        std::vector<ImbaVec4>().swap(mb.tris);
    }
    const int node_count = (int)b->nodes.size();

    // Top level (GpuTopLevelAdapter::build_accel, gpu_adapter.cpp:164-185).
    const int ni = (int)b->instances.size();
    std::vector<BBox> bounds;
    std::vector<float3> centres;
    instance_bounds(b->meshes, b->instances, bounds, centres);
    std::vector<ImbaNode> top;
    b->inst_nodes.clear();
    {
        Bvh2TopWriter w(top, b->inst_nodes, b->instances, layout, node_count);
        BinnedSahBuilder<2, GpuTopCost> tb;
        tb.build(bounds.data(), centres.data(), ni,
                 [&](const BBox& p, int c, const BBox* bx) { w.node(p, c, bx); },
                 [&](const BBox& bb, int c, const int* ids) { w.leaf(bb, c, ids); });
    }
    b->nodes.insert(b->nodes.end(), top.begin(), top.end());
    b->root = node_count;

    // Masks: one descriptor per material id.
    b->mask_buffer.assign(4, 1);
    b->mask_descs.clear();
    const int nmat = std::max(max_material + 1, (int)b->masks.size());
    for (int m = 0; m < nmat; m++) {
        ImbaTransparencyMask desc{1, 1, 0};
        if (m < (int)b->masks.size() && b->masks[m].width > 0) {
            const MaskData& md = b->masks[m];
            desc.width = md.width;
            desc.height = md.height;
            desc.offset = (int)b->mask_buffer.size();
            for (uint8_t t : md.texels) b->mask_buffer.push_back(t != 0);
        }
        b->mask_descs.push_back(desc);
    }
    b->built = true;
    b->cpu_built = false;
    return IMBA_OK;
}

int imba_build_cpu_layout(ImbaSceneBuild* b, int builder, int threads) {
    IMBA_CHECK_ARG(b, "NULL argument");
    IMBA_CHECK_ARG(builder == IMBA_BUILDER_SBVH || builder == IMBA_BUILDER_BINNED_SAH, "unknown builder");
    if (!b->built) return fail(IMBA_ERR_INVALID_ARG, "imba_build_run has not been called");
    const int nm = (int)b->meshes.size();
    std::vector<MeshBvh4> per_mesh(nm);
    std::atomic<int> loaded(0);
    b->cache.resize(nm);
    parallel_over(nm, threads, [&](int m) {
        std::vector<char> raw;
        if (load_cached(b->cache[m], IMBA_CACHE_BVH4, b->tri_layout[m], raw, per_mesh[m].tris)) {
            per_mesh[m].nodes.resize(raw.size() / sizeof(ImbaNode4));
            std::memcpy(per_mesh[m].nodes.data(), raw.data(), raw.size());
            loaded++;
            return;
        }
        build_mesh_bvh4(b->meshes[m], builder, b->tri_layout[m], per_mesh[m]);
        if (!b->cache[m].empty())
            cache_store(b->cache[m].c_str(), IMBA_CACHE_BVH4, per_mesh[m].nodes.data(), (int)per_mesh[m].nodes.size(),
                        per_mesh[m].tris.data(), (int)per_mesh[m].tris.size(), 0, 0, b->tri_layout[m]);
    });
    b->cache_loaded = loaded;
    b->cache_built = nm - loaded;

    // Concatenate with global links: node children > 0 shift by the mesh's first node, leaf
    // links by its first Vec4; 0 (unused slot) stays.
    b->cpu_nodes.clear();
    b->cpu_tris.clear();
    std::vector<int> layout(nm);
    for (int m = 0; m < nm; m++) {
        const int node_base = (int)b->cpu_nodes.size();
        const int tri_base = (int)b->cpu_tris.size();
        layout[m] = node_base;
        for (ImbaNode4 n : per_mesh[m].nodes) {
            for (int j = 0; j < 4; j++) {
                const int c = n.children[j];
                n.children[j] = c > 0 ? c + node_base : c < 0 ? ~(~c + tri_base) : 0;
            }
            b->cpu_nodes.push_back(n);
        }
        b->cpu_tris.insert(b->cpu_tris.end(), per_mesh[m].tris.begin(), per_mesh[m].tris.end());
        std::vector<ImbaNode4>().swap(per_mesh[m].nodes);
        std::vector<ImbaVec4>().swap(per_mesh[m].tris);
    }
    const int node_count = (int)b->cpu_nodes.size();

    // Top level (CpuTopLevelAdapter::build_accel, cpu_adapter.cpp:227-247).
    std::vector<BBox> bounds;
    std::vector<float3> centres;
    instance_bounds(b->meshes, b->instances, bounds, centres);
    std::vector<ImbaNode4> top;
    b->cpu_inst_nodes.clear();
    {
        Bvh4TopWriter w(top, b->cpu_inst_nodes, b->instances, layout, node_count);
        BinnedSahBuilder<4, CpuCost> tb;
        tb.build(bounds.data(), centres.data(), (int)b->instances.size(),
                 [&](const BBox& p, int c, const BBox* bx) { w.node(p, c, bx); },
                 [&](const BBox& bb, int c, const int* ids) { w.leaf(bb, c, ids); });
    }
    b->cpu_nodes.insert(b->cpu_nodes.end(), top.begin(), top.end());
    b->cpu_root = node_count;
    b->cpu_built = true;
    return IMBA_OK;
}

int imba_build_set_cache(ImbaSceneBuild* b, int mesh_id, const char* path) {
    IMBA_CHECK_ARG(b, "NULL argument");
    IMBA_CHECK_ARG(mesh_id >= 0 && mesh_id < (int)b->meshes.size(), "mesh_id out of range");
    if ((int)b->cache.size() < (int)b->meshes.size()) b->cache.resize(b->meshes.size());
    b->cache[mesh_id] = path ? path : "";
    return IMBA_OK;
}

int imba_build_cache_stats(const ImbaSceneBuild* b, int* loaded, int* built) {
    IMBA_CHECK_ARG(b && loaded && built, "NULL argument");
    *loaded = b->cache_loaded;
    *built = b->cache_built;
    return IMBA_OK;
}

int imba_build_get_cpu_desc(const ImbaSceneBuild* b, ImbaCpuSceneDesc* out) {
    IMBA_CHECK_ARG(b && out, "NULL argument");
    if (!b->cpu_built) return fail(IMBA_ERR_INVALID_ARG, "imba_build_cpu_layout has not been called");
    out->root = b->cpu_root;
    out->nodes = b->cpu_nodes.data();
    out->node_count = (int)b->cpu_nodes.size();
    out->instances = b->cpu_inst_nodes.data();
    out->instance_count = (int)b->cpu_inst_nodes.size();
    out->tris = b->cpu_tris.data();
    out->tri_vec4_count = (int)b->cpu_tris.size();
    out->indices = b->indices.data();
    out->index_count = (int)b->indices.size();
    out->texcoords = b->texcoords.data();
    out->texcoord_count = (int)b->texcoords.size();
    out->masks = b->mask_descs.data();
    out->mask_count = (int)b->mask_descs.size();
    out->mask_buffer = b->mask_buffer.data();
    out->mask_buffer_size = (int)b->mask_buffer.size();
    return IMBA_OK;
}

int imba_build_get_desc(const ImbaSceneBuild* b, ImbaSceneDesc* out) {
    IMBA_CHECK_ARG(b && out, "NULL argument");
    if (!b->built) return fail(IMBA_ERR_INVALID_ARG, "imba_build_run has not been called");
    out->root = b->root;
    out->nodes = b->nodes.data();
    out->node_count = (int)b->nodes.size();
    out->instances = b->inst_nodes.data();
    out->instance_count = (int)b->inst_nodes.size();
    out->tris = b->tris.data();
    out->tri_vec4_count = (int)b->tris.size();
    out->indices = b->indices.data();
    out->index_count = (int)b->indices.size();
    out->texcoords = b->texcoords.data();
    out->texcoord_count = (int)b->texcoords.size();
    out->masks = b->mask_descs.data();
    out->mask_count = (int)b->mask_descs.size();
    out->mask_buffer = b->mask_buffer.data();
    out->mask_buffer_size = (int)b->mask_buffer.size();
    return IMBA_OK;
}

int imba_build_get_tri_attrs(const ImbaSceneBuild* b, const float** geom_normals, const int** materials, int* tri_count) {
    IMBA_CHECK_ARG(b && geom_normals && materials && tri_count, "NULL argument");
    if (!b->built) return fail(IMBA_ERR_INVALID_ARG, "imba_build_run has not been called");
    *geom_normals = b->geom_normals.data();
    *materials = b->tri_materials.data();
    *tri_count = (int)b->tri_materials.size();
    return IMBA_OK;
}

}  // extern "C"
