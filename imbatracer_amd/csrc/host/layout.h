// Layout writers: turn a builder's node/leaf stream into the BVH2 traversal arrays.
//
//   Bvh2MeshWriter   <- GpuMeshAdapter::NodeWriter / LeafWriter      core/gpu_adapter.cpp:62-142
//   Bvh2TopWriter    <- GpuTopLevelAdapter::NodeWriter / LeafWriter  core/gpu_adapter.cpp:203-296
//   dummy parent     <- fill_dummy_parent                            core/gpu_adapter.cpp:14-31
//
// Nodes are appended in emission (depth-first) order.  A pending-link stack records which
// child slot of which node the next node or leaf fills; child 1 is pushed before child 0 so
// the next item written becomes child 0.  A tree that is a single leaf gets a dummy parent
// whose right child is IMBA_DUMMY_CHILD with a degenerate box at the origin.
#pragma once

#include <algorithm>
#include <vector>

#include "imba_hip.h"
#include "vec.h"

namespace imba {

inline float bits_as_float(uint32_t b) {
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

inline ImbaBBox to_abi(const BBox& b) {
    ImbaBBox r;
    r.lo_x = b.min.x; r.lo_y = b.min.y; r.lo_z = b.min.z;
    r.hi_x = b.max.x; r.hi_y = b.max.y; r.hi_z = b.max.z;
    return r;
}

inline void set_child(ImbaNode& n, int slot, int value) {
    if (slot == 0) n.left = value; else n.right = value;
}

inline ImbaNode dummy_parent(const BBox& leaf_bb, int leaf_link) {
    ImbaNode n{};
    n.left = leaf_link;
    n.right = IMBA_DUMMY_CHILD;
    n.left_bb = to_abi(leaf_bb);
    n.right_bb.lo_x = 0.0f; n.right_bb.lo_y = 0.0f; n.right_bb.lo_z = 0.0f;
    n.right_bb.hi_x = -0.0f; n.right_bb.hi_y = -0.0f; n.right_bb.hi_z = -0.0f;
    return n;
}

struct PendingLink {
    int parent, slot;
};

// Writes one mesh BVH.  Node links are relative to the start of `nodes`; leaf links are
// ~(index of the first Vec4 in `tris`) relative to the start of `tris`; triangle ids are
// `tri_base` + local triangle index.
struct Bvh2MeshWriter {
    std::vector<ImbaNode>& nodes;
    std::vector<ImbaVec4>& tris;
    const TriSource& mesh;
    int tri_base;
    std::vector<PendingLink> links;
    bool single_leaf = false;

    Bvh2MeshWriter(std::vector<ImbaNode>& n, std::vector<ImbaVec4>& t, const TriSource& m, int base)
        : nodes(n), tris(t), mesh(m), tri_base(base) {}

    void node(const BBox&, int count, const BBox* child) {
        const int i = (int)nodes.size();
        nodes.emplace_back();
        if (!links.empty()) {
            const PendingLink l = links.back();
            links.pop_back();
            set_child(nodes[l.parent], l.slot, i);
        }
        (void)count;  // always 2 for a binary tree
        nodes[i].left_bb = to_abi(child[0]);
        nodes[i].right_bb = to_abi(child[1]);
        links.push_back({i, 1});
        links.push_back({i, 0});
    }

    void leaf(const BBox& bb, int count, const int* ids) {
        const int link = ~(int)tris.size();
        if (links.empty()) {
            nodes.push_back(dummy_parent(bb, link));
            single_leaf = true;
        } else {
            const PendingLink l = links.back();
            links.pop_back();
            set_child(nodes[l.parent], l.slot, link);
        }
        for (int i = 0; i < count; i++) {
            const Tri t = mesh.triangle(ids[i]);
            tris.push_back({t.v0.x, t.v0.y, t.v0.z, 0.0f});
            tris.push_back({t.v1.x, t.v1.y, t.v1.z, bits_as_float((uint32_t)(ids[i] + tri_base))});
            tris.push_back({t.v2.x, t.v2.y, t.v2.z, 0.0f});
        }
        tris.back().w = bits_as_float(IMBA_LEAF_END_BITS);
    }
};

struct InstanceDesc {
    int mesh;
    float mat[3][4];      // object -> world (rows 0..2 of the 4x4)
    float inv_mat[3][4];  // world -> object
};

// Writes the top level: nodes are appended to `nodes` with links offset by `root_offset`;
// leaves append instance records terminated by pad = {-1,-1}.
struct Bvh2TopWriter {
    std::vector<ImbaNode>& nodes;
    std::vector<ImbaInstanceNode>& inst_nodes;
    const std::vector<InstanceDesc>& instances;
    const std::vector<int>& mesh_roots;
    int root_offset;
    std::vector<PendingLink> links;

    Bvh2TopWriter(std::vector<ImbaNode>& n, std::vector<ImbaInstanceNode>& in, const std::vector<InstanceDesc>& inst,
                  const std::vector<int>& roots, int offset)
        : nodes(n), inst_nodes(in), instances(inst), mesh_roots(roots), root_offset(offset) {}

    void node(const BBox&, int, const BBox* child) {
        const int i = (int)nodes.size();
        nodes.emplace_back();
        if (!links.empty()) {
            const PendingLink l = links.back();
            links.pop_back();
            set_child(nodes[l.parent], l.slot, i + root_offset);
        }
        nodes[i].left_bb = to_abi(child[0]);
        nodes[i].right_bb = to_abi(child[1]);
        links.push_back({i, 1});
        links.push_back({i, 0});
    }

    void leaf(const BBox& bb, int count, const int* ids) {
        const int link = ~(int)inst_nodes.size();
        if (links.empty()) {
            nodes.push_back(dummy_parent(bb, link));
        } else {
            const PendingLink l = links.back();
            links.pop_back();
            set_child(nodes[l.parent], l.slot, link);
        }
        for (int j = 0; j < count; j++) {
            const InstanceDesc& d = instances[ids[j]];
            ImbaInstanceNode in{};
            for (int r = 0; r < 3; r++) {
                in.transf[r].x = d.inv_mat[r][0];
                in.transf[r].y = d.inv_mat[r][1];
                in.transf[r].z = d.inv_mat[r][2];
                in.transf[r].w = d.inv_mat[r][3];
            }
            in.id = ids[j];
            in.next = mesh_roots[d.mesh];
            in.pad[0] = 0;
            in.pad[1] = 0;
            inst_nodes.push_back(in);
        }
        inst_nodes.back().pad[0] = -1;
        inst_nodes.back().pad[1] = -1;
    }
};

// ---------------------------------------------------------------------------------------
// CPU layout (BVH4 nodes, 4-triangle packets): CpuMeshAdapter / CpuTopLevelAdapter,
// core/cpu_adapter.cpp:14-206, 221-366.  Child j of a node is filled by the j-th item written
// after it (slots pushed from count-1 down to 0).

inline void set_box4(ImbaNode4& n, int j, const BBox& b) {
    n.min_x[j] = b.min.x; n.min_y[j] = b.min.y; n.min_z[j] = b.min.z;
    n.max_x[j] = b.max.x; n.max_y[j] = b.max.y; n.max_z[j] = b.max.z;
}

// Inner node: `count` child boxes, unused slots empty ([FLT_MAX, -FLT_MAX], child 0).
inline void write_node4(ImbaNode4& n, int count, const BBox* child, std::vector<PendingLink>& links, int self) {
    for (int j = count - 1; j >= 0; j--) {
        set_box4(n, j, child[j]);
        links.push_back({self, j});
    }
    for (int j = 3; j >= count; j--) {
        set_box4(n, j, BBox::empty());
        n.children[j] = 0;
    }
}

// Dummy parent of a single-leaf tree (fill_dummy_parent, cpu_adapter.cpp:14-25): slot 0 holds
// the leaf, slot 1 child 0; the remaining slots stay zero.
inline ImbaNode4 dummy_parent4(const BBox& leaf_bb, int leaf_link) {
    ImbaNode4 n{};
    n.children[0] = leaf_link;
    n.children[1] = 0;
    set_box4(n, 0, leaf_bb);
    return n;
}

// One 13-Vec4 packet of up to four triangles (cpu_adapter.cpp:129-201); n uses the fma cross
// product of the traversal arithmetic.
inline void write_packet4(std::vector<ImbaVec4>& tris, const TriSource& mesh, const int* ids, int c, int tri_base) {
    float raw[13][4];
    for (int j = 0; j < 4; j++) {
        if (j < c) {
            const Tri t = mesh.triangle(ids[j]);
            const float3 e1 = t.v0 - t.v1;
            const float3 e2 = t.v2 - t.v0;
            const float n[3] = {std::fma(e1.y, e2.z, -(e1.z * e2.y)), std::fma(e1.z, e2.x, -(e1.x * e2.z)),
                                std::fma(e1.x, e2.y, -(e1.y * e2.x))};
            const float v[12] = {t.v0.x, t.v0.y, t.v0.z, e1.x, e1.y, e1.z, e2.x, e2.y, e2.z, n[0], n[1], n[2]};
            for (int k = 0; k < 12; k++) raw[k][j] = v[k];
            raw[12][j] = bits_as_float((uint32_t)(ids[j] + tri_base));
        } else {
            for (int k = 0; k < 12; k++) raw[k][j] = 0.0f;
            raw[12][j] = bits_as_float(0x80000000u);
        }
    }
    for (int k = 0; k < 13; k++) tris.push_back({raw[k][0], raw[k][1], raw[k][2], raw[k][3]});
}

struct Bvh4MeshWriter {
    std::vector<ImbaNode4>& nodes;
    std::vector<ImbaVec4>& tris;
    const TriSource& mesh;
    int tri_base;
    std::vector<PendingLink> links;
    bool single_leaf = false;

    Bvh4MeshWriter(std::vector<ImbaNode4>& n, std::vector<ImbaVec4>& t, const TriSource& m, int base)
        : nodes(n), tris(t), mesh(m), tri_base(base) {}

    void node(const BBox&, int count, const BBox* child) {
        const int i = (int)nodes.size();
        nodes.emplace_back();
        if (!links.empty()) {
            const PendingLink l = links.back();
            links.pop_back();
            nodes[l.parent].children[l.slot] = i;
        }
        write_node4(nodes[i], count, child, links, i);
    }

    void leaf(const BBox& bb, int count, const int* ids) {
        const int link = ~(int)tris.size();
        if (links.empty()) {
            nodes.push_back(dummy_parent4(bb, link));
            single_leaf = true;
        } else {
            const PendingLink l = links.back();
            links.pop_back();
            nodes[l.parent].children[l.slot] = link;
        }
        for (int i = 0; i < count; i += 4) write_packet4(tris, mesh, ids + i, std::min(4, count - i), tri_base);
        const float s = bits_as_float(0x80000000u);
        tris.push_back({s, s, s, s});
    }
};

struct Bvh4TopWriter {
    std::vector<ImbaNode4>& nodes;
    std::vector<ImbaInstanceNode>& inst_nodes;
    const std::vector<InstanceDesc>& instances;
    const std::vector<int>& mesh_roots;
    int root_offset;
    std::vector<PendingLink> links;

    Bvh4TopWriter(std::vector<ImbaNode4>& n, std::vector<ImbaInstanceNode>& in, const std::vector<InstanceDesc>& inst,
                  const std::vector<int>& roots, int offset)
        : nodes(n), inst_nodes(in), instances(inst), mesh_roots(roots), root_offset(offset) {}

    void node(const BBox&, int count, const BBox* child) {
        const int i = (int)nodes.size();
        nodes.emplace_back();
        if (!links.empty()) {
            const PendingLink l = links.back();
            links.pop_back();
            nodes[l.parent].children[l.slot] = i + root_offset;
        }
        write_node4(nodes[i], count, child, links, i);
    }

    void leaf(const BBox& bb, int count, const int* ids) {
        const int link = ~(int)inst_nodes.size();
        if (links.empty()) {
            nodes.push_back(dummy_parent4(bb, link));
        } else {
            const PendingLink l = links.back();
            links.pop_back();
            nodes[l.parent].children[l.slot] = link;
        }
        for (int j = 0; j < count; j++) {
            const InstanceDesc& d = instances[ids[j]];
            ImbaInstanceNode in{};
            for (int r = 0; r < 3; r++) {
                in.transf[r].x = d.inv_mat[r][0];
                in.transf[r].y = d.inv_mat[r][1];
                in.transf[r].z = d.inv_mat[r][2];
                in.transf[r].w = d.inv_mat[r][3];
            }
            in.id = ids[j];
            in.next = mesh_roots[d.mesh];
            inst_nodes.push_back(in);
        }
        inst_nodes.back().pad[0] = -1;
        inst_nodes.back().pad[1] = -1;
    }
};

// Cost functions of the reference adapters.
struct GpuMeshCost {  // core/gpu_adapter.cpp:51-58
    static float leaf_cost(int count, float area) { return (float)count * area; }
    static float traversal_cost(float area) { return area * 1.0f; }
};
struct GpuTopCost {   // core/gpu_adapter.cpp:192-199
    static float leaf_cost(int count, float area) { return (float)count * area; }
    static float traversal_cost(float area) { return area * 0.5f; }
};
struct CpuCost {      // core/cpu_adapter.cpp:45-52,257-264
    static float leaf_cost(int count, float area) { return (float)((count - 1) / 4 + 1) * area; }
    static float traversal_cost(float area) { return area * 0.5f; }
};

}  // namespace imba
