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
