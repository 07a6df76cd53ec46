// Four-wide traversal structure derived from a two-level BVH2 scene.
//
// The GPU traversal can walk the scene's BVH2 as given (core/gpu_adapter.cpp layout) or a
// 4-wide tree derived from it at upload: every 4-wide node takes the two children of a BVH2
// node and, while it has fewer than four, replaces its inner child of largest surface area
// by that child's two children.  Leaves (triangle lists, instance lists) and their boxes are
// the BVH2's own, so a ray meets the same triangles with the same arithmetic; only boxes of
// the absorbed inner nodes are no longer tested, and each step tests four boxes instead of
// two.  Node layout: ImbaNode4 (box j of child j as min/max per axis, children[j]: > 0 node,
// < 0 leaf link as in the BVH2, 0 unused slot).  Nodes of one tree are in preorder (a child
// follows its parent); mesh trees come first (in increasing BVH2 root order), then the top
// level.  Instance records are copied with `next` pointing at the mesh's 4-wide root.
#include <algorithm>
#include <cfloat>
#include <cstring>
#include <map>
#include <vector>

#include "status.h"
#include "wide_bvh.h"

namespace imba {

namespace {

struct Entry {
    ImbaBBox box;
    int link;
};

float half_area(const ImbaBBox& b) {
    const float dx = b.hi_x - b.lo_x, dy = b.hi_y - b.lo_y, dz = b.hi_z - b.lo_z;
    return dx * dy + dy * dz + dz * dx;
}

void set_slot(ImbaNode4& n, int j, const ImbaBBox& b) {
    n.min_x[j] = b.lo_x; n.max_x[j] = b.hi_x;
    n.min_y[j] = b.lo_y; n.max_y[j] = b.hi_y;
    n.min_z[j] = b.lo_z; n.max_z[j] = b.hi_z;
}

// Collapses the BVH2 tree at root2; returns the 4-wide root index and updates *pushes with the
// largest number of stack entries a ray can hold inside this tree.
int collapse_tree(const ImbaSceneDesc& d, int root2, std::vector<ImbaNode4>& out, int* pushes) {
    struct Task {
        int node2, parent4, slot, depth_pushes;
    };
    std::vector<Task> work{{root2, -1, 0, 0}};
    int root4 = -1;
    auto inner_ok = [&](int link) { return link >= 0 && link < d.node_count; };
    while (!work.empty()) {
        const Task t = work.back();
        work.pop_back();
        if (!inner_ok(t.node2)) return -1;
        const ImbaNode& n = d.nodes[t.node2];
        Entry list[4];
        int k = 0;
        list[k++] = {n.left_bb, n.left};
        if (n.right != IMBA_DUMMY_CHILD) list[k++] = {n.right_bb, n.right};
        while (k < 4) {
            int best = -1;
            float best_area = -1.0f;
            for (int j = 0; j < k; j++) {
                if (list[j].link < 0) continue;
                const float a = half_area(list[j].box);
                if (a > best_area) {
                    best_area = a;
                    best = j;
                }
            }
            if (best < 0) break;
            if (!inner_ok(list[best].link)) return -1;
            const ImbaNode& c = d.nodes[list[best].link];
            list[best] = {c.left_bb, c.left};
            if (c.right != IMBA_DUMMY_CHILD) list[k++] = {c.right_bb, c.right};
        }
        const int i4 = (int)out.size();
        out.emplace_back();
        ImbaNode4& w = out.back();
        std::memset(&w, 0, sizeof(w));
        for (int j = 0; j < 4; j++) {
            if (j < k) {
                set_slot(w, j, list[j].box);
                w.children[j] = list[j].link < 0 ? list[j].link : 0;  // inner links are set by the child
            } else {
                const ImbaBBox empty = {FLT_MAX, -FLT_MAX, FLT_MAX, -FLT_MAX, FLT_MAX, -FLT_MAX};
                set_slot(w, j, empty);
                w.children[j] = 0;
            }
        }
        w.pad[0] = k;
        if (t.parent4 >= 0) out[t.parent4].children[t.slot] = i4;
        else root4 = i4;
        const int here = t.depth_pushes + (k - 1);
        *pushes = std::max(*pushes, here);
        for (int j = k - 1; j >= 0; j--)
            if (list[j].link >= 0) work.push_back({list[j].link, i4, j, here});
    }
    return root4;
}

}  // namespace

int collapse_scene4(const ImbaSceneDesc& d, WideScene& out) {
    out = WideScene{};
    if (d.root < 0 || d.root >= d.node_count) return fail(IMBA_ERR_BAD_BVH, "root out of range");
    // the instance records the top level reaches, and their meshes' roots
    std::vector<uint8_t> used(d.instance_count, 0);
    std::map<int, int> root_map;
    {
        std::vector<int> work{d.root};
        while (!work.empty()) {
            const int k = work.back();
            work.pop_back();
            if (k < 0 || k >= d.node_count) return fail(IMBA_ERR_BAD_BVH, "top-level link %d out of range", k);
            const ImbaNode& n = d.nodes[k];
            const int ch[2] = {n.left, n.right};
            for (int j = 0; j < 2; j++) {
                if (j == 1 && ch[j] == IMBA_DUMMY_CHILD) continue;
                if (ch[j] >= 0) {
                    work.push_back(ch[j]);
                    continue;
                }
                for (int i = ~ch[j]; i >= 0 && i < d.instance_count; i++) {
                    used[i] = 1;
                    root_map[d.instances[i].next] = -1;
                    if (d.instances[i].pad[0] == -1) break;
                }
            }
        }
    }
    int mesh_pushes = 0;
    for (auto& kv : root_map) {
        kv.second = collapse_tree(d, kv.first, out.nodes, &mesh_pushes);
        if (kv.second < 0) return fail(IMBA_ERR_BAD_BVH, "mesh tree at %d has links out of range", kv.first);
    }
    int top_pushes = 0;
    out.root = collapse_tree(d, d.root, out.nodes, &top_pushes);
    if (out.root < 0) return fail(IMBA_ERR_BAD_BVH, "top level has links out of range");
    out.instances.assign(d.instances, d.instances + d.instance_count);
    for (int i = 0; i < d.instance_count; i++) out.instances[i].next = used[i] ? root_map[d.instances[i].next] : -1;
    // stack entries of the top level stay while a ray is inside a mesh tree
    out.max_stack = top_pushes + mesh_pushes + 2;
    return IMBA_OK;
}

}  // namespace imba

using namespace imba;

struct ImbaWideBvh {
    WideScene w;
};

extern "C" {

int imba_wide_create(const ImbaSceneDesc* host, ImbaWideBvh** out) {
    IMBA_CHECK_ARG(host && out, "NULL argument");
    *out = nullptr;
    auto* h = new ImbaWideBvh();
    int r = collapse_scene4(*host, h->w);
    if (r) {
        delete h;
        return r;
    }
    *out = h;
    return IMBA_OK;
}

int imba_wide_get(const ImbaWideBvh* h, const ImbaNode4** nodes, int* node_count, const ImbaInstanceNode** instances,
                  int* instance_count, int* root, int* max_stack) {
    IMBA_CHECK_ARG(h && nodes && node_count && instances && instance_count && root && max_stack, "NULL argument");
    *nodes = h->w.nodes.data();
    *node_count = (int)h->w.nodes.size();
    *instances = h->w.instances.data();
    *instance_count = (int)h->w.instances.size();
    *root = h->w.root;
    *max_stack = h->w.max_stack;
    return IMBA_OK;
}

int imba_wide_release(ImbaWideBvh* h) {
    delete h;
    return IMBA_OK;
}

}  // extern "C"
