// Four-wide traversal structure derived from a two-level BVH2 scene.
//
// The GPU traversal can walk the scene's BVH2 as given (core/gpu_adapter.cpp layout) or a
// 4-wide tree derived from it at upload: every 4-wide node takes the two children of a BVH2
// node and, while it has fewer than four, replaces its inner child of largest surface area
// by that child's two children.  Leaves (triangle lists, instance lists) and their boxes are
// the BVH2's own, so a ray meets the same triangles with the same arithmetic; only boxes of
// the absorbed inner nodes are no longer tested, and each step tests four boxes instead of
// two.  Node layout: ImbaNode4 (box j of child j as min/max per axis, children[j]: > 0 node,
// < 0 leaf link as in the BVH2, 0 unused slot).  The first kTreeletMaxNodes nodes of a tree
// are its top levels in breadth-first order, the rest follow in preorder (a child always
// follows its parent); mesh trees come first (in increasing BVH2 root order), then the top
// level.  Instance records are copied with `next` pointing at the mesh's 4-wide root.
#include <algorithm>
#include <cmath>
#include <cfloat>
#include <cstring>
#include <deque>
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
// largest number of stack entries a ray can hold inside this tree.  The first kTreeletMaxNodes
// nodes are taken from the front of the work list (breadth first), the others from its back.
int collapse_tree(const ImbaSceneDesc& d, int root2, std::vector<ImbaNode4>& out, int* pushes) {
    struct Task {
        int node2, parent4, slot, depth_pushes;
    };
    std::deque<Task> work{{root2, -1, 0, 0}};
    int root4 = -1;
    const size_t first = out.size();
    auto inner_ok = [&](int link) { return link >= 0 && link < d.node_count; };
    while (!work.empty()) {
        const bool bfs = out.size() - first < (size_t)kTreeletMaxNodes;
        const Task t = bfs ? work.front() : work.back();
        if (bfs) work.pop_front();
        else work.pop_back();
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
        if (out.size() - first < (size_t)kTreeletMaxNodes) {
            for (int j = 0; j < k; j++)
                if (list[j].link >= 0) work.push_back({list[j].link, i4, j, here});
        } else {
            for (int j = k - 1; j >= 0; j--)
                if (list[j].link >= 0) work.push_back({list[j].link, i4, j, here});
        }
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
    // the treelet: the top level, or the mesh when the top level reaches one instance only
    int reached = 0, only = -1;
    for (int i = 0; i < d.instance_count; i++)
        if (used[i]) { reached++; only = i; }
    int base = out.root, end = (int)out.nodes.size();
    if (reached == 1) {  // the only mesh tree spans the nodes before the top level
        base = root_map[d.instances[only].next];
        end = out.root;
    }
    out.treelet_base = base;
    out.treelet_nodes = std::min(kTreeletMaxNodes, end - base);
    out.instances.assign(d.instances, d.instances + d.instance_count);
    for (int i = 0; i < d.instance_count; i++) out.instances[i].next = used[i] ? root_map[d.instances[i].next] : -1;
    // stack entries of the top level stay while a ray is inside a mesh tree
    out.max_stack = top_pushes + mesh_pushes + 2;
    quantize_nodes(out);
    return IMBA_OK;
}

namespace {

float pow2(int e) { return std::ldexp(1.0f, e); }

// Smallest exponent e >= -100 with 255 * 2^e >= extent, and quantized bounds of every child
// such that the dequantized box (computed as the kernel does) contains the exact box.
bool quantize_axis(const float* lo, const float* hi, int k, float origin, float extent, int8_t* exp_out, uint8_t* qlo,
                   uint8_t* qhi) {
    if (!std::isfinite(origin) || !std::isfinite(extent) || extent < 0.0f) return false;
    int e = -100;
    if (extent > 0.0f) {
        int ex = 0;
        std::frexp(extent / 255.0f, &ex);  // extent/255 < 2^ex
        e = std::max(-100, ex);
    }
    for (; e <= 127; e++) {
        const float s = pow2(e);
        bool ok = true;
        for (int j = 0; j < k && ok; j++) {
            if (!std::isfinite(lo[j]) || !std::isfinite(hi[j])) return false;
            long ql = (long)std::floor((double)(lo[j] - origin) / s);
            long qh = (long)std::ceil((double)(hi[j] - origin) / s);
            ql = std::max(0L, std::min(255L, ql));
            qh = std::max(0L, std::min(255L, qh));
            while (ql > 0 && std::fma((float)ql, s, origin) > lo[j]) ql--;
            while (qh < 255 && std::fma((float)qh, s, origin) < hi[j]) qh++;
            if (std::fma((float)ql, s, origin) > lo[j] || std::fma((float)qh, s, origin) < hi[j]) ok = false;
            qlo[j] = (uint8_t)ql;
            qhi[j] = (uint8_t)qh;
        }
        if (ok) {
            *exp_out = (int8_t)e;
            return true;
        }
    }
    return false;
}

}  // namespace

bool quantize_nodes(WideScene& w) {
    w.qnodes.assign(w.nodes.size(), QNode4{});
    for (size_t i = 0; i < w.nodes.size(); i++) {
        const ImbaNode4& n = w.nodes[i];
        QNode4& q = w.qnodes[i];
        std::memset(&q, 0, sizeof(q));
        int k = 0;
        while (k < 4 && n.children[k] != 0) k++;
        q.count = (uint8_t)k;
        for (int j = 0; j < 4; j++) q.children[j] = n.children[j];
        const float* lo[3] = {n.min_x, n.min_y, n.min_z};
        const float* hi[3] = {n.max_x, n.max_y, n.max_z};
        uint8_t* qlo[3] = {q.lo_x, q.lo_y, q.lo_z};
        uint8_t* qhi[3] = {q.hi_x, q.hi_y, q.hi_z};
        for (int a = 0; a < 3; a++) {
            float mn = lo[a][0], mx = hi[a][0];
            for (int j = 1; j < k; j++) {
                mn = std::min(mn, lo[a][j]);
                mx = std::max(mx, hi[a][j]);
            }
            q.origin[a] = mn;
            if (!quantize_axis(lo[a], hi[a], k, mn, mx - mn, &q.exp[a], qlo[a], qhi[a])) {
                w.qnodes.clear();
                return false;
            }
        }
    }
    return true;
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

int imba_wide_get_quantized(const ImbaWideBvh* h, const void** nodes, int* node_count) {
    IMBA_CHECK_ARG(h && nodes && node_count, "NULL argument");
    *nodes = h->w.qnodes.data();
    *node_count = (int)h->w.qnodes.size();
    return IMBA_OK;
}

int imba_wide_release(ImbaWideBvh* h) {
    delete h;
    return IMBA_OK;
}

}  // extern "C"
