// Builder emission streams (imba_bvh_emit_*): the raw node/leaf sequence of one builder run.
#include <cstring>
#include <vector>

#include "bvh_build.h"
#include "imba_hip.h"
#include "layout.h"
#include "status.h"

using namespace imba;

namespace {

struct Emitter {
    std::vector<int32_t> v;
    void f(float x) {
        int32_t b;
        std::memcpy(&b, &x, 4);
        v.push_back(b);
    }
    void box(const BBox& b) {
        f(b.min.x); f(b.min.y); f(b.min.z);
        f(b.max.x); f(b.max.y); f(b.max.z);
    }
    void node(const BBox& parent, int count, const BBox* child) {
        v.push_back(0);
        v.push_back(count);
        box(parent);
        for (int i = 0; i < count; i++) box(child[i]);
    }
    void leaf(const BBox& bb, int count, const int* ids) {
        v.push_back(1);
        v.push_back(count);
        box(bb);
        for (int i = 0; i < count; i++) v.push_back(ids[i]);
    }
};

template <int N, typename Cost>
void emit_mesh(const TriSource& src, int builder, float alpha, Emitter& e) {
    auto on_node = [&](const BBox& p, int c, const BBox* b) { e.node(p, c, b); };
    auto on_leaf = [&](const BBox& bb, int c, const int* ids) { e.leaf(bb, c, ids); };
    if (builder == IMBA_BUILDER_SBVH) {
        SplitBvhBuilder<N, Cost> b;
        b.build(src, on_node, on_leaf, 2, alpha);
    } else {
        BinnedSahBuilder<N, Cost> b;
        b.build_mesh(src, on_node, on_leaf);
    }
}

template <int N, typename Cost>
void emit_boxes(const BBox* boxes, const float3* centres, int n, Emitter& e) {
    BinnedSahBuilder<N, Cost> b;
    b.build(boxes, centres, n, [&](const BBox& p, int c, const BBox* bx) { e.node(p, c, bx); },
            [&](const BBox& bb, int c, const int* ids) { e.leaf(bb, c, ids); });
}

int finish(const Emitter& e, int32_t* out, size_t cap, size_t* len) {
    if (len) *len = e.v.size();
    if (!out || cap < e.v.size()) return fail(IMBA_ERR_CAPACITY, "stream needs %zu entries", e.v.size());
    std::memcpy(out, e.v.data(), e.v.size() * sizeof(int32_t));
    return IMBA_OK;
}

}  // namespace

extern "C" {

int imba_bvh_emit_mesh(const ImbaMeshDesc* m, int builder, int width, int cost, int32_t* out, size_t cap, size_t* len) {
    IMBA_CHECK_ARG(m && m->vertices && m->indices && m->tri_count > 0, "bad mesh");
    IMBA_CHECK_ARG(width == 2 || width == 4, "width must be 2 or 4");
    IMBA_CHECK_ARG(builder == IMBA_BUILDER_SBVH || builder == IMBA_BUILDER_BINNED_SAH, "unknown builder");
    TriSource src;
    src.vertices = m->vertices;
    src.indices = m->indices;
    src.tri_count = m->tri_count;
    Emitter e;
    // Split thresholds of the reference adapters: 1e-5 for the GPU mesh (gpu_adapter.cpp:43
    // uses the default), 1e-4 for the CPU mesh (cpu_adapter.cpp:37).
    if (cost == IMBA_COST_GPU_MESH) {
        if (width == 2) emit_mesh<2, GpuMeshCost>(src, builder, 1e-5f, e);
        else emit_mesh<4, GpuMeshCost>(src, builder, 1e-5f, e);
    } else if (cost == IMBA_COST_GPU_TOP) {
        if (width == 2) emit_mesh<2, GpuTopCost>(src, builder, 1e-5f, e);
        else emit_mesh<4, GpuTopCost>(src, builder, 1e-5f, e);
    } else if (cost == IMBA_COST_CPU) {
        if (width == 2) emit_mesh<2, CpuCost>(src, builder, 1e-4f, e);
        else emit_mesh<4, CpuCost>(src, builder, 1e-4f, e);
    } else {
        return fail(IMBA_ERR_INVALID_ARG, "unknown cost");
    }
    return finish(e, out, cap, len);
}

int imba_bvh_emit_boxes(const float* boxes, const float* centres, int n, int width, int cost, int32_t* out,
                        size_t cap, size_t* len) {
    IMBA_CHECK_ARG(boxes && centres && n > 0, "bad input");
    IMBA_CHECK_ARG(width == 2 || width == 4, "width must be 2 or 4");
    std::vector<BBox> bb(n);
    std::vector<float3> c(n);
    for (int i = 0; i < n; i++) {
        bb[i].min = float3(boxes[6 * i], boxes[6 * i + 1], boxes[6 * i + 2]);
        bb[i].max = float3(boxes[6 * i + 3], boxes[6 * i + 4], boxes[6 * i + 5]);
        c[i] = float3(centres[3 * i], centres[3 * i + 1], centres[3 * i + 2]);
    }
    Emitter e;
    if (cost == IMBA_COST_GPU_MESH) {
        if (width == 2) emit_boxes<2, GpuMeshCost>(bb.data(), c.data(), n, e);
        else emit_boxes<4, GpuMeshCost>(bb.data(), c.data(), n, e);
    } else if (cost == IMBA_COST_GPU_TOP) {
        if (width == 2) emit_boxes<2, GpuTopCost>(bb.data(), c.data(), n, e);
        else emit_boxes<4, GpuTopCost>(bb.data(), c.data(), n, e);
    } else if (cost == IMBA_COST_CPU) {
        if (width == 2) emit_boxes<2, CpuCost>(bb.data(), c.data(), n, e);
        else emit_boxes<4, CpuCost>(bb.data(), c.data(), n, e);
    } else {
        return fail(IMBA_ERR_INVALID_ARG, "unknown cost");
    }
    return finish(e, out, cap, len);
}

}  // extern "C"
