// Harness around the reference's own BVH builders — TEST INFRASTRUCTURE ONLY.
//
// Compiled by oracle/Makefile directly from the reference headers (core/fast_bvh_builder.h,
// core/sbvh_builder.h and what they include) into oracle/_ref/libref_builders.so.  It runs a
// builder with writer callbacks that record the node/leaf emission stream in the format of
// imba_bvh_emit_* (include/imba_hip.h), so tests can compare the library's builders with the
// reference's bit for bit.  The cost functions are those the reference adapters pass as
// template arguments (core/gpu_adapter.cpp:51-58,192-199; core/cpu_adapter.cpp:45-52,257-264).
#include <cstdint>
#include <cstring>
#include <vector>

#include "imbatracer/core/fast_bvh_builder.h"
#include "imbatracer/core/mesh.h"
#include "imbatracer/core/sbvh_builder.h"
#include "imbatracer/render/random.h"

namespace {

struct GpuMeshCost {
    static float leaf_cost(int count, float area) { return count * area; }
    static float traversal_cost(float area) { return area * 1.0f; }
};
struct GpuTopCost {
    static float leaf_cost(int count, float area) { return count * area; }
    static float traversal_cost(float area) { return area * 0.5f; }
};
struct CpuCost {
    static float leaf_cost(int count, float area) { return ((count - 1) / 4 + 1) * area; }
    static float traversal_cost(float area) { return area * 0.5f; }
};

struct Rec {
    std::vector<int32_t>* v;
    void f(float x) {
        int32_t b;
        std::memcpy(&b, &x, 4);
        v->push_back(b);
    }
    void box(const imba::BBox& b) {
        f(b.min.x); f(b.min.y); f(b.min.z);
        f(b.max.x); f(b.max.y); f(b.max.z);
    }
};

struct NodeRec : Rec {
    template <typename BBoxFn>
    void operator()(const imba::BBox& parent, int count, BBoxFn bboxes) {
        v->push_back(0);
        v->push_back(count);
        box(parent);
        for (int i = 0; i < count; i++) box(bboxes(i));
    }
};

struct LeafRec : Rec {
    template <typename RefFn>
    void operator()(const imba::BBox& bb, int count, RefFn refs) {
        v->push_back(1);
        v->push_back(count);
        box(bb);
        for (int i = 0; i < count; i++) v->push_back((int32_t)refs(i));
    }
};

imba::Mesh make_mesh(const float* v4, int nv, const uint32_t* idx4, int nt) {
    imba::Mesh m;
    m.set_vertex_count(nv);
    m.set_index_count(4 * (size_t)nt);
    for (int i = 0; i < nv; i++) m.vertices()[i] = imba::float4(v4[4 * i], v4[4 * i + 1], v4[4 * i + 2], v4[4 * i + 3]);
    std::memcpy(m.indices(), idx4, sizeof(uint32_t) * 4 * (size_t)nt);
    return m;
}

template <int N, typename Cost>
void run_mesh(const imba::Mesh& m, int builder, float alpha, std::vector<int32_t>& out) {
    NodeRec nr;
    LeafRec lr;
    nr.v = &out;
    lr.v = &out;
    if (builder == 0) {
        imba::SplitBvhBuilder<N, Cost> b;
        b.build(m, nr, lr, 2, alpha);
    } else {
        imba::FastBvhBuilder<N, Cost> b;
        b.build(m, nr, lr, 2);
    }
}

template <int N, typename Cost>
void run_boxes(const imba::BBox* bb, const imba::float3* c, int n, std::vector<int32_t>& out) {
    NodeRec nr;
    LeafRec lr;
    nr.v = &out;
    lr.v = &out;
    imba::FastBvhBuilder<N, Cost> b;
    b.build(bb, c, n, nr, lr, 1);
}

int finish(const std::vector<int32_t>& v, int32_t* out, size_t cap, size_t* len) {
    *len = v.size();
    if (!out || cap < v.size()) return 5;
    std::memcpy(out, v.data(), v.size() * sizeof(int32_t));
    return 0;
}

}  // namespace

extern "C" int ref_emit_mesh(const float* v4, int nv, const uint32_t* idx4, int nt, int builder, int width, int cost,
                             int32_t* out, size_t cap, size_t* len) {
    const imba::Mesh m = make_mesh(v4, nv, idx4, nt);
    std::vector<int32_t> v;
    if (cost == 0) {
        if (width == 2) run_mesh<2, GpuMeshCost>(m, builder, 1e-5f, v); else run_mesh<4, GpuMeshCost>(m, builder, 1e-5f, v);
    } else if (cost == 1) {
        if (width == 2) run_mesh<2, GpuTopCost>(m, builder, 1e-5f, v); else run_mesh<4, GpuTopCost>(m, builder, 1e-5f, v);
    } else {
        if (width == 2) run_mesh<2, CpuCost>(m, builder, 1e-4f, v); else run_mesh<4, CpuCost>(m, builder, 1e-4f, v);
    }
    return finish(v, out, cap, len);
}

// The reference generator (render/random.h:15-47): n draws of random_int(0, 0x7FFFFFFF)
// (= MWC64X() mod 0x7FFFFFFF; the raw draw is private), and n random_float() draws of a
// second generator with the same seed.
extern "C" void ref_rng(uint64_t seed, int n, uint32_t* raw, float* uniform) {
    imba::RNG a(seed), b(seed);
    for (int i = 0; i < n; i++) raw[i] = (uint32_t)a.random_int(0, 0x7FFFFFFF);
    for (int i = 0; i < n; i++) uniform[i] = b.random_float();
}

// Random-triangle scene of the CPU-runnable configuration drawn with the reference generator:
// per triangle a centre (3 x random_float()), then three vertices centre + random_float(-0.01, 0.01)
// per coordinate.
extern "C" void ref_random_triangles(uint64_t seed, int n, float* vertices) {
    imba::RNG rng(seed);
    for (int t = 0; t < n; t++) {
        float c[3];
        for (int k = 0; k < 3; k++) c[k] = rng.random_float();
        for (int v = 0; v < 3; v++) {
            float* out = vertices + 4 * (3 * (size_t)t + v);
            for (int k = 0; k < 3; k++) out[k] = c[k] + rng.random_float(-0.01f, 0.01f);
            out[3] = 1.0f;
        }
    }
}

extern "C" int ref_emit_boxes(const float* boxes, const float* centres, int n, int width, int cost, int32_t* out,
                              size_t cap, size_t* len) {
    std::vector<imba::BBox> bb(n);
    std::vector<imba::float3> c(n);
    for (int i = 0; i < n; i++) {
        bb[i] = imba::BBox(imba::float3(boxes[6 * i], boxes[6 * i + 1], boxes[6 * i + 2]),
                           imba::float3(boxes[6 * i + 3], boxes[6 * i + 4], boxes[6 * i + 5]));
        c[i] = imba::float3(centres[3 * i], centres[3 * i + 1], centres[3 * i + 2]);
    }
    std::vector<int32_t> v;
    if (cost == 0) {
        if (width == 2) run_boxes<2, GpuMeshCost>(bb.data(), c.data(), n, v); else run_boxes<4, GpuMeshCost>(bb.data(), c.data(), n, v);
    } else if (cost == 1) {
        if (width == 2) run_boxes<2, GpuTopCost>(bb.data(), c.data(), n, v); else run_boxes<4, GpuTopCost>(bb.data(), c.data(), n, v);
    } else {
        if (width == 2) run_boxes<2, CpuCost>(bb.data(), c.data(), n, v); else run_boxes<4, CpuCost>(bb.data(), c.data(), n, v);
    }
    return finish(v, out, cap, len);
}
