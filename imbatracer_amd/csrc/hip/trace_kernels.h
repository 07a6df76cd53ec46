// Closest-hit and any-hit traversal of two-level BVH2 scenes on gfx950.
//
// Replaces the AnyDSL kernels intersect_gpu_masked_instanced / occluded_gpu_masked_instanced
// that the reference calls at render/scheduling/ray_queue.h:297-308 and 347-358 (their source
// is not part of the reference).  Layouts are those written by the reference's GPU adapters
// (core/gpu_adapter.cpp): 64-byte BVH2 nodes, three Vec4 per triangle with the triangle id in
// v1.w and the leaf end flagged by the bit pattern 0x80000000 in v2.w, 64-byte instance
// records ending a top-level leaf with pad = {-1,-1}.
//
// Per ray the traversal is depth first: both child boxes of a node are tested, the nearer
// child (ties: left) is visited first and the other is pushed.  Entering an instance
// transforms the ray into object space and remembers the stack depth; when the stack falls
// back to that depth the next instance of the same leaf is entered, or the world ray is
// restored.  Arithmetic (fixed, shared with the CPU restatement used as the checker):
//   box:      t = fma(bound, 1/d', -o * (1/d')), d' = d or +-1e-9 when |d| < 1e-9;
//             near = max(max(min x, min y), max(min z, tmin)),
//             far  = min(min(max x, max y), min(max z, t_best)); hit when near <= far
//   triangle: e1 = v0 - v1, e2 = v2 - v0, n = e1 x e2, c = v0 - o, r = d x c,
//             det = n.d, U = r.e2, V = r.e1, T = n.c (cross and dot with fma as below);
//             with signs folded by det: U >= 0, V >= 0, |det| - U - V >= 0, det != 0;
//             t = T / |det|, u = U / |det|
//   cross(a, b) = (fma(a.y, b.z, -(a.z b.y)), fma(a.z, b.x, -(a.x b.z)), fma(a.x, b.y, -(a.y b.x)))
//   dot(a, b)   = fma(a.x, b.x, fma(a.y, b.y, a.z b.z))
//   instance:   o' = M (o, 1), d' = M (d, 0) with row sums fma(m0, x, fma(m1, y, fma(m2, z, m3)))
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace imba {

struct TraceArgs {
    const float4* nodes;  int n_nodes;
    const float4* inst;   int n_inst;
    const float4* tris;   int n_tris4;
    const int4* indices;  int n_tri_ids;     // triangles described by `indices`
    const float2* texcoords; int n_texcoords;
    const int* masks;     int n_masks;       // 3 ints per mask: width, height, offset
    const unsigned char* mask_buf; int mask_buf_size;
    const float4* rays;   int4* hits; int count; int root;
    uint32_t* counts;     // 3 per ray (stats variant), or null
    uint32_t* depth_hist; // stats variant: rays by the deepest stack entry they used (64 bins,
                          // the last one for 63 and deeper), or null
    int* err;             // incremented once per ray that met a malformed structure
    int* spill;           // global stack entries beyond the LDS part
    int spill_depth;      // spill entries per thread
    int* work_counter;    // persistent variant: next ray to fetch
    const int* count_dev; // device ray count read when the launch runs (<= count), or null
    int tl_base, tl_nodes; // quantized tree: nodes [tl_base, tl_base + tl_nodes) are read from
                           // the block's copy in LDS (dynamic shared memory, 64 B per node)
};

// Rays of this launch: count, or the device count clamped to [0, count].
__device__ __forceinline__ int launch_count(const TraceArgs& a) {
    return a.count_dev ? min(max(*a.count_dev, 0), a.count) : a.count;
}

constexpr int kNoNode = INT32_MIN;
constexpr int kDummyChild = 0x76543210;
constexpr uint32_t kLeafEnd = 0x80000000u;

struct RaySpace {
    float ox, oy, oz, dx, dy, dz;
    float idx, idy, idz, oidx, oidy, oidz;
};

__device__ __forceinline__ float safe_dir(float d) { return fabsf(d) < 1e-9f ? copysignf(1e-9f, d) : d; }

__device__ __forceinline__ void setup_space(RaySpace& r) {
    r.idx = 1.0f / safe_dir(r.dx);
    r.idy = 1.0f / safe_dir(r.dy);
    r.idz = 1.0f / safe_dir(r.dz);
    r.oidx = r.ox * r.idx;
    r.oidy = r.oy * r.idy;
    r.oidz = r.oz * r.idz;
}

__device__ __forceinline__ bool box_hit(const RaySpace& r, float lx, float hx, float ly, float hy, float lz, float hz,
                                        float tmin, float tmax, float& tnear) {
    const float t0x = fmaf(lx, r.idx, -r.oidx), t1x = fmaf(hx, r.idx, -r.oidx);
    const float t0y = fmaf(ly, r.idy, -r.oidy), t1y = fmaf(hy, r.idy, -r.oidy);
    const float t0z = fmaf(lz, r.idz, -r.oidz), t1z = fmaf(hz, r.idz, -r.oidz);
    tnear = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), tmin));
    const float tfar = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fminf(fmaxf(t0z, t1z), tmax));
    return tnear <= tfar;
}

// Returns true when the ray meets the triangle (edges included, det != 0); t, u, v are
// then the distance and the barycentric weights of v1 and v2.
__device__ __forceinline__ bool tri_hit(const RaySpace& r, const float4& a, const float4& b, const float4& c,
                                        float& t, float& u, float& v) {
    const float e1x = a.x - b.x, e1y = a.y - b.y, e1z = a.z - b.z;
    const float e2x = c.x - a.x, e2y = c.y - a.y, e2z = c.z - a.z;
    const float nx = fmaf(e1y, e2z, -(e1z * e2y));
    const float ny = fmaf(e1z, e2x, -(e1x * e2z));
    const float nz = fmaf(e1x, e2y, -(e1y * e2x));
    const float cx = a.x - r.ox, cy = a.y - r.oy, cz = a.z - r.oz;
    const float rx = fmaf(r.dy, cz, -(r.dz * cy));
    const float ry = fmaf(r.dz, cx, -(r.dx * cz));
    const float rz = fmaf(r.dx, cy, -(r.dy * cx));
    const float det = fmaf(nx, r.dx, fmaf(ny, r.dy, nz * r.dz));
    const float U = fmaf(rx, e2x, fmaf(ry, e2y, rz * e2z));
    const float V = fmaf(rx, e1x, fmaf(ry, e1y, rz * e1z));
    const float T = fmaf(nx, cx, fmaf(ny, cy, nz * cz));
    const bool neg = det < 0.0f;
    const float ad = fabsf(det);
    const float us = neg ? -U : U;
    const float vs = neg ? -V : V;
    const float w = (ad - us) - vs;
    if (!(us >= 0.0f && vs >= 0.0f && w >= 0.0f && det != 0.0f)) return false;
    const float ts = neg ? -T : T;
    t = ts / ad;
    u = us / ad;
    v = vs / ad;
    return true;
}

// Opacity of the material mask at the hit (texel: wrap the interpolated texcoords to [0,1),
// scale by the mask size, truncate, clamp to the last texel).  Returns 1 when opaque, 0 when
// transparent, and 3 (opaque, bit 1 set) when the scene arrays are inconsistent.
__device__ __forceinline__ int texel_opaque(const TraceArgs& a, int tri, float u, float v) {
    if ((unsigned)tri >= (unsigned)a.n_tri_ids) return 3;
    const int4 ind = a.indices[tri];
    if ((unsigned)ind.w >= (unsigned)a.n_masks) return 3;
    const int w = a.masks[3 * ind.w + 0], h = a.masks[3 * ind.w + 1], off = a.masks[3 * ind.w + 2];
    int texel = 0;
    if (!(w == 1 && h == 1)) {
        if ((unsigned)ind.x >= (unsigned)a.n_texcoords || (unsigned)ind.y >= (unsigned)a.n_texcoords ||
            (unsigned)ind.z >= (unsigned)a.n_texcoords)
            return 3;
        const float2 t0 = a.texcoords[ind.x], t1 = a.texcoords[ind.y], t2 = a.texcoords[ind.z];
        const float w0 = (1.0f - u) - v;
        const float s = fmaf(t0.x, w0, fmaf(t1.x, u, t2.x * v));
        const float q = fmaf(t0.y, w0, fmaf(t1.y, u, t2.y * v));
        const float fs = (s - floorf(s)) * (float)w;
        const float fq = (q - floorf(q)) * (float)h;
        const int xi = (fs >= 0.0f && fs < (float)w) ? (int)fs : (fs >= (float)w ? w - 1 : 0);
        const int yi = (fq >= 0.0f && fq < (float)h) ? (int)fq : (fq >= (float)h ? h - 1 : 0);
        texel = yi * w + xi;
    }
    const long long at = (long long)off + texel;
    if (at < 0 || at >= a.mask_buf_size) return 3;
    return a.mask_buf[at] != 0 ? 1 : 0;
}

// Address-space qualifiers for the device pass (the host pass sees plain pointers).
#if defined(__HIP_DEVICE_COMPILE__)
#define IMBA_AS(n) __attribute__((address_space(n)))
#else
#define IMBA_AS(n)
#endif

// The LDS (address space 3) view of a pointer into a kernel's __shared__ memory: accesses through
// it are ds_read / ds_write, not flat accesses (which the compiler otherwise emits when an LDS
// and a global access are merged into one load through a selected pointer).
template <typename T>
__device__ __forceinline__ IMBA_AS(3) T* lds_view(T* p) {
    return (IMBA_AS(3) T*)p;
}

// The global (address space 1) view of a scene or queue array: loads through it are
// global_load, not flat_load, also where the pointer itself was read from memory (the batched
// kernel's queue table), which the compiler cannot otherwise place.
template <typename T>
__device__ __forceinline__ const IMBA_AS(1) T* gview(const T* p) {
    return (const IMBA_AS(1) T*)p;
}
template <typename T>
__device__ __forceinline__ IMBA_AS(1) T* gview_w(T* p) {
    return (IMBA_AS(1) T*)p;
}

// Per-lane traversal stack: entry 0..sp-1 in LDS ([depth][lane], conflict free), then in the
// global spill buffer.  With IMBA_STACK_TOP_REG=1 the top entry lives in a register instead: a
// pop returns it at once and issues the load that refills it, which has a whole node step to
// arrive; a push writes the previous top out.  Measured on the 21M-triangle PT frame (quantized
// 4-wide, 256 x 16): 970 / 973 against 975 Mrays/s with every entry in LDS (one more VGPR, 4
// more bytes of scratch per lane in the 7-wave variants), so the pops' LDS latency is hidden
// already and the option stays off.
#ifndef IMBA_STACK_TOP_REG
#define IMBA_STACK_TOP_REG 0
#endif

template <int BLOCK, int STACK_LDS>
struct LaneStack {
    int* lds;     // [STACK_LDS][BLOCK]
    int* spill;   // [spill_depth][threads] or null
    int spill_depth;
    int lane_slot;   // threadIdx.x
    int spill_slot;  // global thread index
    int spill_stride;
    int sp;
    int top;          // entry sp - 1 (IMBA_STACK_TOP_REG)
    const uint4* tl;  // LDS copy of the treelet nodes (64 B each, as in memory), or null

    __device__ __forceinline__ void store(int k, int v) {
        if (k < STACK_LDS) lds_view(lds)[k * BLOCK + lane_slot] = v;
        else spill[(size_t)(k - STACK_LDS) * spill_stride + spill_slot] = v;
    }
    __device__ __forceinline__ int load(int k) {
        if (k < STACK_LDS) return lds_view(lds)[k * BLOCK + lane_slot];
        return spill[(size_t)(k - STACK_LDS) * spill_stride + spill_slot];
    }
    __device__ __forceinline__ bool push(int v) {
#if IMBA_STACK_TOP_REG
        if (sp - STACK_LDS >= spill_depth) return false;
        if (sp > 0) store(sp - 1, top);
        top = v;
#else
        // (the capacity test only on the spill path: testing it first for every push measured
        // 14% slower on the deep trees of 256 queues over 10K-4M random triangles)
        if (sp < STACK_LDS) {
            lds_view(lds)[sp * BLOCK + lane_slot] = v;
        } else {
            const int k = sp - STACK_LDS;
            if (k >= spill_depth) return false;
            spill[(size_t)k * spill_stride + spill_slot] = v;
        }
#endif
        sp++;
        return true;
    }
    __device__ __forceinline__ int pop() {
        sp--;
#if IMBA_STACK_TOP_REG
        const int v = top;
        if (sp > 0) top = load(sp - 1);
        return v;
#else
        if (sp < STACK_LDS) return lds_view(lds)[sp * BLOCK + lane_slot];
        return spill[(size_t)(sp - STACK_LDS) * spill_stride + spill_slot];
#endif
    }
};

enum : int { kTop = 1, kInstLast = 2, kBad = 4, kDone = 8 };

// Traversal state of one ray (kept in registers across the steps of the persistent kernel).
struct WorldRay {
    float ox, oy, oz, dx, dy, dz;
    float idx, idy, idz;  // 1 / safe_dir(d), as setup_space computes them
};

struct TraceState {
    WorldRay world;  // world-space ray (its reciprocals are recomputed when the ray returns to it)
    RaySpace cs;     // ray in the space of the current level (world or an instance's object space)
    float tmin;      // the ray's tmax is best_t until a hit
    float best_t, best_u;
    int best_tri, best_inst;
    int node;        // next node / leaf link to process, kNoNode if none
    int tri;         // next triangle record of the current leaf, -1 if none
    int inst, inst_id, inst_sp;  // current instance, its id, stack depth when it was entered
    int flags;
    uint32_t cn, ct, ci;  // visit counts (stats variant)
    uint32_t cd;          // deepest stack entry used (stats variant)
};

// The current space becomes world space (the reciprocals kept from the start: the values
// setup_space gives).
__device__ __forceinline__ void to_world(TraceState& s) {
    s.cs.ox = s.world.ox; s.cs.oy = s.world.oy; s.cs.oz = s.world.oz;
    s.cs.dx = s.world.dx; s.cs.dy = s.world.dy; s.cs.dz = s.world.dz;
    s.cs.idx = s.world.idx; s.cs.idy = s.world.idy; s.cs.idz = s.world.idz;
    s.cs.oidx = s.cs.ox * s.cs.idx;
    s.cs.oidy = s.cs.oy * s.cs.idy;
    s.cs.oidz = s.cs.oz * s.cs.idz;
}

__device__ __forceinline__ void state_init(TraceState& s, const float4& ro, const float4& rd, int root) {
    s.world.ox = ro.x; s.world.oy = ro.y; s.world.oz = ro.z;
    s.world.dx = rd.x; s.world.dy = rd.y; s.world.dz = rd.z;
    s.cs.ox = ro.x; s.cs.oy = ro.y; s.cs.oz = ro.z;
    s.cs.dx = rd.x; s.cs.dy = rd.y; s.cs.dz = rd.z;
    setup_space(s.cs);
    s.world.idx = s.cs.idx; s.world.idy = s.cs.idy; s.world.idz = s.cs.idz;
    s.tmin = ro.w;
    s.best_t = rd.w;
    s.best_u = 0.0f;
    s.best_tri = -1;
    s.best_inst = -1;
    s.node = root;
    s.tri = -1;
    s.inst = -1;
    s.inst_id = -1;
    s.inst_sp = -1;
    s.flags = kTop | kInstLast;
    s.cn = s.ct = s.ci = s.cd = 0;
}

// Enters instance i of the current top-level leaf: the ray goes to its object space.
// Returns the instance's root node, or kNoNode for a malformed record.
template <bool STATS>
__device__ __forceinline__ int enter_instance(const TraceArgs& a, TraceState& s, int i, int sp) {
    if ((unsigned)i >= (unsigned)a.n_inst) {
        s.flags |= kBad | kInstLast;
        return kNoNode;
    }
    const auto* inst = gview(a.inst);
    const float4 m0 = inst[4 * i + 0], m1 = inst[4 * i + 1], m2 = inst[4 * i + 2];
    const int4 meta = gview(reinterpret_cast<const int4*>(a.inst))[4 * i + 3];
    if (STATS) s.ci++;
    const WorldRay& w = s.world;
    s.cs.ox = fmaf(m0.x, w.ox, fmaf(m0.y, w.oy, fmaf(m0.z, w.oz, m0.w)));
    s.cs.oy = fmaf(m1.x, w.ox, fmaf(m1.y, w.oy, fmaf(m1.z, w.oz, m1.w)));
    s.cs.oz = fmaf(m2.x, w.ox, fmaf(m2.y, w.oy, fmaf(m2.z, w.oz, m2.w)));
    s.cs.dx = fmaf(m0.x, w.dx, fmaf(m0.y, w.dy, m0.z * w.dz));
    s.cs.dy = fmaf(m1.x, w.dx, fmaf(m1.y, w.dy, m1.z * w.dz));
    s.cs.dz = fmaf(m2.x, w.dx, fmaf(m2.y, w.dy, m2.z * w.dz));
    if (__float_as_uint(s.cs.dx) == __float_as_uint(w.dx) && __float_as_uint(s.cs.dy) == __float_as_uint(w.dy) &&
        __float_as_uint(s.cs.dz) == __float_as_uint(w.dz)) {
        // the direction is unchanged (a translation or the identity): the world reciprocals
        s.cs.idx = w.idx; s.cs.idy = w.idy; s.cs.idz = w.idz;
        s.cs.oidx = s.cs.ox * s.cs.idx;
        s.cs.oidy = s.cs.oy * s.cs.idy;
        s.cs.oidz = s.cs.oz * s.cs.idz;
    } else {
        setup_space(s.cs);
    }
    s.inst = i;
    s.inst_id = meta.x;
    s.inst_sp = sp;
    s.flags = (s.flags & ~(kTop | kInstLast)) | (meta.z == -1 ? kInstLast : 0);
    if ((unsigned)meta.y >= (unsigned)a.n_nodes) {
        s.flags |= kBad;
        return kNoNode;
    }
    return meta.y;
}

// A leaf link in s.node: top-level leaves enter their first instance, mesh leaves start
// their triangle list.
template <bool STATS>
__device__ __forceinline__ void resolve_leaf(const TraceArgs& a, TraceState& s, int sp) {
    if (s.node < 0 && s.node != kNoNode) {
        if (s.flags & kTop) {
            s.node = enter_instance<STATS>(a, s, ~s.node, sp);
        } else {
            s.tri = ~s.node;
            s.node = kNoNode;
        }
    }
}

// One inner node of the BVH2: both child boxes are tested, the nearer hit child (ties: left)
// is next and the other is pushed.
template <int BLOCK, int STACK_LDS, bool STATS>
__device__ __forceinline__ void node_step2(const TraceArgs& a, TraceState& s, LaneStack<BLOCK, STACK_LDS>& st) {
    const int n = s.node;
    if (n >= a.n_nodes) {
        s.flags |= kBad;
        s.node = kNoNode;
        return;
    }
    const auto* nodes = gview(a.nodes);
    const float4 n0 = nodes[4 * n + 0];
    const float4 n1 = nodes[4 * n + 1];
    const float4 n2 = nodes[4 * n + 2];
    const int4 n3 = gview(reinterpret_cast<const int4*>(a.nodes))[4 * n + 3];
    if (STATS) s.cn++;
    float tl = 0.0f, tr = 0.0f;
    bool hl = box_hit(s.cs, n0.x, n0.y, n0.z, n0.w, n1.x, n1.y, s.tmin, s.best_t, tl);
    bool hr = n3.y != kDummyChild && box_hit(s.cs, n1.z, n1.w, n2.x, n2.y, n2.z, n2.w, s.tmin, s.best_t, tr);
    if (hl && n3.x >= 0 && n3.x <= n) { hl = false; s.flags |= kBad; }
    if (hr && n3.y >= 0 && n3.y <= n) { hr = false; s.flags |= kBad; }
    if (hl && hr) {
        const bool left_first = tl <= tr;
        if (!st.push(left_first ? n3.y : n3.x)) s.flags |= kBad;
        if (STATS) s.cd = max(s.cd, (uint32_t)st.sp);
        s.node = left_first ? n3.x : n3.y;
    } else {
        s.node = hl ? n3.x : (hr ? n3.y : kNoNode);
    }
    resolve_leaf<STATS>(a, s, st.sp);
}

// Ordered exchange of two (distance, link) pairs: the nearer first, the lower slot on a tie.
__device__ __forceinline__ void order2(float& da, int& ca, float& db, int& cb) {
    const bool sw = db < da;
    const float d = sw ? db : da;
    const int c = sw ? cb : ca;
    db = sw ? da : db;
    cb = sw ? ca : cb;
    da = d;
    ca = c;
}

// One node of the 4-wide tree (ImbaNode4 layout: 8 float4 = min_x, max_x, min_y, max_y, min_z,
// max_z, children, count): the used child boxes are tested, the nearest hit child is next and
// the other hit children are pushed, farthest deepest.
template <int BLOCK, int STACK_LDS, bool STATS>
__device__ __forceinline__ void node_step4(const TraceArgs& a, TraceState& s, LaneStack<BLOCK, STACK_LDS>& st) {
    const int n = s.node;
    if (n >= a.n_nodes) {
        s.flags |= kBad;
        s.node = kNoNode;
        return;
    }
    const auto* p = gview(a.nodes + 8 * (size_t)n);
    const float4 lx = p[0], hx = p[1], ly = p[2], hy = p[3], lz = p[4], hz = p[5];
    const int4 ch = gview(reinterpret_cast<const int4*>(a.nodes + 8 * (size_t)n))[6];
    if (STATS) s.cn++;
    float d0, d1, d2, d3;
    bool h0 = ch.x != 0 && box_hit(s.cs, lx.x, hx.x, ly.x, hy.x, lz.x, hz.x, s.tmin, s.best_t, d0);
    bool h1 = ch.y != 0 && box_hit(s.cs, lx.y, hx.y, ly.y, hy.y, lz.y, hz.y, s.tmin, s.best_t, d1);
    bool h2 = ch.z != 0 && box_hit(s.cs, lx.z, hx.z, ly.z, hy.z, lz.z, hz.z, s.tmin, s.best_t, d2);
    bool h3 = ch.w != 0 && box_hit(s.cs, lx.w, hx.w, ly.w, hy.w, lz.w, hz.w, s.tmin, s.best_t, d3);
    if (h0 && ch.x > 0 && ch.x <= n) { h0 = false; s.flags |= kBad; }
    if (h1 && ch.y > 0 && ch.y <= n) { h1 = false; s.flags |= kBad; }
    if (h2 && ch.z > 0 && ch.z <= n) { h2 = false; s.flags |= kBad; }
    if (h3 && ch.w > 0 && ch.w <= n) { h3 = false; s.flags |= kBad; }
    const int cnt = (int)h0 + (int)h1 + (int)h2 + (int)h3;
    const float inf = __int_as_float(0x7f800000);
    float e0 = h0 ? d0 : inf, e1 = h1 ? d1 : inf, e2 = h2 ? d2 : inf, e3 = h3 ? d3 : inf;
    int c0 = ch.x, c1 = ch.y, c2 = ch.z, c3 = ch.w;
    // sorting network; misses (at infinity) end up last
    order2(e0, c0, e1, c1);
    order2(e2, c2, e3, c3);
    order2(e0, c0, e2, c2);
    order2(e1, c1, e3, c3);
    order2(e1, c1, e2, c2);
    if (cnt == 0) {
        s.node = kNoNode;
    } else {
        bool ok = true;
        if (cnt >= 4) ok &= st.push(c3);
        if (cnt >= 3) ok &= st.push(c2);
        if (cnt >= 2) ok &= st.push(c1);
        if (STATS) s.cd = max(s.cd, (uint32_t)st.sp);
        if (!ok) s.flags |= kBad;
        s.node = c0;
    }
    resolve_leaf<STATS>(a, s, st.sp);
}

// Child box j of a quantized 4-wide node, one axis: [fma(q_lo, 2^e, o), fma(q_hi, 2^e, o)].
__device__ __forceinline__ float dequant(uint32_t word, int j, float scale, float origin) {
    return fmaf((float)((word >> (8 * j)) & 255u), scale, origin);
}

// One node of the quantized 4-wide tree (64 B: origin, exponents, children, 8-bit child
// bounds per axis); otherwise as node_step4.  The dequantized boxes contain the exact ones.
template <int BLOCK, int STACK_LDS, bool STATS, bool TLK>
__device__ __forceinline__ void node_step4q(const TraceArgs& a, TraceState& s, LaneStack<BLOCK, STACK_LDS>& st) {
    const int n = s.node;
    if (n >= a.n_nodes) {
        s.flags |= kBad;
        s.node = kNoNode;
        return;
    }
    uint4 w0, w1, w2, w3;
    const unsigned in_tl = (unsigned)(n - a.tl_base);
    if (TLK && in_tl < (unsigned)a.tl_nodes) {
        const uint4* q = st.tl + 4 * in_tl;
        w0 = q[0]; w1 = q[1]; w2 = q[2]; w3 = q[3];
    } else if (TLK) {
        // merged with the branch above into one flat load by the compiler
        const uint4* p = reinterpret_cast<const uint4*>(a.nodes) + 4 * (size_t)n;
        w0 = p[0]; w1 = p[1]; w2 = p[2]; w3 = p[3];
    } else {
        const auto* p = gview(reinterpret_cast<const uint4*>(a.nodes) + 4 * (size_t)n);
        w0 = p[0]; w1 = p[1]; w2 = p[2]; w3 = p[3];
    }
    if (STATS) s.cn++;
    const float ox = __uint_as_float(w0.x), oy = __uint_as_float(w0.y), oz = __uint_as_float(w0.z);
    const float sx = __uint_as_float((uint32_t)((int)(int8_t)(w0.w & 255u) + 127) << 23);
    const float sy = __uint_as_float((uint32_t)((int)(int8_t)((w0.w >> 8) & 255u) + 127) << 23);
    const float sz = __uint_as_float((uint32_t)((int)(int8_t)((w0.w >> 16) & 255u) + 127) << 23);
    const int4 ch = make_int4((int)w1.x, (int)w1.y, (int)w1.z, (int)w1.w);
    float d[4];
    bool h[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int c = j == 0 ? ch.x : j == 1 ? ch.y : j == 2 ? ch.z : ch.w;
        h[j] = c != 0 && box_hit(s.cs, dequant(w2.x, j, sx, ox), dequant(w2.y, j, sx, ox), dequant(w2.z, j, sy, oy),
                                 dequant(w2.w, j, sy, oy), dequant(w3.x, j, sz, oz), dequant(w3.y, j, sz, oz), s.tmin,
                                 s.best_t, d[j]);
        if (h[j] && c > 0 && c <= n) { h[j] = false; s.flags |= kBad; }
    }
    const int cnt = (int)h[0] + (int)h[1] + (int)h[2] + (int)h[3];
    const float inf = __int_as_float(0x7f800000);
    float e0 = h[0] ? d[0] : inf, e1 = h[1] ? d[1] : inf, e2 = h[2] ? d[2] : inf, e3 = h[3] ? d[3] : inf;
    int c0 = ch.x, c1 = ch.y, c2 = ch.z, c3 = ch.w;
    order2(e0, c0, e1, c1);
    order2(e2, c2, e3, c3);
    order2(e0, c0, e2, c2);
    order2(e1, c1, e3, c3);
    order2(e1, c1, e2, c2);
    if (cnt == 0) {
        s.node = kNoNode;
    } else {
        bool ok = true;
        if (cnt >= 4) ok &= st.push(c3);
        if (cnt >= 3) ok &= st.push(c2);
        if (cnt >= 2) ok &= st.push(c1);
        if (STATS) s.cd = max(s.cd, (uint32_t)st.sp);
        if (!ok) s.flags |= kBad;
        s.node = c0;
    }
    resolve_leaf<STATS>(a, s, st.sp);
}

// Advances one ray by one step: bookkeeping (next instance, back to world space, pop) if
// nothing is pending, then one triangle test or one inner node.  Returns true when the ray
// is finished.
template <int BLOCK, int STACK_LDS, bool ANY, bool MASKED, bool STATS, int TREE, bool TLK = true>
__device__ __forceinline__ bool advance(const TraceArgs& a, TraceState& s, LaneStack<BLOCK, STACK_LDS>& st) {
    if (s.tri < 0 && s.node == kNoNode) {
        // the current instance's tree is exhausted: next instance of the leaf or world space
        while (!(s.flags & kTop) && st.sp == s.inst_sp) {
            if (s.flags & kInstLast) {
                s.flags |= kTop;
                to_world(s);
                s.inst_sp = -1;
                break;
            }
            s.node = enter_instance<STATS>(a, s, s.inst + 1, st.sp);
            if (s.node != kNoNode) break;
        }
        if (s.node == kNoNode) {
            if (st.sp == 0) return true;
            s.node = st.pop();
            resolve_leaf<STATS>(a, s, st.sp);
        }
    }
    if (s.tri >= 0) {
        const int t = s.tri;
        if (t + 2 >= a.n_tris4) {
            s.flags |= kBad;
            s.tri = -1;
            return false;
        }
        const auto* tris = gview(a.tris);
        const float4 v0 = tris[t], v1 = tris[t + 1], v2 = tris[t + 2];
        if (STATS) s.ct++;
        s.tri = (__float_as_uint(v2.w) == kLeafEnd) ? -1 : t + 3;
        float th, uh, vh;
        if (tri_hit(s.cs, v0, v1, v2, th, uh, vh) && th >= s.tmin) {
            const int tid = __float_as_int(v1.w);
            if (ANY) {
                if (th <= s.best_t) {
                    int op = 1;
                    if (MASKED) {
                        op = texel_opaque(a, tid, uh, vh);
                        if (op & 2) s.flags |= kBad;
                    }
                    if (op & 1) {
                        s.best_t = th; s.best_u = uh; s.best_tri = tid; s.best_inst = s.inst_id;
                        return true;
                    }
                }
            } else {
                const bool better = th < s.best_t ||
                    (th == s.best_t && ((unsigned)tid < (unsigned)s.best_tri ||
                                        (tid == s.best_tri && (unsigned)s.inst_id < (unsigned)s.best_inst)));
                if (better) {
                    int op = 1;
                    if (MASKED) {
                        op = texel_opaque(a, tid, uh, vh);
                        if (op & 2) s.flags |= kBad;
                    }
                    if (op & 1) {
                        s.best_t = th; s.best_u = uh; s.best_tri = tid; s.best_inst = s.inst_id;
                    }
                }
            }
        }
        return false;
    }
    if (s.node >= 0) {
        if (TREE == 2) node_step4q<BLOCK, STACK_LDS, STATS, TLK>(a, s, st);
        else if (TREE == 1) node_step4<BLOCK, STACK_LDS, STATS>(a, s, st);
        else node_step2<BLOCK, STACK_LDS, STATS>(a, s, st);
    }
    return false;
}

template <int BLOCK, int STACK_LDS>
__device__ __forceinline__ void stack_init(LaneStack<BLOCK, STACK_LDS>& st, int* lds, const TraceArgs& a) {
    st.lds = lds;
    st.spill = a.spill;
    st.spill_depth = a.spill_depth;
    st.lane_slot = threadIdx.x;
    st.spill_slot = blockIdx.x * BLOCK + threadIdx.x;
    st.spill_stride = gridDim.x * BLOCK;
    st.sp = 0;
    st.tl = nullptr;
}

// Copies the treelet nodes of a quantized-tree launch into the block's dynamic LDS (sized by
// the host for a.tl_nodes nodes).  a.tl_nodes is the same for every thread of the grid.  The
// copy keeps the node layout: a word-major copy (fewer LDS bank conflicts between lanes
// reading different nodes) needs a runtime stride that costs the kernel 7 VGPRs and one wave
// per SIMD, more than the conflicts cost.
template <int BLOCK, int STACK_LDS, int TREE>
__device__ __forceinline__ void treelet_init(const TraceArgs& a, LaneStack<BLOCK, STACK_LDS>& st) {
    extern __shared__ uint4 treelet_lds[];
    if (TREE != 2 || a.tl_nodes <= 0) return;
    const uint4* src = reinterpret_cast<const uint4*>(a.nodes) + 4 * (size_t)a.tl_base;
    for (int i = threadIdx.x; i < 4 * a.tl_nodes; i += BLOCK) treelet_lds[i] = src[i];
    __syncthreads();
    st.tl = treelet_lds;
}

template <bool STATS>
__device__ __forceinline__ void write_result(const TraceArgs& a, const TraceState& s, int id) {
    gview_w(a.hits)[id] = make_int4(s.best_tri, s.best_inst, __float_as_int(s.best_t), __float_as_int(s.best_u));
    if (STATS && a.counts) {
        a.counts[3 * (size_t)id + 0] = s.cn;
        a.counts[3 * (size_t)id + 1] = s.ct;
        a.counts[3 * (size_t)id + 2] = s.ci;
    }
    if (STATS && a.depth_hist) atomicAdd(&a.depth_hist[min(s.cd, 63u)], 1u);
    if (s.flags & kBad) atomicAdd(a.err, 1);
}

// One ray per thread (rays strided over the grid when it is smaller than the ray count).
// The 4-wide variants are held to 7 waves per SIMD (72 VGPRs, a few bytes of scratch per
// lane): faster than their natural 6 (77 VGPRs; PT frames 4% faster with the any-hit variant
// held too) and, for closest hit, than 8 (64 VGPRs, more spills: 9% slower).
template <int BLOCK, int STACK_LDS, bool ANY, bool MASKED, bool STATS, int TREE>
__global__ __launch_bounds__(BLOCK, (MASKED || STATS || TREE == 0) ? 1 : 7) void k_trace(TraceArgs a) {
    __shared__ int lds_stack[STACK_LDS * BLOCK];
    const int count = launch_count(a);
    // a block past the rays (a launch sized for the capacity of a queue whose size is read on
    // the device) leaves before it copies the treelet: block-uniform, before any barrier
    if ((int)(blockIdx.x * BLOCK) >= count) return;
    LaneStack<BLOCK, STACK_LDS> st;
    stack_init(st, lds_stack, a);
    treelet_init<BLOCK, STACK_LDS, TREE>(a, st);
    const int stride = gridDim.x * BLOCK;
    for (int id = blockIdx.x * BLOCK + threadIdx.x; id < count; id += stride) {
        TraceState s;
        state_init(s, a.rays[2 * (size_t)id], a.rays[2 * (size_t)id + 1], a.root);
        st.sp = 0;
        while (!advance<BLOCK, STACK_LDS, ANY, MASKED, STATS, TREE>(a, s, st)) {
        }
        write_result<STATS>(a, s, id);
    }
}

// Persistent variant: every wave keeps 64 rays in flight and, whenever at least `refill` of its
// lanes have finished, takes the next consecutive block of ray ids for them with one atomic
// on a global counter (ballot + popcount + mbcnt).  Each iteration advances every active lane
// by one step of `advance`, so the per-ray results are those of k_trace.
// counter[0] = next ray, counter[1] = finished waves; the last wave to finish resets both.
template <int BLOCK, int STACK_LDS, bool ANY, bool MASKED, bool STATS, int TREE>
__global__ __launch_bounds__(BLOCK) void k_trace_dyn(TraceArgs a, int refill) {
    __shared__ int lds_stack[STACK_LDS * BLOCK];
    LaneStack<BLOCK, STACK_LDS> st;
    stack_init(st, lds_stack, a);
    treelet_init<BLOCK, STACK_LDS, TREE>(a, st);
    const int lane = __lane_id();
    int* counter = a.work_counter;
    int id = -1;             // ray of this lane, -1 = none
    bool exhausted = false;  // wave-uniform: no rays left to fetch
    const int count = launch_count(a);
    TraceState s;
    for (;;) {
        const unsigned long long idle = __ballot(id < 0);
        if (!exhausted && __popcll(idle) >= refill) {
            const int n = __popcll(idle);
            const int leader = __ffsll((long long)idle) - 1;
            int base = 0;
            if (lane == leader) base = atomicAdd(&counter[0], n);
            base = __shfl(base, leader, 64);
            if (base + n >= count) exhausted = true;
            if (id < 0) {
                const int r = base + __builtin_amdgcn_mbcnt_hi((unsigned)(idle >> 32),
                                                               __builtin_amdgcn_mbcnt_lo((unsigned)idle, 0u));
                if (r < count) {
                    id = r;
                    state_init(s, a.rays[2 * (size_t)r], a.rays[2 * (size_t)r + 1], a.root);
                    st.sp = 0;
                }
            }
        }
        if (__ballot(id >= 0) == 0) {
            if (exhausted) break;
            continue;
        }
        if (id >= 0 && advance<BLOCK, STACK_LDS, ANY, MASKED, STATS, TREE>(a, s, st)) {
            write_result<STATS>(a, s, id);
            id = -1;
        }
    }
    // the last wave out resets the counters for the launch that reuses this slot
    if (lane == 0) {
        const int waves = gridDim.x * (BLOCK / 64);
        if (atomicAdd(&counter[1], 1) == waves - 1) {
            atomicExch(&counter[0], 0);
            atomicExch(&counter[1], 0);
        }
    }
}

// ---------------------------------------------------------------------------------------
// Batched traversal: many ray queues, each with its own scene, in one launch per kernel
// variant.  The queues of a launch are cut into chunks of 64 consecutive rays; every wave
// takes the next chunk from a global counter, finds its queue by binary search over the
// queues' first chunks (wave-uniform, so the queue's arguments sit in scalar registers) and
// traces one ray per lane with the same per-ray steps as k_trace, so every queue's hits are
// those of a launch of its own.
struct BatchItem {
    TraceArgs a;           // a.count: rays of this launch (set by k_batch_prefix)
    const int* count_dev;  // device ray count read when the launch runs, or null
    int capacity;          // rays (count_dev null) or the upper bound of *count_dev
    int chunk_begin;       // first chunk of this queue within its launch group
    int group;             // launch group: 2 * any_hit + masked
};

struct BatchGroups {
    int first[5];  // items of group g: [first[g], first[g + 1])
};

// One block of 1024 threads: ray counts of this launch, each queue's first chunk, chunks per group.
__global__ __launch_bounds__(1024) void k_batch_prefix(BatchItem* items, BatchGroups groups, int* group_chunks) {
    // thread t owns the queues [b0, b1); items are ordered by group, so a queue's chunk_begin is
    // its exclusive chunk prefix over all queues minus that of its group's first queue
    __shared__ int part[1024];
    __shared__ int base[5];
    const int n_items = groups.first[4];
    const int t = threadIdx.x;
    const int per = (n_items + 1023) / 1024;
    const int b0 = min(n_items, t * per), b1 = min(n_items, b0 + per);
    int sum = 0;
    for (int i = b0; i < b1; i++) {
        int n = items[i].capacity;
        if (items[i].count_dev) n = min(max(*items[i].count_dev, 0), n);
        items[i].a.count = n;
        sum += (n + 63) >> 6;
    }
    part[t] = sum;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {  // inclusive scan of the per-thread sums
        const int v = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    const int total = part[1023];
    int run = part[t] - sum;  // exclusive prefix at b0
    for (int i = b0; i < b1; i++) {
        for (int g = 0; g < 5; g++)
            if (groups.first[g] == i) base[g] = run;
        run += (items[i].a.count + 63) >> 6;
    }
    __syncthreads();
    run = part[t] - sum;
    int g = 0;
    for (int i = b0; i < b1; i++) {
        while (g < 3 && groups.first[g + 1] <= i) g++;
        items[i].chunk_begin = run - base[g];
        run += (items[i].a.count + 63) >> 6;
    }
    if (t < 4) {
        const int f0 = groups.first[t], f1 = groups.first[t + 1];
        group_chunks[t] = f0 == f1 ? 0 : (f1 < n_items ? base[t + 1] : total) - base[t];
    }
}

// TL (quantized tree, tl_max > 0): every wave keeps its own LDS copy of the first
// min(tl_max, a.tl_nodes) treelet nodes of the scene of its current chunk (tl_max x 64 B per
// wave of dynamic LDS), reloaded when a chunk of another scene comes.  The 4-wide variants
// are held to 7 waves per SIMD (72 VGPRs, a few bytes of scratch per lane), faster than their
// natural 6.
template <int BLOCK, int STACK_LDS, bool ANY, bool MASKED, int TREE, bool TL>
__global__ __launch_bounds__(BLOCK, MASKED ? 1 : (TREE == 0 && !ANY ? 8 : 7)) void k_trace_batch(const BatchItem* __restrict__ items, int first, int n_items,
                                                       const int* __restrict__ total_chunks, int* counter, int* spill,
                                                       int spill_depth, int tl_max) {
    extern __shared__ uint4 treelet_lds[];
    __shared__ int lds_stack[STACK_LDS * BLOCK];
    LaneStack<BLOCK, STACK_LDS> st;
    st.lds = lds_stack;
    st.spill = spill;
    st.spill_depth = spill_depth;
    st.lane_slot = threadIdx.x;
    st.spill_slot = blockIdx.x * BLOCK + threadIdx.x;
    st.spill_stride = gridDim.x * BLOCK;
    st.sp = 0;
    st.tl = TL ? treelet_lds + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6) * 4 * tl_max) : nullptr;
    const float4* tl_nodes_of = nullptr;  // scene (node array) whose treelet this wave holds
    int tl_base_of = -1;
    const int lane = __lane_id();
    const int total = __builtin_amdgcn_readfirstlane(*total_chunks);
    for (;;) {
        int c = 0;
        if (lane == 0) c = atomicAdd(&counter[0], 1);
        c = __builtin_amdgcn_readfirstlane(__shfl(c, 0, 64));
        if (c >= total) break;
        // the queue owning chunk c: the last one whose first chunk is <= c (queues without
        // rays share their first chunk with the next queue and are skipped)
        int lo = first, hi = first + n_items - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (__builtin_amdgcn_readfirstlane(items[mid].chunk_begin) <= c) lo = mid;
            else hi = mid - 1;
        }
        lo = __builtin_amdgcn_readfirstlane(lo);
        TraceArgs a = items[lo].a;
        if (TL) {
            a.tl_nodes = min(a.tl_nodes, tl_max);
            if (a.nodes != tl_nodes_of || a.tl_base != tl_base_of) {  // wave-uniform
                // the wave's earlier LDS reads are complete before the copy is overwritten
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                __builtin_amdgcn_wave_barrier();
                uint4* w = const_cast<uint4*>(st.tl);
                const uint4* src = reinterpret_cast<const uint4*>(a.nodes) + 4 * (size_t)a.tl_base;
                for (int i = lane; i < 4 * a.tl_nodes; i += 64) w[i] = src[i];
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                __builtin_amdgcn_wave_barrier();
                tl_nodes_of = a.nodes;
                tl_base_of = a.tl_base;
            }
        } else {
            a.tl_nodes = 0;
        }
        const int id = (c - items[lo].chunk_begin) * 64 + lane;
        if (id < a.count) {
            TraceState s;
            state_init(s, gview(a.rays)[2 * (size_t)id], gview(a.rays)[2 * (size_t)id + 1], a.root);
            st.sp = 0;
            while (!advance<BLOCK, STACK_LDS, ANY, MASKED, false, TREE, TL>(a, s, st)) {
            }
            write_result<false>(a, s, id);
        }
    }
    if (lane == 0) {
        const int waves = gridDim.x * (BLOCK / 64);
        if (atomicAdd(&counter[1], 1) == waves - 1) {
            atomicExch(&counter[0], 0);
            atomicExch(&counter[1], 0);
        }
    }
}

}  // namespace imba
