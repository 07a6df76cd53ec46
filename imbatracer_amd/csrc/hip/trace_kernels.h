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
    uint32_t* counts;     // 3 per ray (stats variant)
    int* err;             // incremented once per ray that met a malformed structure
    int* spill;           // global stack entries beyond the LDS part
    int spill_depth;      // spill entries per thread
    int* work_counter;    // persistent variant: next ray to fetch
};

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
// scale by the mask size, truncate, clamp to the last texel).
__device__ __forceinline__ bool texel_opaque(const TraceArgs& a, int tri, float u, float v, bool& bad) {
    if ((unsigned)tri >= (unsigned)a.n_tri_ids) { bad = true; return true; }
    const int4 ind = a.indices[tri];
    if ((unsigned)ind.w >= (unsigned)a.n_masks) { bad = true; return true; }
    const int w = a.masks[3 * ind.w + 0], h = a.masks[3 * ind.w + 1], off = a.masks[3 * ind.w + 2];
    int texel = 0;
    if (!(w == 1 && h == 1)) {
        if ((unsigned)ind.x >= (unsigned)a.n_texcoords || (unsigned)ind.y >= (unsigned)a.n_texcoords ||
            (unsigned)ind.z >= (unsigned)a.n_texcoords) { bad = true; return true; }
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
    if (at < 0 || at >= a.mask_buf_size) { bad = true; return true; }
    return a.mask_buf[at] != 0;
}

template <int BLOCK, int STACK_LDS>
struct LaneStack {
    int* lds;     // [STACK_LDS][BLOCK]
    int* spill;   // [spill_depth][threads] or null
    int spill_depth;
    int lane_slot;   // threadIdx.x
    int spill_slot;  // global thread index
    int spill_stride;
    int sp;

    __device__ __forceinline__ bool push(int v) {
        if (sp < STACK_LDS) {
            lds[sp * BLOCK + lane_slot] = v;
        } else {
            const int k = sp - STACK_LDS;
            if (k >= spill_depth) return false;
            spill[(size_t)k * spill_stride + spill_slot] = v;
        }
        sp++;
        return true;
    }
    __device__ __forceinline__ int pop() {
        sp--;
        if (sp < STACK_LDS) return lds[sp * BLOCK + lane_slot];
        return spill[(size_t)(sp - STACK_LDS) * spill_stride + spill_slot];
    }
};

struct HitRecord {
    float t, u;
    int tri, inst;
};

// Traverses one ray. Returns the hit record; visit counts are accumulated in c_*.
template <int BLOCK, int STACK_LDS, bool ANY, bool MASKED, bool STATS>
__device__ __forceinline__ HitRecord trace_ray(const TraceArgs& a, float4 ro, float4 rd, LaneStack<BLOCK, STACK_LDS>& st,
                                               uint32_t& c_nodes, uint32_t& c_tris, uint32_t& c_inst, bool& bad) {
    RaySpace world;
    world.ox = ro.x; world.oy = ro.y; world.oz = ro.z;
    world.dx = rd.x; world.dy = rd.y; world.dz = rd.z;
    setup_space(world);
    RaySpace cs = world;
    const float tmin = ro.w;
    const float tmax = rd.w;

    HitRecord best;
    best.t = tmax;
    best.u = 0.0f;
    best.tri = -1;
    best.inst = -1;

    int node = a.root;
    bool top = true;
    int inst_sp = -1;
    int inst = -1, inst_id = -1;
    bool inst_last = true;
    st.sp = 0;

    auto enter_instance = [&](int i) -> int {
        if ((unsigned)i >= (unsigned)a.n_inst) { bad = true; inst_last = true; return kNoNode; }
        const float4 m0 = a.inst[4 * i + 0], m1 = a.inst[4 * i + 1], m2 = a.inst[4 * i + 2];
        const int4 meta = reinterpret_cast<const int4*>(a.inst)[4 * i + 3];
        if (STATS) c_inst++;
        cs.ox = fmaf(m0.x, world.ox, fmaf(m0.y, world.oy, fmaf(m0.z, world.oz, m0.w)));
        cs.oy = fmaf(m1.x, world.ox, fmaf(m1.y, world.oy, fmaf(m1.z, world.oz, m1.w)));
        cs.oz = fmaf(m2.x, world.ox, fmaf(m2.y, world.oy, fmaf(m2.z, world.oz, m2.w)));
        cs.dx = fmaf(m0.x, world.dx, fmaf(m0.y, world.dy, m0.z * world.dz));
        cs.dy = fmaf(m1.x, world.dx, fmaf(m1.y, world.dy, m1.z * world.dz));
        cs.dz = fmaf(m2.x, world.dx, fmaf(m2.y, world.dy, m2.z * world.dz));
        setup_space(cs);
        inst = i;
        inst_id = meta.x;
        inst_last = meta.z == -1;
        inst_sp = st.sp;
        top = false;
        if ((unsigned)meta.y >= (unsigned)a.n_nodes) { bad = true; return kNoNode; }
        return meta.y;
    };

    for (;;) {
        if (node >= 0) {
            if (node >= a.n_nodes) {
                bad = true;
                node = kNoNode;
            } else {
                const float4 n0 = a.nodes[4 * node + 0];
                const float4 n1 = a.nodes[4 * node + 1];
                const float4 n2 = a.nodes[4 * node + 2];
                const int4 n3 = reinterpret_cast<const int4*>(a.nodes)[4 * node + 3];
                if (STATS) c_nodes++;
                float tl = 0.0f, tr = 0.0f;
                bool hl = box_hit(cs, n0.x, n0.y, n0.z, n0.w, n1.x, n1.y, tmin, best.t, tl);
                bool hr = n3.y != kDummyChild && box_hit(cs, n1.z, n1.w, n2.x, n2.y, n2.z, n2.w, tmin, best.t, tr);
                if (hl && n3.x >= 0 && n3.x <= node) { hl = false; bad = true; }
                if (hr && n3.y >= 0 && n3.y <= node) { hr = false; bad = true; }
                if (hl && hr) {
                    const bool left_first = tl <= tr;
                    if (!st.push(left_first ? n3.y : n3.x)) bad = true;
                    node = left_first ? n3.x : n3.y;
                } else if (hl) {
                    node = n3.x;
                } else if (hr) {
                    node = n3.y;
                } else {
                    node = kNoNode;
                }
            }
        } else if (node != kNoNode) {
            const int leaf = ~node;
            if (top) {
                node = enter_instance(leaf);
                if (node != kNoNode) continue;
            } else {
                bool done = false;
                for (int t = leaf;; t += 3) {
                    if (t + 2 >= a.n_tris4) { bad = true; break; }
                    const float4 v0 = a.tris[t], v1 = a.tris[t + 1], v2 = a.tris[t + 2];
                    if (STATS) c_tris++;
                    float th, uh, vh;
                    if (tri_hit(cs, v0, v1, v2, th, uh, vh) && th >= tmin) {
                        const int tri = __float_as_int(v1.w);
                        if (ANY) {
                            if (th <= tmax && (!MASKED || texel_opaque(a, tri, uh, vh, bad))) {
                                best.t = th; best.u = uh; best.tri = tri; best.inst = inst_id;
                                done = true;
                                break;
                            }
                        } else {
                            const bool better = th < best.t ||
                                (th == best.t && ((unsigned)tri < (unsigned)best.tri ||
                                                  (tri == best.tri && (unsigned)inst_id < (unsigned)best.inst)));
                            if (better && (!MASKED || texel_opaque(a, tri, uh, vh, bad))) {
                                best.t = th; best.u = uh; best.tri = tri; best.inst = inst_id;
                            }
                        }
                    }
                    if (__float_as_uint(v2.w) == kLeafEnd) break;
                }
                if (done) break;
                node = kNoNode;
            }
        }

        if (node == kNoNode) {
            // The current instance's tree is exhausted: enter the next instance of the same
            // top-level leaf, or return to the top level in world space.
            while (!top && st.sp == inst_sp) {
                if (inst_last) {
                    top = true;
                    cs = world;
                    inst_sp = -1;
                    break;
                }
                node = enter_instance(inst + 1);
                if (node != kNoNode) break;
            }
            if (node != kNoNode) continue;
            if (st.sp == 0) break;
            node = st.pop();
        }
    }
    return best;
}

// One ray per thread, rays strided over a grid sized to the resident waves.
template <int BLOCK, int STACK_LDS, bool ANY, bool MASKED, bool STATS>
__global__ __launch_bounds__(BLOCK) void k_trace(TraceArgs a) {
    __shared__ int lds_stack[STACK_LDS * BLOCK];
    LaneStack<BLOCK, STACK_LDS> st;
    st.lds = lds_stack;
    st.spill = a.spill;
    st.spill_depth = a.spill_depth;
    st.lane_slot = threadIdx.x;
    st.spill_slot = blockIdx.x * BLOCK + threadIdx.x;
    st.spill_stride = gridDim.x * BLOCK;
    st.sp = 0;
    const int stride = gridDim.x * BLOCK;
    for (int id = blockIdx.x * BLOCK + threadIdx.x; id < a.count; id += stride) {
        const float4 ro = a.rays[2 * (size_t)id];
        const float4 rd = a.rays[2 * (size_t)id + 1];
        uint32_t cn = 0, ct = 0, ci = 0;
        bool bad = false;
        const HitRecord h = trace_ray<BLOCK, STACK_LDS, ANY, MASKED, STATS>(a, ro, rd, st, cn, ct, ci, bad);
        a.hits[id] = make_int4(h.tri, h.inst, __float_as_int(h.t), __float_as_int(h.u));
        if (STATS) {
            a.counts[3 * (size_t)id + 0] = cn;
            a.counts[3 * (size_t)id + 1] = ct;
            a.counts[3 * (size_t)id + 2] = ci;
        }
        if (bad) atomicAdd(a.err, 1);
    }
}

}  // namespace imba
