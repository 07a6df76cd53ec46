/*
 * CPU baseline: two-level BVH4 + 4-triangle-packet traversal — TEST INFRASTRUCTURE ONLY.
 *
 * Restates the reference's CPU path, RayQueue::traverse_cpu / traverse_occluded_cpu
 * (render/scheduling/ray_queue.h:246-265, 296-315), which calls the AnyDSL CPU kernels
 * intersect_cpu_masked_instanced / occluded_cpu_masked_instanced over the arrays of
 * core/cpu_adapter.cpp (BVH4 nodes :56-103, triangle packets :129-206, instance records
 * :339-355).  Those kernels are not part of the reference, so this is a single-ray traversal
 * with 4-wide SSE box and triangle tests (one child or one triangle per lane) and the
 * arithmetic of the GPU path, lane for lane:
 *   - slab test fma(bound, 1/d', -o/d') with d' = +-1e-9 where |d| < 1e-9; the entry plane of
 *     each axis is chosen by the sign of 1/d' (so an empty slot [FLT_MAX, -FLT_MAX] never
 *     overlaps); near = max(max(nx, ny), max(nz, tmin)), far = min(min(fx, fy), min(fz, t));
 *   - child 0 is an unused slot and is never entered (a child always follows its parent);
 *   - the precomputed-normal Moller-Trumbore test of the packets, with the same fma dot and
 *     cross products and divisions as the BVH2 oracle (trace.c);
 *   - closest hit: smallest t, ties to the smaller tri id, then the smaller instance id, so
 *     closest hits equal the BVH2 traversal's bit for bit; any hit: the first hit found
 *     (children nearest first, triangles in packet order), which may be another triangle
 *     than the BVH2 traversal's but agrees on whether the ray is occluded.
 *
 * Pinned by: tests/test_cpu_bvh4.py — closest hits identical to oracle_trace (itself pinned by
 * brute force), occlusion flags identical, on random, instanced and masked scenes; the
 * layout it reads is pinned against the reference's own BVH4 builders there as well.
 *
 * Built with -O3 -mavx2 -mfma -ffp-contract=off: _mm_fmadd_ps/_mm_fmsub_ps are single-rounding
 * fused operations, equal to the scalar fmaf of trace.c.
 */
#include <immintrin.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define STACK_SIZE 512

typedef struct {
    float o[3], d[3];
    __m128 ox, oy, oz, dx, dy, dz;        /* broadcast origin and direction */
    __m128 ix, iy, iz, oix, oiy, oiz;     /* 1/d' and o/d' */
    int negx, negy, negz;                 /* entry plane of each axis is the max plane */
} Space4;

typedef struct {
    float t, u;
    int tri, inst;
} Best4;

static void make_space4(Space4* s) {
    float inv[3], oinv[3];
    for (int k = 0; k < 3; k++) {
        float d = s->d[k];
        if (fabsf(d) < 1e-9f) d = copysignf(1e-9f, d);
        inv[k] = 1.0f / d;
        oinv[k] = s->o[k] * inv[k];
    }
    s->ox = _mm_set1_ps(s->o[0]); s->oy = _mm_set1_ps(s->o[1]); s->oz = _mm_set1_ps(s->o[2]);
    s->dx = _mm_set1_ps(s->d[0]); s->dy = _mm_set1_ps(s->d[1]); s->dz = _mm_set1_ps(s->d[2]);
    s->ix = _mm_set1_ps(inv[0]); s->iy = _mm_set1_ps(inv[1]); s->iz = _mm_set1_ps(inv[2]);
    s->oix = _mm_set1_ps(oinv[0]); s->oiy = _mm_set1_ps(oinv[1]); s->oiz = _mm_set1_ps(oinv[2]);
    s->negx = signbit(inv[0]) != 0;
    s->negy = signbit(inv[1]) != 0;
    s->negz = signbit(inv[2]) != 0;
}

/* Box test of the four child slots; returns the lane mask of entered children and their
 * entry distances. */
static inline int box4(const ImbaNode4* n, const Space4* s, __m128 tmin, __m128 tmax, __m128* tnear) {
    const __m128 nx = _mm_fmsub_ps(_mm_loadu_ps(s->negx ? n->max_x : n->min_x), s->ix, s->oix);
    const __m128 fx = _mm_fmsub_ps(_mm_loadu_ps(s->negx ? n->min_x : n->max_x), s->ix, s->oix);
    const __m128 ny = _mm_fmsub_ps(_mm_loadu_ps(s->negy ? n->max_y : n->min_y), s->iy, s->oiy);
    const __m128 fy = _mm_fmsub_ps(_mm_loadu_ps(s->negy ? n->min_y : n->max_y), s->iy, s->oiy);
    const __m128 nz = _mm_fmsub_ps(_mm_loadu_ps(s->negz ? n->max_z : n->min_z), s->iz, s->oiz);
    const __m128 fz = _mm_fmsub_ps(_mm_loadu_ps(s->negz ? n->min_z : n->max_z), s->iz, s->oiz);
    const __m128 tn = _mm_max_ps(_mm_max_ps(nx, ny), _mm_max_ps(nz, tmin));
    const __m128 tf = _mm_min_ps(_mm_min_ps(fx, fy), _mm_min_ps(fz, tmax));
    const __m128i used = _mm_xor_si128(_mm_cmpeq_epi32(_mm_loadu_si128((const __m128i*)n->children), _mm_setzero_si128()),
                                       _mm_set1_epi32(-1));
    *tnear = tn;
    return _mm_movemask_ps(_mm_and_ps(_mm_cmple_ps(tn, tf), _mm_castsi128_ps(used)));
}

/* Opacity of global triangle `tri` at barycentrics (u, v): the same texel rule as trace.c. */
static int opaque4(const ImbaCpuSceneDesc* sc, int tri, float u, float v, int* bad) {
    if (!sc->indices || !sc->masks || !sc->mask_buffer) return 1;
    if (tri < 0 || tri >= sc->index_count / 4) { *bad = 1; return 1; }
    const int* ind = sc->indices + 4 * (size_t)tri;
    const int m = ind[3];
    if (m < 0 || m >= sc->mask_count) { *bad = 1; return 1; }
    const ImbaTransparencyMask mk = sc->masks[m];
    long texel = 0;
    if (!(mk.width == 1 && mk.height == 1)) {
        for (int k = 0; k < 3; k++)
            if (ind[k] < 0 || ind[k] >= sc->texcoord_count) { *bad = 1; return 1; }
        const ImbaVec2 a = sc->texcoords[ind[0]], b = sc->texcoords[ind[1]], c = sc->texcoords[ind[2]];
        const float w0 = (1.0f - u) - v;
        const float s = fmaf(a.x, w0, fmaf(b.x, u, c.x * v));
        const float q = fmaf(a.y, w0, fmaf(b.y, u, c.y * v));
        const float fs = (s - floorf(s)) * (float)mk.width;
        const float fq = (q - floorf(q)) * (float)mk.height;
        int xi, yi;
        if (fs >= 0.0f && fs < (float)mk.width) xi = (int)fs; else xi = fs >= (float)mk.width ? mk.width - 1 : 0;
        if (fq >= 0.0f && fq < (float)mk.height) yi = (int)fq; else yi = fq >= (float)mk.height ? mk.height - 1 : 0;
        texel = (long)yi * mk.width + xi;
    }
    const long at = (long)mk.offset + texel;
    if (at < 0 || at >= sc->mask_buffer_size) { *bad = 1; return 1; }
    return sc->mask_buffer[at] != 0;
}

static int improves4(float t, int tri, int inst, const Best4* b) {
    if (t < b->t) return 1;
    if (t > b->t) return 0;
    if ((uint32_t)tri != (uint32_t)b->tri) return (uint32_t)tri < (uint32_t)b->tri;
    return (uint32_t)inst < (uint32_t)b->inst;
}

/* Tests the packets of one leaf; returns 1 when an any-hit query is answered. */
static int leaf4(const ImbaCpuSceneDesc* sc, const Space4* s, int first, float tmin, float tmax, int any, int inst_id,
                 Best4* best, int* bad) {
    const __m128i sentinel = _mm_set1_epi32((int)0x80000000u);
    for (int p = first;; p += 13) {
        if (p < 0 || p >= sc->tri_vec4_count) { *bad = 1; return 0; }
        const float* q = &sc->tris[p].x;
        const __m128i head = _mm_loadu_si128((const __m128i*)q);
        if (_mm_movemask_ps(_mm_castsi128_ps(_mm_cmpeq_epi32(head, sentinel))) == 0xF) return 0;
        if (p + 13 > sc->tri_vec4_count) { *bad = 1; return 0; }
        const __m128 v0x = _mm_loadu_ps(q + 0), v0y = _mm_loadu_ps(q + 4), v0z = _mm_loadu_ps(q + 8);
        const __m128 e1x = _mm_loadu_ps(q + 12), e1y = _mm_loadu_ps(q + 16), e1z = _mm_loadu_ps(q + 20);
        const __m128 e2x = _mm_loadu_ps(q + 24), e2y = _mm_loadu_ps(q + 28), e2z = _mm_loadu_ps(q + 32);
        const __m128 nx = _mm_loadu_ps(q + 36), ny = _mm_loadu_ps(q + 40), nz = _mm_loadu_ps(q + 44);
        const __m128 cx = _mm_sub_ps(v0x, s->ox), cy = _mm_sub_ps(v0y, s->oy), cz = _mm_sub_ps(v0z, s->oz);
        /* r = d x c, fma(a1, b2, -(a2 * b1)) and cyclic */
        const __m128 rx = _mm_fmsub_ps(s->dy, cz, _mm_mul_ps(s->dz, cy));
        const __m128 ry = _mm_fmsub_ps(s->dz, cx, _mm_mul_ps(s->dx, cz));
        const __m128 rz = _mm_fmsub_ps(s->dx, cy, _mm_mul_ps(s->dy, cx));
        /* dot(a, b) = fma(a0, b0, fma(a1, b1, a2 * b2)) */
        const __m128 det = _mm_fmadd_ps(nx, s->dx, _mm_fmadd_ps(ny, s->dy, _mm_mul_ps(nz, s->dz)));
        __m128 U = _mm_fmadd_ps(rx, e2x, _mm_fmadd_ps(ry, e2y, _mm_mul_ps(rz, e2z)));
        __m128 V = _mm_fmadd_ps(rx, e1x, _mm_fmadd_ps(ry, e1y, _mm_mul_ps(rz, e1z)));
        __m128 T = _mm_fmadd_ps(nx, cx, _mm_fmadd_ps(ny, cy, _mm_mul_ps(nz, cz)));
        const __m128 zero = _mm_setzero_ps();
        const __m128 sgn = _mm_and_ps(_mm_cmplt_ps(det, zero), _mm_set1_ps(-0.0f));
        U = _mm_xor_ps(U, sgn);
        V = _mm_xor_ps(V, sgn);
        T = _mm_xor_ps(T, sgn);
        const __m128 ad = _mm_andnot_ps(_mm_set1_ps(-0.0f), det);
        __m128 ok = _mm_and_ps(_mm_cmpge_ps(U, zero), _mm_cmpge_ps(V, zero));
        ok = _mm_and_ps(ok, _mm_cmpge_ps(_mm_sub_ps(_mm_sub_ps(ad, U), V), zero));
        ok = _mm_and_ps(ok, _mm_cmpneq_ps(det, zero));
        int mask = _mm_movemask_ps(ok);
        if (!mask) continue;
        float t[4], u[4], v[4];
        int ids[4];
        _mm_storeu_ps(t, _mm_div_ps(T, ad));
        _mm_storeu_ps(u, _mm_div_ps(U, ad));
        _mm_storeu_ps(v, _mm_div_ps(V, ad));
        memcpy(ids, q + 48, 16);
        for (int j = 0; j < 4; j++) {
            if (!(mask >> j & 1) || !(t[j] >= tmin)) continue;
            if (any) {
                if (t[j] <= tmax && opaque4(sc, ids[j], u[j], v[j], bad)) {
                    best->t = t[j]; best->u = u[j]; best->tri = ids[j]; best->inst = inst_id;
                    return 1;
                }
            } else if (improves4(t[j], ids[j], inst_id, best) && opaque4(sc, ids[j], u[j], v[j], bad)) {
                best->t = t[j]; best->u = u[j]; best->tri = ids[j]; best->inst = inst_id;
            }
        }
    }
}

/* Pushes the entered children of node `self` but the nearest, farther ones deeper (ties to
 * the lower slot first), and returns 1 with the nearest in *next; 0 when none is entered.  A child link that does not follow its parent is
 * malformed and is not entered. */
static inline int order_children(const ImbaNode4* n, int self, int mask, __m128 tn, int* stack, int* sp, int* next,
                                 int* bad) {
    float d[4];
    int c[4], k = 0;
    _mm_storeu_ps(d, tn);
    for (int j = 0; j < 4; j++)
        if (mask >> j & 1) {
            if (n->children[j] >= 0 && n->children[j] <= self) { *bad = 1; continue; }
            /* insertion by entry distance, stable in slot order */
            int p = k++;
            while (p > 0 && d[j] < d[c[p - 1]]) { c[p] = c[p - 1]; p--; }
            c[p] = j;
        }
    if (k == 0) return 0;
    for (int i = k - 1; i >= 1; i--) {
        if (*sp < STACK_SIZE) stack[(*sp)++] = n->children[c[i]];
        else *bad = 1;
    }
    *next = n->children[c[0]];
    return 1;
}

/* Traverses one BVH4 from `root` (mesh level when leaves hold triangle packets); returns 1
 * when an any-hit query is answered. */
static int trace_tree(const ImbaCpuSceneDesc* sc, const Space4* s, int root, float tmin, float tmax, int any,
                      int inst_id, Best4* best, int* bad, int top, const Space4* world);

static int instance_leaf(const ImbaCpuSceneDesc* sc, const Space4* world, int first, float tmin, float tmax, int any,
                         Best4* best, int* bad) {
    for (int i = first;; i++) {
        if (i < 0 || i >= sc->instance_count) { *bad = 1; return 0; }
        const ImbaInstanceNode* in = &sc->instances[i];
        const ImbaVec4* m = in->transf;
        const float rows[3][4] = {{m[0].x, m[0].y, m[0].z, m[0].w},
                                  {m[1].x, m[1].y, m[1].z, m[1].w},
                                  {m[2].x, m[2].y, m[2].z, m[2].w}};
        Space4 obj;
        for (int k = 0; k < 3; k++) {
            obj.o[k] = fmaf(rows[k][0], world->o[0], fmaf(rows[k][1], world->o[1], fmaf(rows[k][2], world->o[2], rows[k][3])));
            obj.d[k] = fmaf(rows[k][0], world->d[0], fmaf(rows[k][1], world->d[1], rows[k][2] * world->d[2]));
        }
        make_space4(&obj);
        if (in->next >= 0 && in->next < sc->node_count) {
            if (trace_tree(sc, &obj, in->next, tmin, tmax, any, in->id, best, bad, 0, world)) return 1;
        } else {
            *bad = 1;
        }
        if (in->pad[0] == -1) return 0;
    }
}

static int trace_tree(const ImbaCpuSceneDesc* sc, const Space4* s, int root, float tmin, float tmax, int any,
                      int inst_id, Best4* best, int* bad, int top, const Space4* world) {
    int stack[STACK_SIZE];
    int sp = 0;
    int node = root;
    const __m128 vmin = _mm_set1_ps(tmin);
    for (;;) {
        if (node >= 0) {
            if (node >= sc->node_count) {
                *bad = 1;
                goto pop;
            }
            const ImbaNode4* n = &sc->nodes[node];
            __m128 tn;
            const int mask = box4(n, s, vmin, _mm_set1_ps(best->t), &tn);
            if (!order_children(n, node, mask, tn, stack, &sp, &node, bad)) goto pop;
            continue;
        } else {
            const int leaf = ~node;
            if (top) {
                if (instance_leaf(sc, world, leaf, tmin, tmax, any, best, bad)) return 1;
            } else if (leaf4(sc, s, leaf, tmin, tmax, any, inst_id, best, bad)) {
                return 1;
            }
        }
    pop:
        if (sp == 0) return 0;
        node = stack[--sp];
    }
}

static int trace_one4(const ImbaCpuSceneDesc* sc, const ImbaRay* ray, int any, ImbaHit* out) {
    Space4 world;
    world.o[0] = ray->org.x; world.o[1] = ray->org.y; world.o[2] = ray->org.z;
    world.d[0] = ray->dir.x; world.d[1] = ray->dir.y; world.d[2] = ray->dir.z;
    make_space4(&world);
    Best4 best = {ray->dir.w, 0.0f, -1, -1};
    int bad = 0;
    if (sc->root < 0 || sc->root >= sc->node_count) bad = 1;
    else trace_tree(sc, &world, sc->root, ray->org.w, ray->dir.w, any, -1, &best, &bad, 1, &world);
    out->tri_id = best.tri;
    out->inst_id = best.inst;
    out->tmax = best.t;
    out->u = best.u;
    return bad;
}

int cpu4_trace(const ImbaCpuSceneDesc* scene, const ImbaRay* rays, int count, int any_hit, ImbaHit* hits) {
    int bad = 0;
    for (int i = 0; i < count; i++) bad += trace_one4(scene, &rays[i], any_hit, &hits[i]);
    return bad;
}

typedef struct {
    const ImbaCpuSceneDesc* scene;
    const ImbaRay* rays;
    ImbaHit* hits;
    int count, any, bad;
    int* next;
} Job4;

#define CHUNK 256

static void* worker4(void* p) {
    Job4* j = (Job4*)p;
    for (;;) {
        const int b = __atomic_fetch_add(j->next, CHUNK, __ATOMIC_RELAXED);
        if (b >= j->count) break;
        const int e = b + CHUNK < j->count ? b + CHUNK : j->count;
        for (int i = b; i < e; i++) j->bad += trace_one4(j->scene, &j->rays[i], j->any, &j->hits[i]);
    }
    return NULL;
}

int cpu4_trace_mt(const ImbaCpuSceneDesc* scene, const ImbaRay* rays, int count, int any_hit, ImbaHit* hits,
                  int threads) {
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t th[256];
    Job4 jobs[256];
    int next = 0;
    for (int t = 0; t < threads; t++) {
        jobs[t] = (Job4){scene, rays, hits, count, any_hit, 0, &next};
        pthread_create(&th[t], NULL, worker4, &jobs[t]);
    }
    int bad = 0;
    for (int t = 0; t < threads; t++) {
        pthread_join(th[t], NULL);
        bad += jobs[t].bad;
    }
    return bad;
}
