/*
 * Oracle: two-level BVH2 traversal on the CPU — TEST INFRASTRUCTURE ONLY.
 *
 * Restates, for one ray at a time, what the reference's GPU traversal computes over the
 * arrays written by core/gpu_adapter.cpp (nodes :62-103, triangles :129-141 with the leaf end
 * flagged by 0x80000000 in v2.w, dummy parent :14-31, instance records :276-295) and the
 * calling convention of render/scheduling/ray_queue.h:269-343 (Ray org.w = tmin,
 * dir.w = tmax; Hit tri_id/inst_id/tmax/u, tri_id < 0 = no hit).  The AnyDSL kernel itself
 * is not part of the reference, so the arithmetic and traversal order are fixed rules shared
 * with the HIP kernel (see the header of trace_kernels.h and DESIGN.md):
 *   - slab test with fma and a reciprocal of the direction (|d| < 1e-9 -> +-1e-9);
 *   - edge-inclusive triangle test with a precomputed-normal Moller-Trumbore form, fma dot
 *     and cross products, t = T/|det|, u = U/|det| (u = weight of v1, as integrator.h:89-94
 *     expects);
 *   - closest hit: smallest t, ties to the smaller tri id, then the smaller instance id;
 *   - any hit: the first hit in traversal order (near child first, ties to the left child).
 *
 * Pinned by: brute force in double precision over the same triangles (oracle_brute_closest,
 * tests/test_oracle_trace.py), closed-form single-triangle cases (edges, vertices, tmin/tmax
 * boundaries, axis-parallel rays, det = 0), instancing invariants (identity instance equals
 * no instance; a translated instance equals translated geometry), and visit-count invariants
 * (every triangle reachable by a ray aimed at it).
 *
 * Compiled with -ffp-contract=off: only the explicit fmaf calls fuse.
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define MAX_STACK 512
#define DUMMY_CHILD 0x76543210
#define NO_NODE INT32_MIN

typedef struct {
    float o[3], d[3], inv[3], oinv[3];
} Space;

static uint32_t float_to_bits(float f) { uint32_t b; memcpy(&b, &f, 4); return b; }

static void make_space(Space* s) {
    for (int k = 0; k < 3; k++) {
        float d = s->d[k];
        if (fabsf(d) < 1e-9f) d = copysignf(1e-9f, d);
        s->inv[k] = 1.0f / d;
        s->oinv[k] = s->o[k] * s->inv[k];
    }
}

/* Child box test; returns 1 and the entry distance when the ray overlaps [tmin, tmax]. */
static int slab(const Space* s, const float lo[3], const float hi[3], float tmin, float tmax, float* near_out) {
    float a[3], b[3];
    for (int k = 0; k < 3; k++) {
        a[k] = fmaf(lo[k], s->inv[k], -s->oinv[k]);
        b[k] = fmaf(hi[k], s->inv[k], -s->oinv[k]);
    }
    const float tn = fmaxf(fmaxf(fminf(a[0], b[0]), fminf(a[1], b[1])), fmaxf(fminf(a[2], b[2]), tmin));
    const float tf = fminf(fminf(fmaxf(a[0], b[0]), fmaxf(a[1], b[1])), fminf(fmaxf(a[2], b[2]), tmax));
    *near_out = tn;
    return tn <= tf;
}

static void xprod(const float a[3], const float b[3], float r[3]) {
    r[0] = fmaf(a[1], b[2], -(a[2] * b[1]));
    r[1] = fmaf(a[2], b[0], -(a[0] * b[2]));
    r[2] = fmaf(a[0], b[1], -(a[1] * b[0]));
}
static float dprod(const float a[3], const float b[3]) { return fmaf(a[0], b[0], fmaf(a[1], b[1], a[2] * b[2])); }

/* Edge-inclusive ray/triangle test on the stored vertices. */
static int triangle(const Space* s, const ImbaVec4* v, float* t, float* u, float* w_v2) {
    const float p0[3] = {v[0].x, v[0].y, v[0].z};
    const float p1[3] = {v[1].x, v[1].y, v[1].z};
    const float p2[3] = {v[2].x, v[2].y, v[2].z};
    float e1[3], e2[3], c[3], n[3], r[3];
    for (int k = 0; k < 3; k++) {
        e1[k] = p0[k] - p1[k];
        e2[k] = p2[k] - p0[k];
        c[k] = p0[k] - s->o[k];
    }
    xprod(e1, e2, n);
    xprod(s->d, c, r);
    const float det = dprod(n, s->d);
    float U = dprod(r, e2), V = dprod(r, e1), T = dprod(n, c);
    if (det < 0.0f) { U = -U; V = -V; T = -T; }
    const float ad = fabsf(det);
    if (!(U >= 0.0f && V >= 0.0f && (ad - U) - V >= 0.0f && det != 0.0f)) return 0;
    *t = T / ad;
    *u = U / ad;
    *w_v2 = V / ad;
    return 1;
}

/* Material-mask opacity at barycentrics (u, v) of global triangle `tri`. */
static int opaque(const ImbaSceneDesc* sc, int tri, float u, float v, int* bad) {
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

typedef struct {
    float t, u;
    int tri, inst;
} Best;

static int improves(float t, int tri, int inst, const Best* b) {
    if (t < b->t) return 1;
    if (t > b->t) return 0;
    if ((uint32_t)tri != (uint32_t)b->tri) return (uint32_t)tri < (uint32_t)b->tri;
    return (uint32_t)inst < (uint32_t)b->inst;
}

/* Enters the next instance of the current top-level leaf whose root is valid; returns that
 * root, or NO_NODE when the leaf has no further instance. */
static int enter_next_instance(const ImbaSceneDesc* sc, const Space* world, Space* cur, int* inst, int* inst_id,
                               int* inst_last, uint32_t* n_inst, int* bad) {
    while (!*inst_last) {
        (*inst)++;
        if (*inst < 0 || *inst >= sc->instance_count) { *bad = 1; *inst_last = 1; break; }
        const ImbaInstanceNode* in = &sc->instances[*inst];
        const ImbaVec4* m = in->transf;
        const float rows[3][4] = {{m[0].x, m[0].y, m[0].z, m[0].w},
                                  {m[1].x, m[1].y, m[1].z, m[1].w},
                                  {m[2].x, m[2].y, m[2].z, m[2].w}};
        (*n_inst)++;
        for (int k = 0; k < 3; k++) {
            cur->o[k] = fmaf(rows[k][0], world->o[0], fmaf(rows[k][1], world->o[1], fmaf(rows[k][2], world->o[2], rows[k][3])));
            cur->d[k] = fmaf(rows[k][0], world->d[0], fmaf(rows[k][1], world->d[1], rows[k][2] * world->d[2]));
        }
        make_space(cur);
        *inst_id = in->id;
        *inst_last = in->pad[0] == -1;
        if (in->next >= 0 && in->next < sc->node_count) return in->next;
        *bad = 1;
    }
    return NO_NODE;
}

/* One ray.  Returns 1 when the structure was malformed somewhere on the ray's path. */
static int trace_one(const ImbaSceneDesc* sc, const ImbaRay* ray, int any, ImbaHit* out, uint32_t* cnt) {
    int stack[MAX_STACK];
    int sp = 0, bad = 0;
    Space world, cur;
    world.o[0] = ray->org.x; world.o[1] = ray->org.y; world.o[2] = ray->org.z;
    world.d[0] = ray->dir.x; world.d[1] = ray->dir.y; world.d[2] = ray->dir.z;
    make_space(&world);
    cur = world;
    const float tmin = ray->org.w, tmax = ray->dir.w;
    Best best = {tmax, 0.0f, -1, -1};
    uint32_t n_nodes = 0, n_tris = 0, n_inst = 0;

    /* in_instance: the ray is in object space of instance `inst`, whose tree started when
     * the stack held inst_base_sp entries. */
    int in_instance = 0, inst_base_sp = -1, inst = -1, inst_id = -1, inst_last = 1;
    int node = sc->root;

    for (;;) {
        if (node >= 0) {
            if (node >= sc->node_count) {
                bad = 1;
                node = NO_NODE;
            } else {
                const ImbaNode* n = &sc->nodes[node];
                n_nodes++;
                const float llo[3] = {n->left_bb.lo_x, n->left_bb.lo_y, n->left_bb.lo_z};
                const float lhi[3] = {n->left_bb.hi_x, n->left_bb.hi_y, n->left_bb.hi_z};
                const float rlo[3] = {n->right_bb.lo_x, n->right_bb.lo_y, n->right_bb.lo_z};
                const float rhi[3] = {n->right_bb.hi_x, n->right_bb.hi_y, n->right_bb.hi_z};
                float nl = 0.0f, nr = 0.0f;
                int hl = slab(&cur, llo, lhi, tmin, best.t, &nl);
                int hr = n->right != DUMMY_CHILD && slab(&cur, rlo, rhi, tmin, best.t, &nr);
                if (hl && n->left >= 0 && n->left <= node) { hl = 0; bad = 1; }
                if (hr && n->right >= 0 && n->right <= node) { hr = 0; bad = 1; }
                if (hl && hr) {
                    const int first = nl <= nr ? n->left : n->right;
                    const int second = nl <= nr ? n->right : n->left;
                    if (sp < MAX_STACK) stack[sp++] = second; else bad = 1;
                    node = first;
                } else if (hl) {
                    node = n->left;
                } else if (hr) {
                    node = n->right;
                } else {
                    node = NO_NODE;
                }
            }
        } else if (node != NO_NODE) {
            const int leaf = ~node;
            if (!in_instance) {
                /* top-level leaf: a list of instance records */
                in_instance = 1;
                inst_base_sp = sp;
                inst = leaf - 1;
                inst_last = 0;
                node = enter_next_instance(sc, &world, &cur, &inst, &inst_id, &inst_last, &n_inst, &bad);
                if (node != NO_NODE) continue;
            } else {
                int finished = 0;
                for (int t = leaf;; t += 3) {
                    if (t < 0 || t + 2 >= sc->tri_vec4_count) { bad = 1; break; }
                    const ImbaVec4* v = &sc->tris[t];
                    n_tris++;
                    float th, uh, vh;
                    if (triangle(&cur, v, &th, &uh, &vh) && th >= tmin) {
                        int32_t tri;
                        const uint32_t b = float_to_bits(v[1].w);
                        memcpy(&tri, &b, 4);
                        if (any) {
                            if (th <= tmax && opaque(sc, tri, uh, vh, &bad)) {
                                best.t = th; best.u = uh; best.tri = tri; best.inst = inst_id;
                                finished = 1;
                                break;
                            }
                        } else if (improves(th, tri, inst_id, &best) && opaque(sc, tri, uh, vh, &bad)) {
                            best.t = th; best.u = uh; best.tri = tri; best.inst = inst_id;
                        }
                    }
                    if (float_to_bits(v[2].w) == IMBA_LEAF_END_BITS) break;
                }
                if (finished) break;
                node = NO_NODE;
            }
        }

        if (node != NO_NODE) continue;
        /* The current instance's tree is exhausted: next instance, or back to world space. */
        if (in_instance && sp == inst_base_sp) {
            node = enter_next_instance(sc, &world, &cur, &inst, &inst_id, &inst_last, &n_inst, &bad);
            if (node != NO_NODE) continue;
            in_instance = 0;
            inst_base_sp = -1;
            cur = world;
        }
        if (sp == 0) break;
        node = stack[--sp];
    }

    out->tri_id = best.tri;
    out->inst_id = best.inst;
    out->tmax = best.t;
    out->u = best.u;
    if (cnt) { cnt[0] = n_nodes; cnt[1] = n_tris; cnt[2] = n_inst; }
    return bad;
}

int oracle_trace(const ImbaSceneDesc* scene, const ImbaRay* rays, int count, int any_hit, ImbaHit* hits,
                 uint32_t* counts) {
    int bad = 0;
    for (int i = 0; i < count; i++) bad += trace_one(scene, &rays[i], any_hit, &hits[i], counts ? counts + 3 * (size_t)i : NULL);
    return bad;
}

typedef struct {
    const ImbaSceneDesc* scene;
    const ImbaRay* rays;
    ImbaHit* hits;
    int count, any, tid, threads, bad;
} Job;

static void* worker(void* p) {
    Job* j = (Job*)p;
    for (int b = j->tid * 64; b < j->count; b += j->threads * 64) {
        const int e = b + 64 < j->count ? b + 64 : j->count;
        for (int i = b; i < e; i++) j->bad += trace_one(j->scene, &j->rays[i], j->any, &j->hits[i], NULL);
    }
    return NULL;
}

int oracle_trace_mt(const ImbaSceneDesc* scene, const ImbaRay* rays, int count, int any_hit, ImbaHit* hits,
                    int threads) {
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t th[256];
    Job jobs[256];
    for (int t = 0; t < threads; t++) {
        jobs[t] = (Job){scene, rays, hits, count, any_hit, t, threads, 0};
        pthread_create(&th[t], NULL, worker, &jobs[t]);
    }
    int bad = 0;
    for (int t = 0; t < threads; t++) {
        pthread_join(th[t], NULL);
        bad += jobs[t].bad;
    }
    return bad;
}

/* Double-precision Moller-Trumbore (1997) over world-space triangles, edges inclusive. */
void oracle_brute_closest(const float* tv, int tri_count, const ImbaRay* rays, int count, int* tri_out, double* t_out) {
    for (int i = 0; i < count; i++) {
        const double o[3] = {rays[i].org.x, rays[i].org.y, rays[i].org.z};
        const double d[3] = {rays[i].dir.x, rays[i].dir.y, rays[i].dir.z};
        const double tmin = rays[i].org.w, tmax = rays[i].dir.w;
        double best = INFINITY;
        int best_id = -1;
        for (int k = 0; k < tri_count; k++) {
            const float* p = tv + 9 * (size_t)k;
            const double v0[3] = {p[0], p[1], p[2]}, v1[3] = {p[3], p[4], p[5]}, v2[3] = {p[6], p[7], p[8]};
            double e1[3], e2[3], s[3], h[3], q[3];
            for (int c = 0; c < 3; c++) { e1[c] = v1[c] - v0[c]; e2[c] = v2[c] - v0[c]; s[c] = o[c] - v0[c]; }
            h[0] = d[1] * e2[2] - d[2] * e2[1];
            h[1] = d[2] * e2[0] - d[0] * e2[2];
            h[2] = d[0] * e2[1] - d[1] * e2[0];
            const double a = e1[0] * h[0] + e1[1] * h[1] + e1[2] * h[2];
            if (a == 0.0) continue;
            const double f = 1.0 / a;
            const double u = f * (s[0] * h[0] + s[1] * h[1] + s[2] * h[2]);
            if (u < 0.0 || u > 1.0) continue;
            q[0] = s[1] * e1[2] - s[2] * e1[1];
            q[1] = s[2] * e1[0] - s[0] * e1[2];
            q[2] = s[0] * e1[1] - s[1] * e1[0];
            const double v = f * (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]);
            if (v < 0.0 || u + v > 1.0) continue;
            const double t = f * (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]);
            if (t < tmin || t > tmax) continue;
            if (t < best) { best = t; best_id = k; }
        }
        tri_out[i] = best_id;
        t_out[i] = best;
    }
}
