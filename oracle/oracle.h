/*
 * CPU oracle — TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C restatement of the reference's hot path, used by tests/, by the smoke check and by
 * the CPU-baseline leg of bench.py to check or time the HIP library.  Nothing in the library
 * (imbatracer_amd/) links, loads or calls this code.
 *
 * Pins (what each function is checked against) are listed in the header of each .c file.
 */
#ifndef IMBA_ORACLE_H
#define IMBA_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "imba_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- traversal (trace.c) ---- */
/* Closest hit (any_hit = 0) or first hit (any_hit = 1) of rays[0..count) against the scene;
 * counts (nullable): 3 per ray (nodes, triangles, instances).  Returns the number of rays
 * that met a malformed structure. */
int oracle_trace(const ImbaSceneDesc* scene, const ImbaRay* rays, int count, int any_hit, ImbaHit* hits,
                 uint32_t* counts);
/* Same, split over `threads` host threads (rays interleaved in blocks of 64). */
int oracle_trace_mt(const ImbaSceneDesc* scene, const ImbaRay* rays, int count, int any_hit, ImbaHit* hits,
                    int threads);

/* ---- CPU baseline (cpu_bvh4.c): the reference's CPU layout, BVH4 + 4-triangle packets ---- */
/* Closest hit or any hit over an ImbaCpuSceneDesc (imba_build_cpu_layout); returns the number
 * of rays that met a malformed structure. */
int cpu4_trace(const ImbaCpuSceneDesc* scene, const ImbaRay* rays, int count, int any_hit, ImbaHit* hits);
/* Same on `threads` host threads taking chunks of 256 rays from a shared counter. */
int cpu4_trace_mt(const ImbaCpuSceneDesc* scene, const ImbaRay* rays, int count, int any_hit, ImbaHit* hits,
                  int threads);

/* Brute force in double precision over explicit world-space triangles (9 floats each,
 * ids = index).  Closest t in [tmin, tmax] (edges inclusive), ties to the lower id. */
void oracle_brute_closest(const float* tri_verts, int tri_count, const ImbaRay* rays, int count,
                          int* tri_out, double* t_out);

/* ---- queue compaction (compact.c) ---- */
/* RayQueue::compact_hits exactly as the reference runs it (serial swap partition); returns
 * the hit count. Arrays are modified in place. */
int oracle_compact_hits(ImbaRay* rays, ImbaHit* hits, void* states, int state_stride, int count);
/* RayQueue::compact_rays: returns the new queue size (count if nothing was removed). */
int oracle_compact_rays(ImbaRay* rays, void* states, int state_stride, int count);

/* RayQueue::sort_by_material on one thread: dir.w = material, `sorted` grouped by material in
 * queue order, starts[k] = first position of material k (num_mats + 1 entries, nullable). */
void oracle_sort_by_material(const int* indices, int tri_count, ImbaRay* rays, const ImbaHit* hits, int count,
                             int num_mats, int* sorted, int* starts);

/* ---- photon hash grid (hashgrid.c) ---- */
typedef struct OracleGrid {
    int photon_count;
    int table_size;
    float bbox_min[3], bbox_max[3];
    float radius, radius_sqr, cell_size, inv_cell_size;
    int* cell_ends;    /* table_size */
    int* order;        /* photon_count: photon index per slot */
    float* positions;  /* 3 per slot */
} OracleGrid;

/* Serial HashGrid::build; table grows as max(prev table_size, 2 * count). */
void oracle_grid_build(OracleGrid* g, const float* positions, int stride_floats, int count, float radius);
/* HashGrid::query; writes up to k photon indices (original numbering) in result order. */
int oracle_grid_query(const OracleGrid* g, const float* pos, int k, int* out_idx, float* out_dist);
void oracle_grid_query_many(const OracleGrid* g, const float* q_base, int stride_floats, int n, int k,
                            int* out_idx, float* out_dist, int* out_count);
void oracle_grid_free(OracleGrid* g);

#ifdef __cplusplus
}
#endif

#endif
