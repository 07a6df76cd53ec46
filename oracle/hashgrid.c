/*
 * Oracle: photon hash grid — TEST INFRASTRUCTURE ONLY.
 *
 * Serial restatement of HashGrid (rangesearch/rangesearch.h):
 *   build  :25-75   cell = 2r; bounds of the photon positions grown by 0.1% of the extents
 *                   on each side; bucket table size max(previous, 2 * count) (grow only);
 *                   counting sort: count per bucket, exclusive scan, scatter (the reference's
 *                   atomic scatter visits photons in input order when run serially, so a
 *                   bucket lists its photons by increasing index).
 *   query  :77-129  early out when the position is outside the grown bounds; the bucket of
 *                   the position's cell and the buckets of the 7 neighbours towards the
 *                   nearer half in x, y, z (cells may share a bucket: its photons are then
 *                   visited twice); keep photons with |q - p|^2 <= r^2, insertion-sorted by
 *                   distance and capped at k; when full a candidate is dropped only if the
 *                   current k-th distance is strictly smaller, otherwise it replaces the
 *                   k-th entry and moves before every strictly larger distance.
 *   hash   :137-148 ((x * 73856093) ^ (y * 19349663) ^ (z * 83492791)) mod table size, on
 *                   uint32 of the floored cell coordinates.
 *
 * Pinned by: brute force over all photons on small sets (tests/test_oracle_hashgrid.py):
 * every photon within r of the query whose cell is among the 8 probed cells is returned
 * (with multiplicity when buckets coincide), distances are sorted, the k cap keeps the k
 * nearest with the stated tie rule, queries outside the bounds return nothing; the table
 * size history; a hand-computed hash value.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static int bucket_of(const OracleGrid* g, uint32_t x, uint32_t y, uint32_t z) {
    return (int)(((x * 73856093u) ^ (y * 19349663u) ^ (z * 83492791u)) % (uint32_t)g->table_size);
}

static int bucket_of_point(const OracleGrid* g, const float* p) {
    int c[3];
    for (int k = 0; k < 3; k++) c[k] = (int)floorf(g->inv_cell_size * (p[k] - g->bbox_min[k]));
    return bucket_of(g, (uint32_t)c[0], (uint32_t)c[1], (uint32_t)c[2]);
}

void oracle_grid_build(OracleGrid* g, const float* pos, int stride, int count, float radius) {
    g->radius = radius;
    g->radius_sqr = radius * radius;
    g->cell_size = radius * 2.0f;
    g->inv_cell_size = 1.0f / g->cell_size;
    if (g->table_size < count * 2) {
        free(g->cell_ends);
        g->table_size = count * 2;
        g->cell_ends = (int*)malloc(sizeof(int) * (size_t)(g->table_size > 0 ? g->table_size : 1));
    }
    g->photon_count = count;
    free(g->order);
    free(g->positions);
    g->order = (int*)malloc(sizeof(int) * (size_t)(count > 0 ? count : 1));
    g->positions = (float*)malloc(sizeof(float) * 3 * (size_t)(count > 0 ? count : 1));

    float lo[3] = {3.402823466e38f, 3.402823466e38f, 3.402823466e38f};
    float hi[3] = {-3.402823466e38f, -3.402823466e38f, -3.402823466e38f};
    for (int i = 0; i < count; i++) {
        const float* p = pos + (size_t)i * stride;
        for (int k = 0; k < 3; k++) {
            lo[k] = lo[k] < p[k] ? lo[k] : p[k];
            hi[k] = hi[k] > p[k] ? hi[k] : p[k];
        }
    }
    for (int k = 0; k < 3; k++) {
        const float e = hi[k] - lo[k];
        g->bbox_max[k] = hi[k] + e * 0.001f;
        g->bbox_min[k] = lo[k] - e * 0.001f;
    }

    memset(g->cell_ends, 0, sizeof(int) * (size_t)g->table_size);
    int* bucket = (int*)malloc(sizeof(int) * (size_t)(count > 0 ? count : 1));
    for (int i = 0; i < count; i++) {
        bucket[i] = bucket_of_point(g, pos + (size_t)i * stride);
        g->cell_ends[bucket[i]]++;
    }
    int sum = 0;
    for (int c = 0; c < g->table_size; c++) {
        const int n = g->cell_ends[c];
        g->cell_ends[c] = sum;
        sum += n;
    }
    for (int i = 0; i < count; i++) {
        const int slot = g->cell_ends[bucket[i]]++;
        g->order[slot] = i;
        memcpy(g->positions + 3 * (size_t)slot, pos + (size_t)i * stride, 3 * sizeof(float));
    }
    free(bucket);
}

int oracle_grid_query(const OracleGrid* g, const float* q, int k, int* out_idx, float* out_dist) {
    if (k <= 0 || g->photon_count <= 0) return 0;
    for (int a = 0; a < 3; a++)
        if (!(q[a] >= g->bbox_min[a] && q[a] <= g->bbox_max[a])) return 0;
    float cell[3], fl[3];
    int base[3], other[3];
    for (int a = 0; a < 3; a++) {
        cell[a] = g->inv_cell_size * (q[a] - g->bbox_min[a]);
        fl[a] = floorf(cell[a]);
        base[a] = (int)fl[a];
        other[a] = base[a] + ((cell[a] - fl[a]) < 0.5f ? -1 : 1);
    }
    int count = 0;
    for (int j = 0; j < 8; j++) {
        const int x = (j & 4) ? other[0] : base[0];
        const int y = (j & 2) ? other[1] : base[1];
        const int z = (j & 1) ? other[2] : base[2];
        const int b = bucket_of(g, (uint32_t)x, (uint32_t)y, (uint32_t)z);
        const int begin = b == 0 ? 0 : g->cell_ends[b - 1];
        const int end = g->cell_ends[b];
        for (int s = begin; s < end; s++) {
            const float* p = g->positions + 3 * (size_t)s;
            const float dx = q[0] - p[0], dy = q[1] - p[1], dz = q[2] - p[2];
            const float d2 = dx * dx + dy * dy + dz * dz;
            if (!(d2 <= g->radius_sqr)) continue;
            if (count == k) {
                if (out_dist[count - 1] < d2) continue;
            } else {
                count++;
            }
            int l = count - 1;
            while (l > 0 && out_dist[l - 1] > d2) {
                out_dist[l] = out_dist[l - 1];
                out_idx[l] = out_idx[l - 1];
                l--;
            }
            out_dist[l] = d2;
            out_idx[l] = g->order[s];
        }
    }
    return count;
}

void oracle_grid_free(OracleGrid* g) {
    free(g->cell_ends);
    free(g->order);
    free(g->positions);
    memset(g, 0, sizeof(*g));
}

/* The query above for n positions (query q at q_base + q * stride_floats), one thread; results
 * at out_idx[q * k], out_dist[q * k] and out_count[q].  Used as the serial CPU baseline. */
void oracle_grid_query_many(const OracleGrid* g, const float* q_base, int stride_floats, int n, int k, int* out_idx,
                            float* out_dist, int* out_count) {
    for (int q = 0; q < n; q++)
        out_count[q] = oracle_grid_query(g, q_base + (size_t)q * stride_floats, k, out_idx + (size_t)q * k,
                                         out_dist + (size_t)q * k);
}
