/*
 * Oracle: ray-queue compaction — TEST INFRASTRUCTURE ONLY.
 *
 * oracle_compact_hits restates RayQueue::compact_hits (render/scheduling/ray_queue.h:161-185):
 * one serial pass; the first miss position is remembered and every later hit is swapped
 * into it, so hits keep their order and misses end up permuted.
 * oracle_compact_rays restates RayQueue::compact_rays (ray_queue.h:188-207): survivors
 * (state.pixel_id >= 0, the first int of the state) are moved forward in order; hits are
 * not moved; the size only changes when at least one ray was removed.
 *
 * oracle_sort_by_material restates RayQueue::sort_by_material (ray_queue.h:211-244) on one
 * thread (queue order inside a material).
 *
 * Pinned by: brute force in tests/test_oracle_compact.py (a hand-worked 6-element example
 * from the reference's loop, the stable order of the hit prefix / survivors, the multiset
 * of the miss suffix, the returned counts, all-hit / all-miss / empty queues; for the material
 * sort a hand-worked example and the permutation / grouping / stability properties).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static void swap_bytes(void* a, void* b, size_t n) {
    unsigned char tmp[256];
    unsigned char* pa = (unsigned char*)a;
    unsigned char* pb = (unsigned char*)b;
    while (n > 0) {
        const size_t k = n < sizeof(tmp) ? n : sizeof(tmp);
        memcpy(tmp, pa, k);
        memcpy(pa, pb, k);
        memcpy(pb, tmp, k);
        pa += k; pb += k; n -= k;
    }
}

int oracle_compact_hits(ImbaRay* rays, ImbaHit* hits, void* states, int stride, int count) {
    unsigned char* st = (unsigned char*)states;
    int first_miss = -1;
    for (int i = 0; i < count; i++) {
        if (hits[i].tri_id < 0 && first_miss == -1) {
            first_miss = i;
        } else if (hits[i].tri_id >= 0 && first_miss != -1) {
            ImbaHit h = hits[first_miss]; hits[first_miss] = hits[i]; hits[i] = h;
            ImbaRay r = rays[first_miss]; rays[first_miss] = rays[i]; rays[i] = r;
            swap_bytes(st + (size_t)first_miss * stride, st + (size_t)i * stride, (size_t)stride);
            first_miss++;
        }
    }
    return first_miss != -1 ? first_miss : count;
}

int oracle_compact_rays(ImbaRay* rays, void* states, int stride, int count) {
    unsigned char* st = (unsigned char*)states;
    int first_dead = -1;
    for (int i = 0; i < count; i++) {
        int pixel;
        memcpy(&pixel, st + (size_t)i * stride, sizeof(int));
        if (pixel < 0 && first_dead == -1) {
            first_dead = i;
        } else if (pixel >= 0 && first_dead != -1) {
            memcpy(st + (size_t)first_dead * stride, st + (size_t)i * stride, (size_t)stride);
            rays[first_dead] = rays[i];
            first_dead++;
        }
    }
    return first_dead != -1 ? first_dead : count;
}

/*
 * RayQueue::sort_by_material (ray_queue.h:211-244) run on one thread: material of every hit
 * (indices[4 tri_id + 3], pt.cpp:85-91) written to dir.w, per-material counts, exclusive
 * scan, scatter in queue order.  With one thread the reference's scatter visits entries in
 * increasing index, which fixes the order inside a material.  Entries without a valid material
 * take key num_mats (last).  starts: num_mats + 1 ints (nullable).
 */
void oracle_sort_by_material(const int* indices, int tri_count, ImbaRay* rays, const ImbaHit* hits, int count,
                             int num_mats, int* sorted, int* starts) {
    int* cnt = (int*)calloc((size_t)num_mats + 2, sizeof(int));
    for (int i = 0; i < count; i++) {
        const int tri = hits[i].tri_id;
        int mat = -1;
        if (tri >= 0 && tri < tri_count) mat = indices[4 * (size_t)tri + 3];
        const int key = (mat >= 0 && mat < num_mats) ? mat : num_mats;
        memcpy(&rays[i].dir.w, &key, sizeof(int));
        cnt[key]++;
    }
    int accum = 0;
    for (int k = 0; k <= num_mats; k++) {
        const int c = cnt[k];
        cnt[k] = accum;
        if (starts) starts[k] = accum;
        accum += c;
    }
    for (int i = 0; i < count; i++) {
        int key;
        memcpy(&key, &rays[i].dir.w, sizeof(int));
        sorted[cnt[key]++] = i;
    }
    free(cnt);
}
