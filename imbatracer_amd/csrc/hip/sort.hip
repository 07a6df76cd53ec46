// RayQueue::sort_by_material (render/scheduling/ray_queue.h:211-244) on the GPU.
//
// The reference counts hits per material with atomics, scans the counts on one thread and
// scatters the queue indices with a parallel loop, so the order inside a material is whatever
// the threads produce.  Here the key of entry i is the material of its hit triangle
// (indices[4 tri_id + 3], the lookup of pt.cpp:85-91 and vcm.cpp:199-205), written to
// rays[i].dir.w as the reference does, and the indices are sorted by key with a stable LSD
// radix sort of 8-bit digits: entries of one material stay in queue order.  Entries without a
// valid material (a miss, a triangle id or material out of range) get key num_mats and end up
// after all materials.
//
// One pass over a digit:
//   1. k_digit_hist: per tile of 1024 entries, a 256-bin histogram (LDS atomics), stored
//      digit-major so that one exclusive scan gives every (digit, tile) its first position;
//   2. k_scan: exclusive scan of the 256 x tiles counts (one block);
//   3. k_digit_scatter: per tile, in four chunks of 256 entries in order, each wave ranks its
//      lanes among the lanes with the same digit (eight ballots, one per digit bit), the waves'
//      per-digit counts in LDS give the offsets of the earlier waves, and a running per-digit
//      count carries over between chunks.
// Traffic per entry and pass: key and index read twice, written once (16 B), plus the hit
// (16 B), the index record (4 B) and the ray's dir.w (4 B) once for the keys.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.h"

namespace imba {
namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kItems = 4;
constexpr int kTile = kBlock * kItems;

__global__ __launch_bounds__(kBlock) void k_mat_keys(const int4* __restrict__ hits, const int* __restrict__ indices,
                                                     int n_tris, float4* rays, int count, int num_mats, int* keys,
                                                     int* vals) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= count) return;
    const int tri = hits[i].x;
    int mat = -1;
    if (tri >= 0 && tri < n_tris) mat = indices[4 * (size_t)tri + 3];
    const int key = (mat >= 0 && mat < num_mats) ? mat : num_mats;
    reinterpret_cast<float*>(rays)[8 * (size_t)i + 7] = __int_as_float(key);  // ray.dir.w
    keys[i] = key;
    vals[i] = i;
}

__global__ __launch_bounds__(kBlock) void k_digit_hist(const int* __restrict__ keys, int n, int shift, int n_tiles,
                                                       int* counts) {
    __shared__ int hist[256];
    hist[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * kTile;
#pragma unroll
    for (int it = 0; it < kItems; it++) {
        const int i = base + it * kBlock + (int)threadIdx.x;
        if (i < n) atomicAdd(&hist[(keys[i] >> shift) & 255], 1);
    }
    __syncthreads();
    counts[(size_t)threadIdx.x * n_tiles + blockIdx.x] = hist[threadIdx.x];
}

// Exclusive scan of m ints in place (one block of 1024 threads, each over a contiguous part).
__global__ __launch_bounds__(1024) void k_scan(int* v, int m) {
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int per = (m + 1023) / 1024;
    const int b0 = min(m, t * per), b1 = min(m, b0 + per);
    int s = 0;
    for (int b = b0; b < b1; b++) s += v[b];
    part[t] = s;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        const int x = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += x;
        __syncthreads();
    }
    int run = part[t] - s;
    for (int b = b0; b < b1; b++) {
        const int c = v[b];
        v[b] = run;
        run += c;
    }
}

__global__ __launch_bounds__(kBlock) void k_digit_scatter(const int* __restrict__ keys_in, const int* __restrict__ vals_in,
                                                          int n, int shift, int n_tiles, const int* __restrict__ offsets,
                                                          int* keys_out, int* vals_out) {
    __shared__ int wave_count[kWaves][256];
    __shared__ int running[256];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const unsigned long long lt = lane ? (~0ull >> (64 - lane)) : 0ull;  // lanes below this one
    running[threadIdx.x] = offsets[(size_t)threadIdx.x * n_tiles + blockIdx.x];
    const int base = blockIdx.x * kTile;
    for (int it = 0; it < kItems; it++) {
        for (int w = 0; w < kWaves; w++) wave_count[w][threadIdx.x] = 0;
        __syncthreads();
        const int i = base + it * kBlock + (int)threadIdx.x;
        const bool valid = i < n;
        const int key = valid ? keys_in[i] : 0;
        const int val = valid ? vals_in[i] : 0;
        const int digit = (key >> shift) & 255;
        // lanes of this wave with the same digit
        unsigned long long same = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const unsigned long long bits = __ballot((digit >> b) & 1);
            same &= ((digit >> b) & 1) ? bits : ~bits;
        }
        const int rank = __popcll(same & lt);
        if (valid && rank == 0) wave_count[wave][digit] = __popcll(same);
        __syncthreads();
        if (valid) {
            int before = 0;
            for (int w = 0; w < wave; w++) before += wave_count[w][digit];
            const int pos = running[digit] + before + rank;
            keys_out[pos] = key;
            vals_out[pos] = val;
        }
        __syncthreads();
        int add = 0;
        for (int w = 0; w < kWaves; w++) add += wave_count[w][threadIdx.x];
        running[threadIdx.x] += add;
        __syncthreads();
    }
}

// bucket_start[k] = first position whose key is >= k, for k = 0..num_mats.
__global__ void k_bucket_starts(const int* __restrict__ keys, int n, int num_mats, int* starts) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > num_mats) return;
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (keys[mid] < k) lo = mid + 1;
        else hi = mid;
    }
    starts[k] = lo;
}

thread_local Scratch t_sort[64];

}  // namespace
}  // namespace imba

using namespace imba;

extern "C" {

int imba_sort_by_material(const int* indices, int tri_count, ImbaRay* rays, const ImbaHit* hits, int count,
                          int num_mats, int* sorted, int* bucket_start, void* stream) {
    IMBA_CHECK_ARG(count >= 0 && tri_count >= 0, "negative count");
    IMBA_CHECK_ARG(num_mats > 0 && num_mats < (1 << 24), "num_mats must be in [1, 2^24)");
    IMBA_CHECK_ARG(count == 0 || (indices && rays && hits && sorted), "NULL argument");
    bool sync;
    hipStream_t s = resolve_stream(stream, &sync);
    if (count == 0) {
        if (bucket_start) {
            hipLaunchKernelGGL(k_bucket_starts, dim3(ceil_div(num_mats + 1, 256)), dim3(256), 0, s, (const int*)sorted,
                               0, num_mats, bucket_start);
            IMBA_HIP_TRY(hipGetLastError());
            if (sync) IMBA_HIP_TRY(hipStreamSynchronize(s));
        }
        return IMBA_OK;
    }
    int bits = 1;
    while ((1 << bits) <= num_mats) bits++;  // keys go up to num_mats
    const int passes = (bits + 7) / 8;
    const int n_tiles = ceil_div(count, kTile);
    // scratch: keys and values twice, digit counts
    const size_t n4 = (size_t)count * 4;
    const size_t need = 4 * n4 + (size_t)256 * n_tiles * 4;
    int dev = 0;
    IMBA_HIP_TRY(hipGetDevice(&dev));
    Scratch& sc = t_sort[dev & 63];
    int r = grow_buffer(sc, need);
    if (r) return r;
    int* keys_a = (int*)sc.ptr;
    int* keys_b = keys_a + count;
    int* vals_a = keys_b + count;
    int* vals_b = vals_a + count;
    int* counts = vals_b + count;
    hipLaunchKernelGGL(k_mat_keys, dim3(ceil_div(count, kBlock)), dim3(kBlock), 0, s, (const int4*)hits, indices,
                       tri_count, (float4*)rays, count, num_mats, keys_a, vals_a);
    for (int p = 0; p < passes; p++) {
        const int shift = 8 * p;
        const bool last = p == passes - 1;
        int* ko = keys_b;
        int* vo = last ? sorted : vals_b;
        hipLaunchKernelGGL(k_digit_hist, dim3(n_tiles), dim3(kBlock), 0, s, keys_a, count, shift, n_tiles, counts);
        hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, counts, 256 * n_tiles);
        hipLaunchKernelGGL(k_digit_scatter, dim3(n_tiles), dim3(kBlock), 0, s, keys_a, vals_a, count, shift, n_tiles,
                           counts, ko, vo);
        std::swap(keys_a, keys_b);
        std::swap(vals_a, vals_b);
    }
    if (bucket_start)
        hipLaunchKernelGGL(k_bucket_starts, dim3(ceil_div(num_mats + 1, 256)), dim3(256), 0, s, keys_a, count,
                           num_mats, bucket_start);
    IMBA_HIP_TRY(hipGetLastError());
    if (sync) IMBA_HIP_TRY(hipStreamSynchronize(s));
    return IMBA_OK;
}

}  // extern "C"
