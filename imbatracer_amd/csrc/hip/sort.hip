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
// The sort is radix_sort_pairs (radix.h) with tiles of 1024 entries.
// Traffic per entry and pass: key and index read twice, written once (16 B), plus the hit
// (16 B), the index record (4 B) and the ray's dir.w (4 B) once for the keys.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.h"
#include "radix.h"

namespace imba {
namespace {

constexpr int kBlock = 256;
constexpr int kItems = 4;

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
    // scratch: keys and values twice, digit counts
    const size_t n4 = (size_t)count * 4;
    const size_t need = 4 * n4 + radix_counts_size<kItems>(count) * 4;
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
    radix_sort_pairs<kItems>(keys_a, vals_a, keys_b, vals_b, count, bits, counts, sorted, &keys_a, s);
    if (bucket_start)
        hipLaunchKernelGGL(k_bucket_starts, dim3(ceil_div(num_mats + 1, 256)), dim3(256), 0, s, keys_a, count,
                           num_mats, bucket_start);
    IMBA_HIP_TRY(hipGetLastError());
    if (sync) IMBA_HIP_TRY(hipStreamSynchronize(s));
    return IMBA_OK;
}

}  // extern "C"
