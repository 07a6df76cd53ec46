// Stable ray-queue compaction on the GPU.
//
// RayQueue::compact_hits (render/scheduling/ray_queue.h:161-185) and compact_rays (:188-207)
// run as a serial loop over the queue on one host thread in the reference.  Here the queue
// is compacted out of place by three launches:
//   1. count:   each block counts the kept entries of its tile (wave ballot + popcount);
//   2. scan:    one block turns the block counts into exclusive offsets and the total;
//   3. scatter: each block re-derives the rank of every kept entry inside its tile
//               (ballot, mbcnt, wave totals in LDS) and moves ray / hit / state records.
// Kept entries keep their order.  compact_hits also writes the removed (missed) entries
// after the kept ones, in order; the reference's swap loop leaves that tail permuted, and
// callers only use the hit prefix (they shrink the queue to the returned count).
// Traffic per entry: key read twice (4 B), record read and written once (32 + 16 + S B).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "common.h"

namespace imba {
namespace {

constexpr int kBlock = 256;
constexpr int kItems = 8;
constexpr int kTile = kBlock * kItems;
constexpr int kWaves = kBlock / 64;

enum { kModeHits = 0, kModeRays = 1 };

struct CompactArgs {
    const float4* rays_in;
    const int4* hits_in;
    const int* states_in;
    int stride_dw;  // state record size in dwords
    int n;               // entries (or, with count_dev, the capacity)
    const int* count_dev;  // device entry count read when the launch runs, or null
    float4* rays_out;
    int4* hits_out;
    int* states_out;
    int* block_offsets;  // nblocks exclusive offsets, then the total at [nblocks]
    int* count_out;      // device copy of the total, or null
};

// Entries of this launch: n, or the device count clamped to [0, n].
__device__ __forceinline__ int entries(const CompactArgs& a) {
    return a.count_dev ? min(max(*a.count_dev, 0), a.n) : a.n;
}

template <int MODE>
__device__ __forceinline__ bool kept(const CompactArgs& a, int i) {
    if (MODE == kModeHits) return a.hits_in[i].x >= 0;
    return a.states_in[(size_t)i * a.stride_dw] >= 0;
}

__device__ __forceinline__ int lane_rank(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

template <int MODE>
__global__ __launch_bounds__(kBlock) void k_count(CompactArgs a) {
    __shared__ int wave_sum[kWaves];
    const int base = blockIdx.x * kTile;
    const int n = entries(a);
    int cnt = 0;
#pragma unroll
    for (int it = 0; it < kItems; it++) {
        const int i = base + it * kBlock + (int)threadIdx.x;
        const bool f = i < n && kept<MODE>(a, i);
        cnt += __popcll(__ballot(f));
    }
    // every lane of a wave holds the same count
    if ((threadIdx.x & 63) == 0) wave_sum[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        int s = 0;
        for (int w = 0; w < kWaves; w++) s += wave_sum[w];
        a.block_offsets[blockIdx.x] = s;
    }
}

// Exclusive scan of nb block counts in place; the total goes to block_offsets[nb] (and to
// *count_out when given).
__global__ __launch_bounds__(1024) void k_scan_offsets(int* offs, int nb, int* count_out) {
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int per = (nb + 1023) / 1024;
    const int b0 = t * per;
    const int b1 = min(nb, b0 + per);
    int s = 0;
    for (int b = b0; b < b1; b++) s += offs[b];
    part[t] = s;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        const int v = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - s;  // exclusive prefix of this thread's chunk
    for (int b = b0; b < b1; b++) {
        const int c = offs[b];
        offs[b] = run;
        run += c;
    }
    if (t == 1023) {
        offs[nb] = part[1023];
        if (count_out) *count_out = part[1023];
    }
}

template <int MODE>
__device__ __forceinline__ void move_record(const CompactArgs& a, int src, int dst) {
    a.rays_out[2 * (size_t)dst] = a.rays_in[2 * (size_t)src];
    a.rays_out[2 * (size_t)dst + 1] = a.rays_in[2 * (size_t)src + 1];
    if (MODE == kModeHits) a.hits_out[dst] = a.hits_in[src];
    const int* s = a.states_in + (size_t)src * a.stride_dw;
    int* d = a.states_out + (size_t)dst * a.stride_dw;
    if (a.stride_dw == 4) {
        *reinterpret_cast<int4*>(d) = *reinterpret_cast<const int4*>(s);
    } else {
        for (int k = 0; k < a.stride_dw; k++) d[k] = s[k];
    }
}

template <int MODE>
__global__ __launch_bounds__(kBlock) void k_scatter(CompactArgs a, int nblocks) {
    __shared__ int wave_cnt[kWaves];
    const int base = blockIdx.x * kTile;
    const int n = entries(a);
    const int total = a.block_offsets[nblocks];
    int kept_before = a.block_offsets[blockIdx.x];  // kept entries before this chunk
    int removed_before = base - kept_before;         // removed entries before this chunk
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    for (int it = 0; it < kItems; it++) {
        const int chunk = base + it * kBlock;
        const int i = chunk + (int)threadIdx.x;
        const bool valid = i < n;
        const bool f = valid && kept<MODE>(a, i);
        const unsigned long long m = __ballot(f);
        const int r = lane_rank(m);
        if (lane == 0) wave_cnt[wave] = __popcll(m);
        __syncthreads();
        int woff = 0, chunk_kept = 0;
#pragma unroll
        for (int w = 0; w < kWaves; w++) {
            const int c = wave_cnt[w];
            woff += w < wave ? c : 0;
            chunk_kept += c;
        }
        if (f) {
            move_record<MODE>(a, i, kept_before + woff + r);
        } else if (valid && MODE == kModeHits) {
            const int kept_in_chunk_before = woff + r;
            const int removed_in_chunk_before = (int)threadIdx.x - kept_in_chunk_before;
            move_record<MODE>(a, i, total + removed_before + removed_in_chunk_before);
        }
        const int chunk_valid = max(0, min(kBlock, n - chunk));
        kept_before += chunk_kept;
        removed_before += chunk_valid - chunk_kept;
        __syncthreads();
    }
}

thread_local Scratch t_offsets[64];

// Block offsets of the allocation-free compaction (imba_compact_hits_noalloc): one array per
// device for the life of the process, kNoAllocMaxBlocks + 1 ints (256 KiB + 4 B), enough for
// 134,217,728 entries.  Host code never passes this name to a kernel (on the host it is only a
// stand-in for the symbol); it uses the device address hipGetSymbolAddress gave for the
// current device.
constexpr int kNoAllocMaxBlocks = 65536;
static_assert((long long)kNoAllocMaxBlocks * kTile == IMBA_COMPACT_NOALLOC_MAX_COUNT, "header limit");
static __device__ __align__(256) int scratch[kNoAllocMaxBlocks + 1];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int* g_scratch_ptr[kMaxDevices] = {};

// The current device's copy of `scratch`, looked up once per device.
int scratch_for_device(int dev, int** out) {
    if (dev < 0 || dev >= kMaxDevices) return fail(IMBA_ERR_INVALID_ARG, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        IMBA_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(IMBA_ERR_HIP, "hipGetSymbolAddress returned no address for the compaction scratch");
        g_scratch_ptr[dev] = (int*)p;
    }
    *out = g_scratch_ptr[dev];
    return IMBA_OK;
}

template <int MODE>
int run_compact(CompactArgs a, int* new_count, void* stream, bool static_scratch = false) {
    if (a.n < 0) return fail(IMBA_ERR_INVALID_ARG, "count < 0");
    if (a.stride_dw <= 0) return fail(IMBA_ERR_INVALID_ARG, "state stride must be a positive multiple of 4 bytes");
    if (static_scratch && a.n > kNoAllocMaxBlocks * kTile)
        return fail(IMBA_ERR_CAPACITY, "count %d exceeds the %d entries of the allocation-free compaction", a.n,
                    kNoAllocMaxBlocks * kTile);
    if (a.n == 0) {
        if (new_count) *new_count = 0;
        return IMBA_OK;
    }
    if (!a.rays_in || !a.states_in || !a.rays_out || !a.states_out || (MODE == kModeHits && (!a.hits_in || !a.hits_out)))
        return fail(IMBA_ERR_INVALID_ARG, "NULL buffer");
    int dev = 0;
    IMBA_HIP_TRY(hipGetDevice(&dev));
    const int nb = ceil_div(a.n, kTile);
    if (static_scratch) {
        int r = scratch_for_device(dev, &a.block_offsets);
        if (r) return r;
    } else {
        Scratch& sc = t_offsets[dev & 63];
        int r = grow_buffer(sc, sizeof(int) * ((size_t)nb + 1));
        if (r) return r;
        a.block_offsets = (int*)sc.ptr;
    }
    bool sync;
    hipStream_t s;
    if (static_scratch) {
        // no per-thread stream is created here: NULL is the device's null stream
        s = (hipStream_t)stream;
        sync = stream == nullptr;
    } else {
        s = resolve_stream(stream, &sync);
    }
    hipLaunchKernelGGL(k_count<MODE>, dim3(nb), dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL(k_scan_offsets, dim3(1), dim3(1024), 0, s, a.block_offsets, nb, (int*)nullptr);
    hipLaunchKernelGGL(k_scatter<MODE>, dim3(nb), dim3(kBlock), 0, s, a, nb);
    IMBA_HIP_TRY(hipGetLastError());
    if (new_count) {
        int total = 0;
        IMBA_HIP_TRY(hipMemcpyAsync(&total, a.block_offsets + nb, sizeof(int), hipMemcpyDeviceToHost, s));
        IMBA_HIP_TRY(hipStreamSynchronize(s));
        // compact_rays keeps the size when nothing was removed (ray_queue.h:203-206): the total
        // is then the input count either way.
        *new_count = total;
    } else if (sync) {
        IMBA_HIP_TRY(hipStreamSynchronize(s));
    }
    return IMBA_OK;
}

}  // namespace
}  // namespace imba

using namespace imba;

extern "C" {

size_t imba_compact_workspace_size(int capacity) {
    return sizeof(int) * ((size_t)std::max(1, ceil_div(capacity > 0 ? capacity : 0, kTile)) + 1);
}

int imba_compact_rays_dev(const ImbaRay* rays_in, const void* states_in, int state_stride, int capacity,
                          const int* count_dev, ImbaRay* rays_out, void* states_out, int* new_count_dev,
                          void* workspace, void* stream) {
    IMBA_CHECK_ARG(state_stride > 0 && state_stride % 4 == 0, "state_stride must be a positive multiple of 4");
    IMBA_CHECK_ARG(capacity >= 0, "capacity < 0");
    IMBA_CHECK_ARG(count_dev && new_count_dev && workspace, "count_dev, new_count_dev and workspace are required");
    IMBA_CHECK_ARG(capacity == 0 || (rays_in && states_in && rays_out && states_out), "NULL buffer");
    CompactArgs a{};
    a.rays_in = (const float4*)rays_in;
    a.states_in = (const int*)states_in;
    a.stride_dw = state_stride / 4;
    a.n = capacity;
    a.count_dev = count_dev;
    a.rays_out = (float4*)rays_out;
    a.states_out = (int*)states_out;
    a.block_offsets = (int*)workspace;
    a.count_out = new_count_dev;
    const int nb = std::max(1, ceil_div(capacity, kTile));
    bool sync;
    hipStream_t s = resolve_stream(stream, &sync);
    hipLaunchKernelGGL(k_count<kModeRays>, dim3(nb), dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL(k_scan_offsets, dim3(1), dim3(1024), 0, s, a.block_offsets, nb, new_count_dev);
    hipLaunchKernelGGL(k_scatter<kModeRays>, dim3(nb), dim3(kBlock), 0, s, a, nb);
    IMBA_HIP_TRY(hipGetLastError());
    if (sync) IMBA_HIP_TRY(hipStreamSynchronize(s));
    return IMBA_OK;
}

int imba_compact_hits(const ImbaRay* rays_in, const ImbaHit* hits_in, const void* states_in, int state_stride, int count,
                      ImbaRay* rays_out, ImbaHit* hits_out, void* states_out, int* new_count, void* stream) {
    IMBA_CHECK_ARG(state_stride > 0 && state_stride % 4 == 0, "state_stride must be a positive multiple of 4");
    CompactArgs a{};
    a.rays_in = (const float4*)rays_in;
    a.hits_in = (const int4*)hits_in;
    a.states_in = (const int*)states_in;
    a.stride_dw = state_stride / 4;
    a.n = count;
    a.rays_out = (float4*)rays_out;
    a.hits_out = (int4*)hits_out;
    a.states_out = (int*)states_out;
    return run_compact<kModeHits>(a, new_count, stream);
}

int imba_compact_hits_noalloc(const ImbaRay* rays_in, const ImbaHit* hits_in, const void* states_in, int state_stride,
                              int count, ImbaRay* rays_out, ImbaHit* hits_out, void* states_out, int* new_count,
                              void* stream) {
    IMBA_CHECK_ARG(state_stride > 0 && state_stride % 4 == 0, "state_stride must be a positive multiple of 4");
    CompactArgs a{};
    a.rays_in = (const float4*)rays_in;
    a.hits_in = (const int4*)hits_in;
    a.states_in = (const int*)states_in;
    a.stride_dw = state_stride / 4;
    a.n = count;
    a.rays_out = (float4*)rays_out;
    a.hits_out = (int4*)hits_out;
    a.states_out = (int*)states_out;
    return run_compact<kModeHits>(a, new_count, stream, true);
}


int imba_compact_rays(const ImbaRay* rays_in, const void* states_in, int state_stride, int count, ImbaRay* rays_out,
                      void* states_out, int* new_count, void* stream) {
    IMBA_CHECK_ARG(state_stride > 0 && state_stride % 4 == 0, "state_stride must be a positive multiple of 4");
    CompactArgs a{};
    a.rays_in = (const float4*)rays_in;
    a.states_in = (const int*)states_in;
    a.stride_dw = state_stride / 4;
    a.n = count;
    a.rays_out = (float4*)rays_out;
    a.states_out = (int*)states_out;
    return run_compact<kModeRays>(a, new_count, stream);
}

}  // extern "C"
