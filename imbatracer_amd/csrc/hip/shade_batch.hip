// Batched path-tracing shading step and stable queue compaction over many ray queues.
//
// The reference shades and compacts every queue on the host after its traversal
// (render/scheduling/queue_scheduler.h:177-275, ray_queue.h:188-207).  Here one step of many
// queues -- any scenes, any sizes, sizes held in device memory -- takes five launches:
//   k_shade_prefix    one block: each queue's entry count (its device count, clamped to the
//                     capacity), and exclusive prefixes of its shading chunks (256 entries) and
//                     compaction tiles (2048 entries);
//   k_bounce_batch    one block per chunk: the chunk's queue by binary search over the chunk
//                     prefix, then pt_shade_entry (the single-queue shading step: emitter
//                     contribution, next-event shadow ray into the queue's shadow slots,
//                     Lambertian bounce) on each entry;
//   k_ccount_batch    one block per tile: survivors (pixel_id >= 0) in the tile;
//   k_cscan_batch     one block per queue: exclusive scan of its tile counts, the total to the
//                     queue's next device count;
//   k_cscatter_batch  one block per tile: each survivor's rank (ballot, mbcnt, wave counts in
//                     LDS) and its ray and state moved to the queue's output.
// Survivors keep their order, so every queue's result equals its own imba_pt_shade +
// imba_compact_rays_dev.  After the shadow queues' (batched, any-hit) traversal,
// k_shadow_batch adds the unoccluded shadow contributions of every queue (imba_pt_shadow).
// Nothing here allocates or sets memory: a bounce loop over the batch can be captured in a
// graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "common.h"
#include "pt_shade.h"

struct ImbaShadeBatch {
    int device = 0;
    int n = 0;                          // items
    int cap = 0;                        // items the device table holds room for
    ImbaShadeItem* d_items = nullptr;
    int* d_info = nullptr;              // entries, chunk prefix, tile prefix: 3 x (cap + 1)
    int* d_tile_counts = nullptr;
    long long tiles_cap = 0;
    long long max_chunks = 0;           // bounce chunks of the capacities
    long long max_tiles = 0;            // compaction tiles of the capacities
    ImbaShadeItem* staging[2] = {nullptr, nullptr};  // pinned host copies of the table
    hipEvent_t staged[2] = {nullptr, nullptr};
    int next_staging = 0;
    bool launched = false;  // k_shade_prefix has run at least once (its info feeds k_shadow_batch)
    bool captured = false;  // launched into a graph: the device tables may no longer move
    long long captured_chunks = 0;  // grid sizes recorded into graphs (the largest)
    long long captured_tiles = 0;
};

namespace imba {
namespace {

constexpr int kChunk = 256;                 // entries per bounce block
constexpr int kCBlock = 256;
constexpr int kCItems = 8;
constexpr int kCTile = kCBlock * kCItems;   // entries per compaction tile
constexpr int kCWaves = kCBlock / 64;
constexpr int kStateDw = sizeof(PtState) / 4;

__device__ __forceinline__ int item_entries(const ImbaShadeItem& it) {
    return it.count_dev ? min(max(*it.count_dev, 0), it.capacity) : it.capacity;
}

// The item owning chunk (or tile) c: the last one whose first chunk is <= c (items without
// entries share their first chunk with the next item and are skipped).
__device__ __forceinline__ int find_item(const int* begin, int n_items, int c) {
    int lo = 0, hi = n_items - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (begin[mid] <= c) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(1024) void k_shade_prefix(const ImbaShadeItem* items, int n_items, int* info) {
    int* ent = info;
    int* cb = info + (n_items + 1);
    int* tb = info + 2 * (n_items + 1);
    __shared__ int pc[1024], pt[1024];
    const int t = threadIdx.x;
    const int per = (n_items + 1023) / 1024;
    const int b0 = min(n_items, t * per), b1 = min(n_items, b0 + per);
    int sc = 0, st = 0;
    for (int q = b0; q < b1; q++) {
        const int e = item_entries(items[q]);
        ent[q] = e;
        sc += (e + kChunk - 1) / kChunk;
        st += (e + kCTile - 1) / kCTile;
    }
    pc[t] = sc;
    pt[t] = st;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        const int x = t >= d ? pc[t - d] : 0;
        const int y = t >= d ? pt[t - d] : 0;
        __syncthreads();
        pc[t] += x;
        pt[t] += y;
        __syncthreads();
    }
    int rc = pc[t] - sc, rt = pt[t] - st;
    for (int q = b0; q < b1; q++) {
        cb[q] = rc;
        tb[q] = rt;
        rc += (ent[q] + kChunk - 1) / kChunk;
        rt += (ent[q] + kCTile - 1) / kCTile;
    }
    if (t == 1023) {
        cb[n_items] = pc[1023];
        tb[n_items] = pt[1023];
        ent[n_items] = 0;
    }
}

__global__ __launch_bounds__(kChunk) void k_bounce_batch(const ImbaShadeItem* items, int n_items, const int* info) {
    const int* ent = info;
    const int* cb = info + (n_items + 1);
    const int c = blockIdx.x;
    if (c >= cb[n_items]) return;
    const int q = find_item(cb, n_items, c);
    const int i = (c - cb[q]) * kChunk + (int)threadIdx.x;
    if (i >= ent[q]) return;
    const ImbaShadeItem& it = items[q];
    PtShadeArgs a;
    a.rays = (const float4*)it.rays;
    a.hits = (const int4*)it.hits;
    a.states = (PtState*)it.states;
    a.rays_out = (float4*)it.rays;
    a.shadow_rays = (float4*)it.shadow_rays;
    a.shadow_states = (ShadowState*)it.shadow_states;
    a.image = it.image;
    a.sh = it.shading;
    pt_shade_entry(a, i);
}

// The shadow contributions of every item with shadow slots: the same chunks as k_bounce_batch
// (the entry counts k_shade_prefix read for the preceding shading launch).
__global__ __launch_bounds__(kChunk) void k_shadow_batch(const ImbaShadeItem* items, int n_items, const int* info) {
    const int* ent = info;
    const int* cb = info + (n_items + 1);
    const int c = blockIdx.x;
    if (c >= cb[n_items]) return;
    const int q = find_item(cb, n_items, c);
    const int i = (c - cb[q]) * kChunk + (int)threadIdx.x;
    if (i >= ent[q]) return;
    const ImbaShadeItem& it = items[q];
    if (!it.shadow_states || !it.shadow_hits || !it.image) return;
    pt_shadow_entry((const int4*)it.shadow_hits, (const ShadowState*)it.shadow_states, it.image, i);
}

__device__ __forceinline__ bool survives(const ImbaShadeItem& it, int i) {
    return reinterpret_cast<const int*>(it.states)[(size_t)i * kStateDw] >= 0;
}

__global__ __launch_bounds__(kCBlock) void k_ccount_batch(const ImbaShadeItem* items, int n_items, const int* info,
                                                          int* tile_counts) {
    __shared__ int wave_sum[kCWaves];
    const int* ent = info;
    const int* tb = info + 2 * (n_items + 1);
    const int t = blockIdx.x;
    if (t >= tb[n_items]) return;
    const int q = find_item(tb, n_items, t);
    const ImbaShadeItem& it = items[q];
    const int n = ent[q];
    const int base = (t - tb[q]) * kCTile;
    int cnt = 0;
#pragma unroll
    for (int k = 0; k < kCItems; k++) {
        const int i = base + k * kCBlock + (int)threadIdx.x;
        cnt += __popcll(__ballot(i < n && survives(it, i)));
    }
    if ((threadIdx.x & 63) == 0) wave_sum[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        int s = 0;
        for (int w = 0; w < kCWaves; w++) s += wave_sum[w];
        tile_counts[t] = s;
    }
}

// One block per item: exclusive scan of the item's tile counts in place, total to its next count.
__global__ __launch_bounds__(kCBlock) void k_cscan_batch(const ImbaShadeItem* items, int n_items, const int* info,
                                                         int* tile_counts) {
    __shared__ int part[kCBlock];
    const int* tb = info + 2 * (n_items + 1);
    const int q = blockIdx.x;
    const int t0 = tb[q], t1 = tb[q + 1];
    const int nt = t1 - t0;
    const int t = threadIdx.x;
    const int per = (nt + kCBlock - 1) / kCBlock;
    const int b0 = min(nt, t * per), b1 = min(nt, b0 + per);
    int s = 0;
    for (int b = b0; b < b1; b++) s += tile_counts[t0 + b];
    part[t] = s;
    __syncthreads();
    for (int d = 1; d < kCBlock; d <<= 1) {
        const int x = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += x;
        __syncthreads();
    }
    int run = part[t] - s;
    for (int b = b0; b < b1; b++) {
        const int c = tile_counts[t0 + b];
        tile_counts[t0 + b] = run;
        run += c;
    }
    if (t == kCBlock - 1 && items[q].next_count_dev) *items[q].next_count_dev = part[kCBlock - 1];
}

__global__ __launch_bounds__(kCBlock) void k_cscatter_batch(const ImbaShadeItem* items, int n_items, const int* info,
                                                            const int* tile_offsets) {
    __shared__ int wave_cnt[kCWaves];
    const int* ent = info;
    const int* tb = info + 2 * (n_items + 1);
    const int t = blockIdx.x;
    if (t >= tb[n_items]) return;
    const int q = find_item(tb, n_items, t);
    const ImbaShadeItem& it = items[q];
    const int n = ent[q];
    const int base = (t - tb[q]) * kCTile;
    int kept_before = tile_offsets[t];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const float4* rin = (const float4*)it.rays;
    float4* rout = (float4*)it.rays_out;
    const PtState* s_in = (const PtState*)it.states;
    PtState* s_out = (PtState*)it.states_out;
    for (int k = 0; k < kCItems; k++) {
        const int i = base + k * kCBlock + (int)threadIdx.x;
        const bool f = i < n && survives(it, i);
        const unsigned long long m = __ballot(f);
        const int r = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
        if (lane == 0) wave_cnt[wave] = __popcll(m);
        __syncthreads();
        int woff = 0, chunk_kept = 0;
#pragma unroll
        for (int w = 0; w < kCWaves; w++) {
            const int c = wave_cnt[w];
            woff += w < wave ? c : 0;
            chunk_kept += c;
        }
        if (f) {
            const size_t d = (size_t)(kept_before + woff + r);
            rout[2 * d] = rin[2 * (size_t)i];
            rout[2 * d + 1] = rin[2 * (size_t)i + 1];
            s_out[d] = s_in[i];
        }
        kept_before += chunk_kept;
        __syncthreads();
    }
}

int set_items(ImbaShadeBatch* b, const ImbaShadeItem* items, int n, hipStream_t stream) {
    long long chunks = 0, tiles = 0;
    for (int i = 0; i < n; i++) {
        const ImbaShadeItem& it = items[i];
        if (it.capacity < 0) return fail(IMBA_ERR_INVALID_ARG, "item %d: capacity < 0", i);
        if (it.capacity > 0 && (!it.rays || !it.hits || !it.states || !it.rays_out || !it.states_out))
            return fail(IMBA_ERR_INVALID_ARG, "item %d: NULL buffer", i);
        if (it.capacity > 0) {
            const int r = check_shading(&it.shading);
            if (r) return r;
        }
        if ((it.shadow_rays == nullptr) != (it.shadow_states == nullptr))
            return fail(IMBA_ERR_INVALID_ARG, "item %d: shadow_rays and shadow_states go together", i);
        if (it.capacity > 0 && (it.rays_out == it.rays || it.states_out == it.states))
            return fail(IMBA_ERR_INVALID_ARG, "item %d: the compaction output must be other buffers", i);
        chunks += ceil_div(it.capacity, kChunk);
        tiles += ceil_div(it.capacity, kCTile);
    }
    if (chunks > INT32_MAX) return fail(IMBA_ERR_INVALID_ARG, "too many entries for one batch");
    // A graph replays its launches with the table addresses they were recorded with, so the
    // tables of a captured batch cannot be reallocated.
    if (b->captured && (n > b->cap || tiles > b->tiles_cap || chunks > b->captured_chunks ||
                        tiles > b->captured_tiles))
        return fail(IMBA_ERR_CAPACITY, "shade batch was launched into a graph with grids for %lld chunks / %lld tiles "
                    "and tables for %d items; %lld / %lld / %d do not fit", b->captured_chunks, b->captured_tiles,
                    b->cap, chunks, tiles, n);
    if (n > b->cap) {
        IMBA_HIP_TRY(hipDeviceSynchronize());
        if (b->d_items) {
            int r = retire_buffer(b->d_items);
            if (r) return r;
        }
        if (b->d_info) {
            int r = retire_buffer(b->d_info);
            if (r) return r;
        }
        b->d_items = nullptr;
        b->d_info = nullptr;
        for (int k = 0; k < 2; k++) {
            if (b->staging[k]) IMBA_HIP_TRY(hipHostFree(b->staging[k]));
            b->staging[k] = nullptr;
        }
        const int cap = std::max(n, 2 * b->cap);
        IMBA_HIP_TRY(hipMalloc((void**)&b->d_items, sizeof(ImbaShadeItem) * (size_t)cap));
        IMBA_HIP_TRY(hipMalloc((void**)&b->d_info, sizeof(int) * 3 * ((size_t)cap + 1)));
        for (int k = 0; k < 2; k++) {
            IMBA_HIP_TRY(hipHostMalloc((void**)&b->staging[k], sizeof(ImbaShadeItem) * (size_t)cap));
            if (!b->staged[k]) IMBA_HIP_TRY(hipEventCreateWithFlags(&b->staged[k], hipEventDisableTiming));
        }
        b->cap = cap;
    }
    if (tiles > b->tiles_cap) {
        IMBA_HIP_TRY(hipDeviceSynchronize());
        if (b->d_tile_counts) {
            int r = retire_buffer(b->d_tile_counts);
            if (r) return r;
        }
        b->d_tile_counts = nullptr;
        IMBA_HIP_TRY(hipMalloc((void**)&b->d_tile_counts, sizeof(int) * (size_t)std::max(tiles, 1LL)));
        b->tiles_cap = tiles;
    }
    const size_t bytes = sizeof(ImbaShadeItem) * (size_t)n;
    if (n > 0) {
        if (stream) {
            const int k = b->next_staging;
            b->next_staging ^= 1;
            IMBA_HIP_TRY(hipEventSynchronize(b->staged[k]));
            std::memcpy(b->staging[k], items, bytes);
            IMBA_HIP_TRY(hipMemcpyAsync(b->d_items, b->staging[k], bytes, hipMemcpyHostToDevice, stream));
            IMBA_HIP_TRY(hipEventRecord(b->staged[k], stream));
        } else {
            IMBA_HIP_TRY(hipMemcpy(b->d_items, items, bytes, hipMemcpyHostToDevice));
        }
    }
    b->n = n;
    b->max_chunks = chunks;
    b->max_tiles = tiles;
    return IMBA_OK;
}

void destroy(ImbaShadeBatch* b) {
    if (!b) return;
    hipFree(b->d_items);
    hipFree(b->d_info);
    hipFree(b->d_tile_counts);
    for (int k = 0; k < 2; k++) {
        if (b->staging[k]) hipHostFree(b->staging[k]);
        if (b->staged[k]) hipEventDestroy(b->staged[k]);
    }
    delete b;
}

}  // namespace
}  // namespace imba

using namespace imba;

extern "C" {

int imba_shade_batch_create(const ImbaShadeItem* items, int count, ImbaShadeBatch** out) {
    IMBA_CHECK_ARG(out && count >= 0 && (count == 0 || items), "bad arguments");
    *out = nullptr;
    ImbaShadeBatch* b = new ImbaShadeBatch();
    if (hipGetDevice(&b->device) != hipSuccess) {
        delete b;
        return fail(IMBA_ERR_NO_DEVICE, "no HIP device");
    }
    const int r = set_items(b, items, count, nullptr);
    if (r) {
        destroy(b);
        return r;
    }
    *out = b;
    return IMBA_OK;
}

int imba_shade_batch_update(ImbaShadeBatch* b, const ImbaShadeItem* items, int count, void* stream) {
    IMBA_CHECK_ARG(b && count >= 0 && (count == 0 || items), "bad arguments");
    return set_items(b, items, count, (hipStream_t)stream);
}

int imba_shade_batch_launch(ImbaShadeBatch* b, void* stream) {
    IMBA_CHECK_ARG(b, "batch is NULL");
    if (b->n == 0) return IMBA_OK;
    bool sync;
    hipStream_t s = resolve_stream(stream, &sync);
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    IMBA_HIP_TRY(hipStreamIsCapturing(s, &cs));
    if (cs != hipStreamCaptureStatusNone) {
        b->captured = true;
        b->captured_chunks = std::max(b->captured_chunks, b->max_chunks);
        b->captured_tiles = std::max(b->captured_tiles, b->max_tiles);
    }
    const int n = b->n;
    b->launched = true;
    hipLaunchKernelGGL(k_shade_prefix, dim3(1), dim3(1024), 0, s, (const ImbaShadeItem*)b->d_items, n, b->d_info);
    if (b->max_chunks > 0)
        hipLaunchKernelGGL(k_bounce_batch, dim3((unsigned)b->max_chunks), dim3(kChunk), 0, s,
                           (const ImbaShadeItem*)b->d_items, n, (const int*)b->d_info);
    if (b->max_tiles > 0)
        hipLaunchKernelGGL(k_ccount_batch, dim3((unsigned)b->max_tiles), dim3(kCBlock), 0, s,
                           (const ImbaShadeItem*)b->d_items, n, (const int*)b->d_info, b->d_tile_counts);
    hipLaunchKernelGGL(k_cscan_batch, dim3(n), dim3(kCBlock), 0, s, (const ImbaShadeItem*)b->d_items, n,
                       (const int*)b->d_info, b->d_tile_counts);
    if (b->max_tiles > 0)
        hipLaunchKernelGGL(k_cscatter_batch, dim3((unsigned)b->max_tiles), dim3(kCBlock), 0, s,
                           (const ImbaShadeItem*)b->d_items, n, (const int*)b->d_info, (const int*)b->d_tile_counts);
    IMBA_HIP_TRY(hipGetLastError());
    if (sync) IMBA_HIP_TRY(hipStreamSynchronize(s));
    return IMBA_OK;
}

int imba_shade_batch_shadow(ImbaShadeBatch* b, void* stream) {
    IMBA_CHECK_ARG(b, "batch is NULL");
    if (b->n == 0 || b->max_chunks == 0) return IMBA_OK;
    if (!b->launched) return fail(IMBA_ERR_INVALID_ARG, "imba_shade_batch_shadow needs a preceding imba_shade_batch_launch");
    bool sync;
    hipStream_t s = resolve_stream(stream, &sync);
    hipLaunchKernelGGL(k_shadow_batch, dim3((unsigned)b->max_chunks), dim3(kChunk), 0, s,
                       (const ImbaShadeItem*)b->d_items, b->n, (const int*)b->d_info);
    IMBA_HIP_TRY(hipGetLastError());
    if (sync) IMBA_HIP_TRY(hipStreamSynchronize(s));
    return IMBA_OK;
}

int imba_shade_batch_release(ImbaShadeBatch* b) {
    destroy(b);
    return IMBA_OK;
}

}  // extern "C"
