// Photon hash grid on the GPU: counting-sort build and k-capped range query.
//
// HashGrid (rangesearch/rangesearch.h:21-158) builds on the host with TBB: bounds reduce,
// atomic count per bucket, serial exclusive scan, atomic scatter.  Here:
//   bounds   k_bounds (per-block min/max) + k_params (combine, grow by 0.1%, derived values)
//   keys     k_grid_keys: bucket of each photon, paired with the photon index
//   sort     stable LSD radix sort of the pairs by bucket (radix.h): the slot order is by
//            bucket, then by photon index, which is deterministic and equals a serial run of
//            the reference's scatter
//   ends     k_mark_ends writes slot + 1 at the last slot of each bucket, an inclusive max-scan
//            (k_max_tiles / k_max_top / k_max_apply) carries it over the empty buckets
//   gather   positions copied to slot order as float4 (x, y, z, index bits)
// The query visits the same 8 buckets in the same order as the reference (duplicates
// included), one query per wave, and keeps the k-capped, distance-sorted list in lanes 0..k-1
// with the reference's insertion rule.  Squared distances are computed without fused
// multiply-add, as on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>

#include "common.h"
#include "grid_query.h"
#include "radix.h"

namespace imba {
namespace {

constexpr int kBoundsBlocks = 1024;
constexpr int kBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;
constexpr int kSortItems = 16;  // radix tiles of 4096 pairs

__device__ __forceinline__ const float* photon_pos(const unsigned char* base, int stride, int off, int i) {
    return reinterpret_cast<const float*>(base + (size_t)i * stride + off);
}

__global__ __launch_bounds__(kBlock) void k_bounds(const unsigned char* ph, int stride, int off, int n, float* partial) {
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const float* p = photon_pos(ph, stride, off, i);
        for (int k = 0; k < 3; k++) {
            const float v = p[k];
            lo[k] = lo[k] < v ? lo[k] : v;
            hi[k] = hi[k] > v ? hi[k] : v;
        }
    }
    __shared__ float red[6][kBlock];
    for (int k = 0; k < 3; k++) { red[k][threadIdx.x] = lo[k]; red[3 + k][threadIdx.x] = hi[k]; }
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            for (int k = 0; k < 3; k++) {
                const float a = red[k][threadIdx.x], b = red[k][threadIdx.x + s];
                red[k][threadIdx.x] = a < b ? a : b;
                const float c = red[3 + k][threadIdx.x], d = red[3 + k][threadIdx.x + s];
                red[3 + k][threadIdx.x] = c > d ? c : d;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) partial[6 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// One block: the per-block bounds reduced to the photon bounds (min / max are exact in any
// order), then the grid parameters.
__global__ __launch_bounds__(kBlock) void k_params(const float* partial, int nparts, GridParams* gp, float radius,
                                                   int table_size, int count) {
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int b = threadIdx.x; b < nparts; b += kBlock)
        for (int k = 0; k < 3; k++) {
            const float a = partial[6 * b + k], c = partial[6 * b + 3 + k];
            lo[k] = lo[k] < a ? lo[k] : a;
            hi[k] = hi[k] > c ? hi[k] : c;
        }
    __shared__ float red[6][kBlock];
    for (int k = 0; k < 3; k++) { red[k][threadIdx.x] = lo[k]; red[3 + k][threadIdx.x] = hi[k]; }
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            for (int k = 0; k < 3; k++) {
                const float a = red[k][threadIdx.x], b = red[k][threadIdx.x + s];
                red[k][threadIdx.x] = a < b ? a : b;
                const float c = red[3 + k][threadIdx.x], d = red[3 + k][threadIdx.x + s];
                red[3 + k][threadIdx.x] = c > d ? c : d;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    for (int k = 0; k < 3; k++) { lo[k] = red[k][0]; hi[k] = red[3 + k][0]; }
    for (int k = 0; k < 3; k++) {
        const float e = hi[k] - lo[k];
        gp->bmax[k] = hi[k] + e * 0.001f;
        gp->bmin[k] = lo[k] - e * 0.001f;
    }
    gp->radius = radius;
    gp->radius_sqr = radius * radius;
    gp->cell_size = radius * 2.0f;
    gp->inv_cell_size = 1.0f / gp->cell_size;
    gp->table_size = table_size;
    gp->count = count;
}

// Zero fill (a kernel rather than a memset node, so the build can be captured in a graph).
__global__ __launch_bounds__(kBlock) void k_zero(int* p, int n) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) p[i] = 0;
}


__global__ __launch_bounds__(kBlock) void k_grid_keys(const unsigned char* ph, int stride, int off, int n,
                                                      const GridParams* gp, int* keys, int* vals) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float* p = photon_pos(ph, stride, off, i);
    const float inv = gp->inv_cell_size;
    const int cx = (int)floorf(inv * (p[0] - gp->bmin[0]));
    const int cy = (int)floorf(inv * (p[1] - gp->bmin[1]));
    const int cz = (int)floorf(inv * (p[2] - gp->bmin[2]));
    keys[i] = bucket_of((uint32_t)cx, (uint32_t)cy, (uint32_t)cz, gp->table_size);
    vals[i] = i;
}

// ends[b] = 1 + the last slot of bucket b, for the buckets that hold photons (the others
// were zeroed).
__global__ __launch_bounds__(kBlock) void k_mark_ends(const int* keys, int n, int* ends) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    const int b = keys[s];
    if (s == n - 1 || keys[s + 1] != b) ends[b] = s + 1;
}

// Inclusive max-scan of `n` ints in place: per-tile maxima, a one-block exclusive max-scan of
// the maxima, then each tile is scanned again from its offset.  Values are >= 0.
__global__ __launch_bounds__(kBlock) void k_max_tiles(const int* in, int n, int* tile_max) {
    __shared__ int ws[kBlock / 64];
    const int base = blockIdx.x * kScanTile;
    int m = 0;
    for (int it = 0; it < kScanItems; it++) {
        const int i = base + it * kBlock + threadIdx.x;
        if (i < n) m = max(m, in[i]);
    }
    for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_down(m, o, 64));
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < kBlock / 64; w++) t = max(t, ws[w]);
        tile_max[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(1024) void k_max_top(int* v, int nt) {
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int per = (nt + 1023) / 1024;
    const int b0 = min(nt, t * per), b1 = min(nt, b0 + per);
    int m = 0;
    for (int b = b0; b < b1; b++) m = max(m, v[b]);
    part[t] = m;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        const int x = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] = max(part[t], x);
        __syncthreads();
    }
    int run = t > 0 ? part[t - 1] : 0;
    for (int b = b0; b < b1; b++) {
        const int c = v[b];
        v[b] = run;
        run = max(run, c);
    }
}

__global__ __launch_bounds__(kBlock) void k_max_apply(int* v, int n, const int* tile_offsets) {
    __shared__ int wmax[kBlock / 64];
    const int base = blockIdx.x * kScanTile;
    int run = tile_offsets[blockIdx.x];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int it = 0; it < kScanItems; it++) {
        const int i = base + it * kBlock + threadIdx.x;
        int x = i < n ? v[i] : 0;
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(x, o, 64);
            if (lane >= o) x = max(x, y);
        }
        if (lane == 63) wmax[wave] = x;
        __syncthreads();
        int wm = 0, total = 0;
        for (int w = 0; w < kBlock / 64; w++) {
            if (w < wave) wm = max(wm, wmax[w]);
            total = max(total, wmax[w]);
        }
        if (i < n) v[i] = max(run, max(wm, x));
        run = max(run, total);
        __syncthreads();
    }
}

__global__ __launch_bounds__(kBlock) void k_gather_pos(const unsigned char* ph, int stride, int off, int n, const int* order,
                                                       float4* pos4) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    const int i = order[s];
    const float* p = photon_pos(ph, stride, off, i);
    pos4[s] = make_float4(p[0], p[1], p[2], __int_as_float(i));
}

// Bounds of every non-empty bucket's photons (the queries skip buckets that lie beyond their
// limit; an empty bucket's entry is never read), in the order-preserving integer form of
// grid_query.h so that atomic min / max can combine them.  First the first slot of every bucket
// run in the sorted keys resets its bucket's entry; then every wave takes 64 consecutive slots,
// reduces the runs of equal keys within the wave (shift-down doubling: keys are sorted, so equal
// keys are contiguous) and the first lane of each run combines the run's bounds into the bucket's
// entry with six atomics.  Hot buckets thus cost one atomic per 64 photons.
__global__ __launch_bounds__(kBlock) void k_bucket_box_reset(const int* keys_sorted, int n, uint32_t* box) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    const int b = keys_sorted[s];
    if (s > 0 && keys_sorted[s - 1] == b) return;
    uint32_t* e = box + 6 * (size_t)b;
    e[0] = e[1] = e[2] = 0xFFFFFFFFu;  // min over nothing
    e[3] = e[4] = e[5] = 0u;           // max over nothing
}

__global__ __launch_bounds__(kBlock) void k_bucket_box(const int* keys_sorted, int n, const float4* pos4, uint32_t* box) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool valid = s < n;
    const int key = valid ? keys_sorted[s] : -1;
    uint32_t lo[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, hi[3] = {0u, 0u, 0u};
    if (valid) {
        const float4 p = pos4[s];
        lo[0] = hi[0] = ordered_bits(p.x);
        lo[1] = hi[1] = ordered_bits(p.y);
        lo[2] = hi[2] = ordered_bits(p.z);
    }
    for (int off = 1; off < 64; off <<= 1) {
        const int k2 = __shfl_down(key, off, 64);
        const bool take = lane + off < 64 && k2 == key;
        for (int a = 0; a < 3; a++) {
            const uint32_t l2 = __shfl_down(lo[a], off, 64), h2 = __shfl_down(hi[a], off, 64);
            if (take) {
                lo[a] = min(lo[a], l2);
                hi[a] = max(hi[a], h2);
            }
        }
    }
    const int kprev = __shfl_up(key, 1, 64);
    if (!valid || (lane > 0 && kprev == key)) return;  // not the first lane of its run
    uint32_t* e = box + 6 * (size_t)key;
    for (int a = 0; a < 3; a++) {
        atomicMin(e + a, lo[a]);
        atomicMax(e + 3 + a, hi[a]);
    }
}

// One query per wave (the scan and insertion of query_wave_list); the list goes to out_idx /
// out_dist.
__global__ __launch_bounds__(kBlock) void k_query_wave(const GridParams* gp, const int* ends, const float4* pos4,
                                                       const uint32_t* box, const unsigned char* qs, int qstride,
                                                       int qoff, int nq, int k, int* out_idx, float* out_dist,
                                                       int* out_count) {
    const int q = blockIdx.x * (kBlock / 64) + (int)(threadIdx.x >> 6);
    if (q >= nq) return;  // whole waves leave together
    const int lane = (int)(threadIdx.x & 63);
    const float* qp = photon_pos(qs, qstride, qoff, q);
    float my_dist;
    int my_idx;
    const int count = query_wave_list<false>(*gp, ends, pos4, box, qp[0], qp[1], qp[2], k, my_dist, my_idx);
    if (lane < count) {
        out_idx[(size_t)q * k + lane] = my_idx;
        if (out_dist) out_dist[(size_t)q * k + lane] = my_dist;
    }
    if (lane == 0) out_count[q] = count;
}

// k > 64: the same scan and insertion rule with the k-capped list in LDS, one query per wave
// and one wave per block (dynamic LDS: k distances, then k photon indices).  Inserting at
// position pos moves the entries [pos, count - 1) one slot up, 64 at a time from the top: the
// wave reads 64 entries, then writes them one slot higher, so every entry is read before it is
// overwritten.  The insertion position is the number of entries <= the candidate, counted over
// the list by all lanes.
constexpr int kQueryMaxK = 4096;

__global__ __launch_bounds__(64) void k_query_lds(const GridParams* gp, const int* ends, const float4* pos4,
                                                  const unsigned char* qs, int qstride, int qoff, int nq, int k,
                                                  int* out_idx, float* out_dist, int* out_count) {
    extern __shared__ float list_lds[];
    float* ld = list_lds;
    int* li = reinterpret_cast<int*>(list_lds + k);
    const int q = blockIdx.x;
    if (q >= nq) return;
    const int lane = (int)threadIdx.x;
    const float* qp = photon_pos(qs, qstride, qoff, q);
    const float qx = qp[0], qy = qp[1], qz = qp[2];
    const GridParams g = *gp;
    int count = 0;
    const bool inside = qx >= g.bmin[0] && qy >= g.bmin[1] && qz >= g.bmin[2] &&
                        qx <= g.bmax[0] && qy <= g.bmax[1] && qz <= g.bmax[2];
    if (inside && g.count > 0) {
        const float cx = g.inv_cell_size * (qx - g.bmin[0]);
        const float cy = g.inv_cell_size * (qy - g.bmin[1]);
        const float cz = g.inv_cell_size * (qz - g.bmin[2]);
        const float fx = floorf(cx), fy = floorf(cy), fz = floorf(cz);
        const int px = (int)fx, py = (int)fy, pz = (int)fz;
        const int pxo = px + ((cx - fx) < 0.5f ? -1 : 1);
        const int pyo = py + ((cy - fy) < 0.5f ? -1 : 1);
        const int pzo = pz + ((cz - fz) < 0.5f ? -1 : 1);
        int begin = 0, size = 0;
        if (lane < 8) {
            const int x = (lane & 4) ? pxo : px;
            const int y = (lane & 2) ? pyo : py;
            const int z = (lane & 1) ? pzo : pz;
            const int b = bucket_of((uint32_t)x, (uint32_t)y, (uint32_t)z, g.table_size);
            begin = b == 0 ? 0 : ends[b - 1];
            size = ends[b] - begin;
        }
        int pre = size;
        for (int o = 1; o < 8; o <<= 1) {
            const int y = __shfl_up(pre, o, 64);
            if (lane >= o) pre += y;
        }
        pre -= size;
        int beg_j[8], pre_j[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            beg_j[j] = __builtin_amdgcn_readlane(begin, j);
            pre_j[j] = __builtin_amdgcn_readlane(pre, j);
        }
        const int total = __builtin_amdgcn_readlane(pre + size, 7);
        float kth = 0.0f;
        for (int base = 0; base < total; base += 64) {
            const int c = base + lane;
            const bool valid = c < total;
            float d2 = 0.0f;
            int pid = -1;
            if (valid) {
                int slot = 0;
#pragma unroll
                for (int j = 0; j < 8; j++)
                    if (c >= pre_j[j]) slot = beg_j[j] + (c - pre_j[j]);
                const float4 p = pos4[slot];
                const float dx = qx - p.x, dy = qy - p.y, dz = qz - p.z;
                d2 = dx * dx + dy * dy + dz * dz;
                pid = __float_as_int(p.w);
            }
            const float lim = count == k ? kth : g.radius_sqr;
            unsigned long long m = __ballot(valid && d2 <= lim && d2 <= g.radius_sqr);
            while (m) {
                const int src = __builtin_ctzll(m);
                m &= m - 1;
                const float dd = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(d2), src));
                const int id = __builtin_amdgcn_readlane(pid, src);
                if (count == k) {
                    if (kth < dd) continue;
                } else {
                    count++;
                }
                const int keep = count - 1;  // entries that stay (the last one drops when full)
                int le = 0;
                for (int j = lane; j < keep; j += 64) le += ld[j] <= dd ? 1 : 0;
                for (int o = 32; o > 0; o >>= 1) le += __shfl_xor(le, o, 64);
                const int pos = le;
                for (int top = keep - 1; top >= pos; top -= 64) {
                    const int j = top - lane;
                    const bool act = j >= pos;
                    float v = 0.0f;
                    int w = 0;
                    if (act) {
                        v = ld[j];
                        w = li[j];
                    }
                    __syncthreads();
                    if (act) {
                        ld[j + 1] = v;
                        li[j + 1] = w;
                    }
                    __syncthreads();
                }
                if (lane == 0) {
                    ld[pos] = dd;
                    li[pos] = id;
                }
                __syncthreads();
                if (count == k) kth = ld[k - 1];
            }
        }
    }
    for (int j = lane; j < count; j += 64) {
        out_idx[(size_t)q * k + j] = li[j];
        if (out_dist) out_dist[(size_t)q * k + j] = ld[j];
    }
    if (lane == 0) out_count[q] = count;
}

template <typename T>
int ensure(T*& p, int& cap_now, int want) {
    if (p && cap_now >= want) return IMBA_OK;
    if (p) IMBA_HIP_TRY(hipFree(p));
    p = nullptr;
    IMBA_HIP_TRY(hipMalloc((void**)&p, sizeof(T) * (size_t)std::max(want, 1)));
    cap_now = want;
    return IMBA_OK;
}

}  // namespace
}  // namespace imba

using namespace imba;

extern "C" {

int imba_hashgrid_create(ImbaHashGrid** out) {
    IMBA_CHECK_ARG(out, "out is NULL");
    ImbaHashGrid* g = new ImbaHashGrid();
    if (hipGetDevice(&g->device) != hipSuccess) g->device = 0;
    if (hipMalloc((void**)&g->d_params, sizeof(GridParams)) != hipSuccess ||
        hipMalloc((void**)&g->partial, sizeof(float) * 6 * kBoundsBlocks) != hipSuccess) {
        imba_hashgrid_release(g);
        return fail(IMBA_ERR_OUT_OF_MEMORY, "hash grid parameters");
    }
    GridParams zero{};
    if (hipMemcpy(g->d_params, &zero, sizeof(zero), hipMemcpyHostToDevice) != hipSuccess) {
        imba_hashgrid_release(g);
        return fail(IMBA_ERR_HIP, "hash grid parameters");
    }
    *out = g;
    return IMBA_OK;
}

int imba_hashgrid_release(ImbaHashGrid* g) {
    if (!g) return IMBA_OK;
    void* ptrs[] = {g->cell_ends, g->order, g->keys_a, g->vals_a, g->keys_b, g->vals_b, g->pos4, g->d_params,
                    g->partial, g->scan_tmp, g->digits, g->bucket_box};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete g;
    return IMBA_OK;
}

int imba_hashgrid_build(ImbaHashGrid* g, const void* photons, int stride, int pos_offset, int count, float radius,
                        void* stream) {
    IMBA_CHECK_ARG(g, "grid is NULL");
    IMBA_CHECK_ARG(count >= 0 && count <= (1 << 29), "count out of range");
    IMBA_CHECK_ARG(count == 0 || photons, "photons is NULL");
    IMBA_CHECK_ARG(stride >= 12 && stride % 4 == 0 && pos_offset >= 0 && pos_offset % 4 == 0 && pos_offset + 12 <= stride,
                   "stride/pos_offset must be multiples of 4 with the position inside the record");
    IMBA_CHECK_ARG(radius > 0.0f, "radius must be positive");
    bool sync;
    hipStream_t s = resolve_stream(stream, &sync);
    // Bucket table: max(previous size, 2 * count) (rangesearch.h:33-34).
    const int table = std::max(g->table_size, 2 * count);
    if (table > g->table_cap) {
        int r = ensure(g->cell_ends, g->table_cap, table);
        if (r) return r;
    }
    if (table > g->box_cap) {  // 6 uint32 per bucket
        int c = 0, r = ensure(g->bucket_box, c, 6 * table);
        if (r) return r;
        g->box_cap = table;
    }
    g->table_size = table;
    if (count > g->cap) {
        int c = 0, r;
        if ((r = ensure(g->order, c, count))) return r;
        c = 0; if ((r = ensure(g->keys_a, c, count))) return r;
        c = 0; if ((r = ensure(g->vals_a, c, count))) return r;
        c = 0; if ((r = ensure(g->keys_b, c, count))) return r;
        c = 0; if ((r = ensure(g->vals_b, c, count))) return r;
        c = 0; if ((r = ensure(g->pos4, c, count))) return r;
        g->cap = count;
    }
    const int ntiles = ceil_div(std::max(table, 1), kScanTile);
    if (ntiles > g->scan_cap) {
        int r = ensure(g->scan_tmp, g->scan_cap, ntiles);
        if (r) return r;
    }
    const int ndigits = (int)radix_counts_size<kSortItems>(std::max(count, 1));
    if (ndigits > g->digits_cap) {
        int r = ensure(g->digits, g->digits_cap, ndigits);
        if (r) return r;
    }
    g->count = count;
    const unsigned char* ph = (const unsigned char*)photons;
    const int nparts = count > 0 ? std::min(kBoundsBlocks, ceil_div(count, kBlock)) : 0;
    if (count > 0) hipLaunchKernelGGL(k_bounds, dim3(nparts), dim3(kBlock), 0, s, ph, stride, pos_offset, count, g->partial);
    hipLaunchKernelGGL(k_params, dim3(1), dim3(kBlock), 0, s, g->partial, nparts, g->d_params, radius, table, count);
    if (table > 0)
        hipLaunchKernelGGL(k_zero, dim3(std::min(ceil_div(table, kBlock), 4096)), dim3(kBlock), 0, s, g->cell_ends,
                           table);
    if (count > 0) {
        hipLaunchKernelGGL(k_grid_keys, dim3(ceil_div(count, kBlock)), dim3(kBlock), 0, s, ph, stride, pos_offset,
                           count, g->d_params, g->keys_a, g->vals_a);
        int bits = 1;
        while (bits < 31 && (1 << bits) < table) bits++;  // buckets are < table
        int* keys_sorted = nullptr;
        radix_sort_pairs<kSortItems>(g->keys_a, g->vals_a, g->keys_b, g->vals_b, count, bits, g->digits, g->order,
                                     &keys_sorted, s);
        hipLaunchKernelGGL(k_mark_ends, dim3(ceil_div(count, kBlock)), dim3(kBlock), 0, s, keys_sorted, count,
                           g->cell_ends);
        hipLaunchKernelGGL(k_gather_pos, dim3(ceil_div(count, kBlock)), dim3(kBlock), 0, s, ph, stride, pos_offset,
                           count, g->order, g->pos4);
        hipLaunchKernelGGL(k_bucket_box_reset, dim3(ceil_div(count, kBlock)), dim3(kBlock), 0, s, keys_sorted, count,
                           g->bucket_box);
        hipLaunchKernelGGL(k_bucket_box, dim3(ceil_div(count, kBlock)), dim3(kBlock), 0, s, keys_sorted, count, g->pos4,
                           g->bucket_box);
    }
    if (table > 0) {
        hipLaunchKernelGGL(k_max_tiles, dim3(ntiles), dim3(kBlock), 0, s, g->cell_ends, table, g->scan_tmp);
        hipLaunchKernelGGL(k_max_top, dim3(1), dim3(1024), 0, s, g->scan_tmp, ntiles);
        hipLaunchKernelGGL(k_max_apply, dim3(ntiles), dim3(kBlock), 0, s, g->cell_ends, table, g->scan_tmp);
    }
    IMBA_HIP_TRY(hipGetLastError());
    if (sync) IMBA_HIP_TRY(hipStreamSynchronize(s));
    return IMBA_OK;
}

int imba_hashgrid_query(const ImbaHashGrid* g, const void* queries, int query_stride, int query_offset, int nq, int k,
                        int* out_idx, float* out_dist, int* out_count, void* stream) {
    IMBA_CHECK_ARG(g, "grid is NULL");
    if (k < 1 || k > kQueryMaxK) return fail(IMBA_ERR_INVALID_ARG, "k must be in [1, %d]", kQueryMaxK);
    IMBA_CHECK_ARG(nq >= 0, "nq < 0");
    IMBA_CHECK_ARG(query_stride >= 12 && query_stride % 4 == 0 && query_offset >= 0 && query_offset % 4 == 0 &&
                       query_offset + 12 <= query_stride,
                   "query stride/offset must be multiples of 4 with the position inside the record");
    if (nq == 0) return IMBA_OK;
    IMBA_CHECK_ARG(queries && out_idx && out_count, "NULL buffer");
    if (g->table_size == 0) return fail(IMBA_ERR_INVALID_ARG, "grid has not been built");
    bool sync;
    hipStream_t s = resolve_stream(stream, &sync);
    const unsigned char* qs = (const unsigned char*)queries;
    if (k <= 64)
        hipLaunchKernelGGL(k_query_wave, dim3(ceil_div(nq, kBlock / 64)), dim3(kBlock), 0, s, g->d_params,
                           g->cell_ends, g->pos4, g->bucket_box, qs, query_stride, query_offset, nq, k, out_idx,
                           out_dist, out_count);
    else
        hipLaunchKernelGGL(k_query_lds, dim3(nq), dim3(64), (size_t)k * 8, s, g->d_params, g->cell_ends, g->pos4, qs,
                           query_stride, query_offset, nq, k, out_idx, out_dist, out_count);
    IMBA_HIP_TRY(hipGetLastError());
    if (sync) IMBA_HIP_TRY(hipStreamSynchronize(s));
    return IMBA_OK;
}

int imba_hashgrid_reset(ImbaHashGrid* g) {
    IMBA_CHECK_ARG(g, "grid is NULL");
    g->table_size = 0;
    g->count = 0;
    return IMBA_OK;
}

int imba_hashgrid_info(const ImbaHashGrid* g, ImbaHashGridInfo* info) {
    IMBA_CHECK_ARG(g && info, "NULL argument");
    GridParams p{};
    IMBA_HIP_TRY(hipMemcpy(&p, g->d_params, sizeof(p), hipMemcpyDeviceToHost));
    info->photon_count = g->count;
    info->table_size = g->table_size;
    for (int k = 0; k < 3; k++) {
        info->bbox_min[k] = p.bmin[k];
        info->bbox_max[k] = p.bmax[k];
    }
    info->radius = p.radius;
    info->radius_sqr = p.radius_sqr;
    info->cell_size = p.cell_size;
    info->inv_cell_size = p.inv_cell_size;
    info->cell_ends = g->cell_ends;
    info->order = g->order;
    info->positions = (const float*)g->pos4;
    return IMBA_OK;
}

__global__ void k_gather_records(const unsigned char* src, int stride, int n, const int* order, unsigned char* dst) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    const int i = order[s];
    const int* a = reinterpret_cast<const int*>(src + (size_t)i * stride);
    int* b = reinterpret_cast<int*>(dst + (size_t)s * stride);
    for (int k = 0; k < stride / 4; k++) b[k] = a[k];
}

int imba_hashgrid_gather(const ImbaHashGrid* g, const void* src, int stride, void* dst, void* stream) {
    IMBA_CHECK_ARG(g && src && dst, "NULL argument");
    IMBA_CHECK_ARG(stride > 0 && stride % 4 == 0, "stride must be a positive multiple of 4");
    if (g->count == 0) return IMBA_OK;
    bool sync;
    hipStream_t s = resolve_stream(stream, &sync);
    hipLaunchKernelGGL(k_gather_records, dim3(ceil_div(g->count, 256)), dim3(256), 0, s, (const unsigned char*)src,
                       stride, g->count, g->order, (unsigned char*)dst);
    IMBA_HIP_TRY(hipGetLastError());
    if (sync) IMBA_HIP_TRY(hipStreamSynchronize(s));
    return IMBA_OK;
}

}  // extern "C"
