// Stable LSD radix sort of (int key, int value) pairs by 8-bit digits, shared by the material
// sort and the hash-grid build.
//
// One pass over a digit:
//   1. k_radix_hist: per tile of 256 x ITEMS entries, a 256-bin histogram (LDS atomics),
//      stored digit-major so that one exclusive scan gives every (digit, tile) its first
//      position;
//   2. k_radix_scan_sums / _top / _apply: exclusive scan of the 256 x tiles counts;
//   3. k_radix_scatter: per tile, in ITEMS chunks of 256 entries in order, each wave ranks its
//      lanes among the lanes with the same digit (eight ballots, one per digit bit), the waves'
//      per-digit counts in LDS give the offsets of the earlier waves, and a running per-digit
//      count carries over between chunks.
// Entries with equal digits keep their order, so after the passes over all key bits the pairs
// are ordered by key and, for equal keys, as they came in.
// Traffic per entry and pass: key read twice, value once, both written once (20 B).
#pragma once

#include <hip/hip_runtime.h>

#include <utility>

namespace imba {
namespace {

constexpr int kRadixBlock = 256;
constexpr int kRadixWaves = kRadixBlock / 64;

template <int ITEMS>
__global__ __launch_bounds__(kRadixBlock) void k_radix_hist(const int* __restrict__ keys, int n, int shift,
                                                            int n_tiles, int* counts) {
    __shared__ int hist[256];
    hist[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * kRadixBlock * ITEMS;
#pragma unroll
    for (int it = 0; it < ITEMS; it++) {
        const int i = base + it * kRadixBlock + (int)threadIdx.x;
        if (i < n) atomicAdd(&hist[(keys[i] >> shift) & 255], 1);
    }
    __syncthreads();
    counts[(size_t)threadIdx.x * n_tiles + blockIdx.x] = hist[threadIdx.x];
}

// Exclusive scan of m ints in place, in three coalesced launches: per-segment sums
// (kRadixScanSeg ints per block), a one-block scan of the segment sums, then every segment
// scanned again from its offset.
constexpr int kRadixScanItems = 8;
constexpr int kRadixScanSeg = kRadixBlock * kRadixScanItems;

__device__ __forceinline__ int radix_block_sum(int v, int* ws) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    int t = 0;
    for (int w = 0; w < kRadixWaves; w++) t += ws[w];
    __syncthreads();
    return t;
}

__global__ __launch_bounds__(kRadixBlock) void k_radix_scan_sums(const int* v, int m, int* seg_sums) {
    __shared__ int ws[kRadixWaves];
    const int base = blockIdx.x * kRadixScanSeg;
    int s = 0;
#pragma unroll
    for (int it = 0; it < kRadixScanItems; it++) {
        const int i = base + it * kRadixBlock + (int)threadIdx.x;
        if (i < m) s += v[i];
    }
    s = radix_block_sum(s, ws);
    if (threadIdx.x == 0) seg_sums[blockIdx.x] = s;
}

// Exclusive scan of the (few) segment sums, one block.
__global__ __launch_bounds__(1024) void k_radix_scan_top(int* sums, int nseg) {
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int per = (nseg + 1023) / 1024;
    const int b0 = min(nseg, t * per), b1 = min(nseg, b0 + per);
    int s = 0;
    for (int b = b0; b < b1; b++) s += sums[b];
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
        const int c = sums[b];
        sums[b] = run;
        run += c;
    }
}

__global__ __launch_bounds__(kRadixBlock) void k_radix_scan_apply(int* v, int m, const int* seg_offsets) {
    __shared__ int wsum[kRadixWaves];
    const int base = blockIdx.x * kRadixScanSeg;
    int run = seg_offsets[blockIdx.x];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int it = 0; it < kRadixScanItems; it++) {
        const int i = base + it * kRadixBlock + (int)threadIdx.x;
        const int c = i < m ? v[i] : 0;
        int x = c;  // inclusive wave scan
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[wave] = x;
        __syncthreads();
        int woff = 0, total = 0;
        for (int w = 0; w < kRadixWaves; w++) {
            woff += w < wave ? wsum[w] : 0;
            total += wsum[w];
        }
        if (i < m) v[i] = run + woff + x - c;
        run += total;
        __syncthreads();
    }
}

template <int ITEMS>
__global__ __launch_bounds__(kRadixBlock) void k_radix_scatter(const int* __restrict__ keys_in,
                                                               const int* __restrict__ vals_in, int n, int shift,
                                                               int n_tiles, const int* __restrict__ offsets,
                                                               int* keys_out, int* vals_out) {
    __shared__ int wave_count[kRadixWaves][256];
    __shared__ int running[256];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const unsigned long long lt = lane ? (~0ull >> (64 - lane)) : 0ull;  // lanes below this one
    running[threadIdx.x] = offsets[(size_t)threadIdx.x * n_tiles + blockIdx.x];
    const int base = blockIdx.x * kRadixBlock * ITEMS;
    for (int it = 0; it < ITEMS; it++) {
        for (int w = 0; w < kRadixWaves; w++) wave_count[w][threadIdx.x] = 0;
        __syncthreads();
        const int i = base + it * kRadixBlock + (int)threadIdx.x;
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
        for (int w = 0; w < kRadixWaves; w++) add += wave_count[w][threadIdx.x];
        running[threadIdx.x] += add;
        __syncthreads();
    }
}

template <int ITEMS>
inline int radix_tiles(int n) {
    return (n + kRadixBlock * ITEMS - 1) / (kRadixBlock * ITEMS);
}

// Ints of the `counts` workspace radix_sort_pairs needs for n pairs: the digit counts, then
// the scan's segment sums.
template <int ITEMS>
inline size_t radix_counts_size(int n) {
    const size_t m = (size_t)256 * radix_tiles<ITEMS>(n);
    return m + (m + kRadixScanSeg - 1) / kRadixScanSeg + 1;
}

// Sorts n pairs (keys in [0, 2^bits)) from (keys_a, vals_a) using (keys_b, vals_b) as the
// other buffer; `counts` holds radix_counts_size<ITEMS>(n) ints.  The last pass writes the
// values to vals_final (may be vals_b); *keys_sorted is set to the buffer holding the sorted
// keys.  Launch errors are left for the caller's hipGetLastError.
template <int ITEMS>
void radix_sort_pairs(int* keys_a, int* vals_a, int* keys_b, int* vals_b, int n, int bits, int* counts,
                      int* vals_final, int** keys_sorted, hipStream_t s) {
    const int passes = bits <= 8 ? 1 : (bits + 7) / 8;
    const int n_tiles = radix_tiles<ITEMS>(n);
    const int m = 256 * n_tiles;
    const int nseg = (m + kRadixScanSeg - 1) / kRadixScanSeg;
    int* seg = counts + m;
    for (int p = 0; p < passes; p++) {
        const int shift = 8 * p;
        const bool last = p == passes - 1;
        int* vo = last ? vals_final : vals_b;
        hipLaunchKernelGGL(k_radix_hist<ITEMS>, dim3(n_tiles), dim3(kRadixBlock), 0, s, keys_a, n, shift, n_tiles,
                           counts);
        hipLaunchKernelGGL(k_radix_scan_sums, dim3(nseg), dim3(kRadixBlock), 0, s, counts, m, seg);
        hipLaunchKernelGGL(k_radix_scan_top, dim3(1), dim3(1024), 0, s, seg, nseg);
        hipLaunchKernelGGL(k_radix_scan_apply, dim3(nseg), dim3(kRadixBlock), 0, s, counts, m, seg);
        hipLaunchKernelGGL(k_radix_scatter<ITEMS>, dim3(n_tiles), dim3(kRadixBlock), 0, s, keys_a, vals_a, n, shift,
                           n_tiles, counts, keys_b, vo);
        std::swap(keys_a, keys_b);
        std::swap(vals_a, vals_b);
    }
    *keys_sorted = keys_a;
}

}  // namespace
}  // namespace imba
