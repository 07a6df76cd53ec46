// Photon hash grid: parameters, device-side state and the one-query-per-wave scan shared by
// the index query (hashgrid.hip, HashGrid::query of rangesearch/rangesearch.h:79-129) and the
// fused vertex-merging launch (vcm.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace imba {

struct GridParams {
    float bmin[3], bmax[3];
    float radius, radius_sqr, cell_size, inv_cell_size;
    int table_size;
    int count;
};

}  // namespace imba

struct ImbaHashGrid {
    int device = 0;
    int table_size = 0;  // grows only (until imba_hashgrid_reset)
    int table_cap = 0;   // buckets allocated
    int count = 0;
    int cap = 0;         // photon capacity of the per-photon arrays
    int* cell_ends = nullptr;
    int* order = nullptr;
    int* keys_a = nullptr;     // sort buffers (photon_count entries each)
    int* vals_a = nullptr;
    int* keys_b = nullptr;
    int* vals_b = nullptr;
    float4* pos4 = nullptr;    // positions in slot order: x, y, z, photon index bits
    uint32_t* bucket_box = nullptr;  // per bucket: bounds of its photons (lo xyz, hi xyz) as
                                     // ordered_bits, written for the non-empty buckets of the last
                                     // build
    int box_cap = 0;
    imba::GridParams* d_params = nullptr;
    float* partial = nullptr;  // 6 per bounds block
    int* scan_tmp = nullptr;   // tile maxima of the ends scan
    int scan_cap = 0;
    int* digits = nullptr;     // radix digit counts
    int digits_cap = 0;
};

namespace imba {

// The hash of rangesearch/rangesearch.h:136-139 (three primes, xor, modulo the table size).
__device__ __forceinline__ int bucket_of(uint32_t x, uint32_t y, uint32_t z, int table) {
    return (int)(((x * 73856093u) ^ (y * 19349663u) ^ (z * 83492791u)) % (uint32_t)table);
}

constexpr int kDppWaveShr1 = 0x138;  // DPP control: lane i reads lane i - 1 of the wave

// A float's bits mapped to an unsigned integer of the same order (negative floats reversed
// below the positive ones), so that integer min / max are float min / max; and back.
__device__ __forceinline__ uint32_t ordered_bits(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float from_ordered_bits(uint32_t e) {
    return __uint_as_float((e & 0x80000000u) ? (e & 0x7FFFFFFFu) : ~e);
}

// Squared distance from (qx, qy, qz) to the box of a bucket's photons: a lower bound of every
// photon's squared distance as the scan computes it (float subtraction, squaring and addition
// are monotonic, so a photon inside the box is no nearer than the box).  Only meaningful for a
// bucket that holds photons.
__device__ __forceinline__ float box_dist2(const uint32_t* box, int b, float qx, float qy, float qz) {
    const uint2* e = reinterpret_cast<const uint2*>(box + 6 * (size_t)b);
    const uint2 e0 = e[0], e1 = e[1], e2 = e[2];
    const float lox = from_ordered_bits(e0.x), loy = from_ordered_bits(e0.y), loz = from_ordered_bits(e1.x);
    const float hix = from_ordered_bits(e1.y), hiy = from_ordered_bits(e2.x), hiz = from_ordered_bits(e2.y);
    const float gx = fmaxf(fmaxf(lox - qx, qx - hix), 0.0f);
    const float gy = fmaxf(fmaxf(loy - qy, qy - hiy), 0.0f);
    const float gz = fmaxf(fmaxf(loz - qz, qz - hiz), 0.0f);
    return gx * gx + gy * gy + gz * gz;
}

// One query per wave.  The eight probed buckets are visited in the reference's order, each
// bucket's photons scanned 64 at a time (lane l takes photon base + l, so loads are contiguous
// across the wave).  Candidates within the limit at the start of the chunk (the radius, or the
// current k-th distance once the list is full) are then inserted one at a time in visiting
// order with the reference's rule; the list itself lives in lanes 0..k-1 (squared distance and
// id per lane), so any k up to 64 fits.  The limit only shrinks, so the candidates it rejects
// would be rejected by the exact test as well: the results are those of the one-at-a-time
// scan.  For the same reason a bucket whose photon box lies beyond the limit when its turn
// comes is skipped whole (none of its photons would pass).  The id is the photon index, or with
// SLOTS the photon's slot in the grid order (pos4[slot]).  Returns the list length
// (wave-uniform); lanes >= count hold (0, -1).  Every lane of the wave must call it.
template <bool SLOTS>
__device__ __forceinline__ int query_wave_list(const GridParams& g, const int* ends, const float4* pos4,
                                               const uint32_t* box, float qx, float qy, float qz, int k,
                                               float& my_dist, int& my_id) {
    const int lane = (int)(threadIdx.x & 63);
    my_dist = 0.0f;
    my_id = -1;
    int count = 0;
    const bool inside = qx >= g.bmin[0] && qy >= g.bmin[1] && qz >= g.bmin[2] &&
                        qx <= g.bmax[0] && qy <= g.bmax[1] && qz <= g.bmax[2];
    if (!(inside && g.count > 0)) return 0;
    const float cx = g.inv_cell_size * (qx - g.bmin[0]);
    const float cy = g.inv_cell_size * (qy - g.bmin[1]);
    const float cz = g.inv_cell_size * (qz - g.bmin[2]);
    const float fx = floorf(cx), fy = floorf(cy), fz = floorf(cz);
    const int px = (int)fx, py = (int)fy, pz = (int)fz;
    const int pxo = px + ((cx - fx) < 0.5f ? -1 : 1);
    const int pyo = py + ((cy - fy) < 0.5f ? -1 : 1);
    const int pzo = pz + ((cz - fz) < 0.5f ? -1 : 1);
    // lane j < 8: probe j's bucket range and the squared distance to its photons' box
    int begin = 0, size = 0;
    float lb = 0.0f;
    if (lane < 8) {
        const int x = (lane & 4) ? pxo : px;
        const int y = (lane & 2) ? pyo : py;
        const int z = (lane & 1) ? pzo : pz;
        const int b = bucket_of((uint32_t)x, (uint32_t)y, (uint32_t)z, g.table_size);
        begin = b == 0 ? 0 : ends[b - 1];
        size = ends[b] - begin;
        lb = box && size > 0 ? box_dist2(box, b, qx, qy, qz) : 0.0f;
    }
    float kth = 0.0f;  // dist[k - 1] once the list is full (wave-uniform)
    for (int j = 0; j < 8; j++) {
        const int beg = __builtin_amdgcn_readlane(begin, j);
        const int n = __builtin_amdgcn_readlane(size, j);
        const float lbj = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(lb), j));
        if (n == 0 || lbj > (count == k ? kth : g.radius_sqr)) continue;
        // photon c of this bucket (x = +inf past its end, which no test accepts); the next
        // chunk's load is issued before the current chunk's insertions
        auto fetch = [&](int c) {
            if (c >= n) return make_float4(__int_as_float(0x7f800000), 0.0f, 0.0f, 0.0f);
            return pos4[beg + c];
        };
        float4 p_next = fetch(lane);
        for (int base = 0; base < n; base += 64) {
            const bool valid = base + lane < n;
            const float4 p = p_next;
            if (base + 64 < n) p_next = fetch(base + 64 + lane);
            float d2 = 0.0f;
            int pid = -1;
            if (valid) {
                const float dx = qx - p.x, dy = qy - p.y, dz = qz - p.z;
                d2 = dx * dx + dy * dy + dz * dz;
                pid = SLOTS ? beg + base + lane : __float_as_int(p.w);
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
                // keep the first count-1 entries, insert dd after every entry <= dd
                const int keep = count - 1;
                const int pos = __popcll(__ballot(lane < keep && my_dist <= dd));
                // the entries one lane down (wave shift right by one lane, DPP)
                const float up_d = __int_as_float(__builtin_amdgcn_update_dpp(
                    __float_as_int(my_dist), __float_as_int(my_dist), kDppWaveShr1, 0xF, 0xF, false));
                const int up_i = __builtin_amdgcn_update_dpp(my_id, my_id, kDppWaveShr1, 0xF, 0xF, false);
                if (lane > pos && lane <= keep) { my_dist = up_d; my_id = up_i; }
                if (lane == pos) { my_dist = dd; my_id = id; }
                if (count == k) kth = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(my_dist), k - 1));
            }
        }
    }
    return count;
}

}  // namespace imba
