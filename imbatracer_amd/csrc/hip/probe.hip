// Read-bandwidth probe for roofline reporting: how fast the chip streams a buffer of a given
// size out of the level of the memory hierarchy that holds it (L2 for a few MB, the Infinity
// Cache for a few tens to hundreds of MB, HBM beyond).  A scene whose nodes and triangles fit
// in the caches is bounded by these rates rather than by HBM.
//
// Every pass reads the whole buffer once with 16-byte loads, rotated: pass r reads element
// (k + (r mod 16) * shift) mod n at its k-th position, shift = n / 16 + 1.  The passes are
// dealt to G <= 8 groups of blocks (group q takes passes q, q + G, ...), G as large as fits the
// 8 resident 256-thread blocks of a CU, so that a small buffer still keeps every CU busy.  All
// blocks are resident at once, each on one CU reading one slice of the buffer; the groups
// running together read different rotations, and a CU reads an address again only after
// 16 / G steps, i.e. after reading 16 x bytes / CUs of other data (128 KB for 2 MB on 256 CUs,
// four times the 32 KB vector L1), so every load goes to the L2 or beyond.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.h"

namespace imba {
namespace {

constexpr int kProbeBlock = 256;

__global__ __launch_bounds__(kProbeBlock) void k_read_probe(const uint4* __restrict__ p, long long n, int passes,
                                                           int groups, int blocks_per_group, uint4* sink) {
    const int group = blockIdx.x / blocks_per_group;
    const long long g = (long long)(blockIdx.x - group * blocks_per_group) * kProbeBlock + threadIdx.x;
    const long long stride = (long long)blocks_per_group * kProbeBlock;
    const long long shift = n / 16 + 1;
    uint4 acc = make_uint4(0u, 0u, 0u, 0u);
    for (int r = group; r < passes; r += groups) {
        long long off = (long long)(r & 15) * shift;  // < n + 16
        if (off >= n) off -= n;
        long long i = g + off;
        if (i >= n) i -= n;
        for (long long k = g; k < n; k += stride) {
            const uint4 v = p[i];
            acc.x ^= v.x; acc.y += v.y; acc.z ^= v.z; acc.w += v.w;
            i += stride;
            if (i >= n) i -= n;
        }
    }
    // keeps the loads alive; the condition is practically never true and the slot is in bounds
    if ((acc.x ^ acc.y ^ acc.z ^ acc.w) == 0x9E3779B9u) sink[(long long)blockIdx.x * kProbeBlock + threadIdx.x] = acc;
}

}  // namespace
}  // namespace imba

using namespace imba;

extern "C" int imba_read_bandwidth(const void* buffer, size_t bytes, int passes, void* sink, size_t sink_bytes,
                                   void* stream, float* ms) {
    IMBA_CHECK_ARG(buffer && sink && ms, "NULL argument");
    IMBA_CHECK_ARG(passes > 0 && bytes >= 16, "passes must be positive and the buffer at least 16 bytes");
    const DeviceProps* p = device_props();
    if (!p) return fail(IMBA_ERR_NO_DEVICE, "no HIP device");
    const long long n = (long long)(bytes / 16);
    const long long slots = 8LL * p->cu_count;  // resident 256-thread blocks
    const int per_group = (int)std::min<long long>((n + kProbeBlock - 1) / kProbeBlock, slots);
    const int groups = (int)std::max<long long>(1, std::min<long long>({(long long)passes, slots / per_group, 8}));
    const long long threads = (long long)groups * per_group * kProbeBlock;
    if (sink_bytes < (size_t)threads * 16)
        return fail(IMBA_ERR_CAPACITY, "sink needs %lld bytes", threads * 16);
    bool sync;
    hipStream_t s = resolve_stream(stream, &sync);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    IMBA_HIP_TRY(hipEventCreate(&e0));
    IMBA_HIP_TRY(hipEventCreate(&e1));
    IMBA_HIP_TRY(hipEventRecord(e0, s));
    hipLaunchKernelGGL(k_read_probe, dim3((unsigned)(groups * per_group)), dim3(kProbeBlock), 0, s,
                       (const uint4*)buffer, n, passes, groups, per_group, (uint4*)sink);
    hipError_t err = hipGetLastError();
    if (err == hipSuccess) err = hipEventRecord(e1, s);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    if (err == hipSuccess) err = hipEventElapsedTime(ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    IMBA_HIP_TRY(err);
    return IMBA_OK;
}
