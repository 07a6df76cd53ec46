// Host side of the traversal kernels: launch configuration, scratch, C ABI entry points and
// the reference FFI (traversal_gpu.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "common.h"
#include "scene.h"
#include "trace_kernels.h"
#include "../host/wide_bvh.h"

namespace imba {

namespace {

// Spill stacks (stack entries beyond the LDS part), one buffer per device and stream: launches
// on one stream are ordered, so they can share it without waiting.
std::mutex g_scratch_mutex;
std::map<std::pair<int, hipStream_t>, Scratch> g_spill;

// Work counters of the persistent kernel, per device: kCounterSlots slots handed out round
// robin to uncaptured launches (each launch's last wave sets its slot back to zero, so a slot
// is reused only after 1024 later launches), and kCaptureSlots slots of which every launch
// recorded into a graph gets one of its own for good (a graph replays its launches with the
// slot they were recorded with, so no later launch may share it).
constexpr int kCounterSlots = 1024;
constexpr int kCaptureSlots = 4096;
std::mutex g_counter_mutex;
int* g_counters[64] = {nullptr};
unsigned g_counter_next[64] = {0};
unsigned g_capture_next[64] = {0};

int counter_slot(int dev, bool capturing, int** slot) {
    std::lock_guard<std::mutex> lock(g_counter_mutex);
    if (!g_counters[dev]) {
        int* p = nullptr;
        const size_t bytes = sizeof(int) * 2 * (kCounterSlots + kCaptureSlots);
        IMBA_HIP_TRY(hipMalloc((void**)&p, bytes));
        IMBA_HIP_TRY(hipMemset(p, 0, bytes));
        IMBA_HIP_TRY(hipDeviceSynchronize());
        g_counters[dev] = p;
    }
    if (capturing) {
        if (g_capture_next[dev] >= (unsigned)kCaptureSlots)
            return fail(IMBA_ERR_CAPACITY, "more than %d persistent launches recorded into graphs", kCaptureSlots);
        *slot = g_counters[dev] + 2 * (kCounterSlots + g_capture_next[dev]++);
    } else {
        *slot = g_counters[dev] + 2 * (g_counter_next[dev]++ % kCounterSlots);
    }
    return IMBA_OK;
}

bool is_capturing(hipStream_t s) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

// Allocation is not allowed while `s` is being captured into a graph.
int not_capturing(hipStream_t s, const char* what) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    IMBA_HIP_TRY(hipStreamIsCapturing(s, &cs));
    if (cs != hipStreamCaptureStatusNone)
        return fail(IMBA_ERR_INVALID_ARG, "graph capture: this launch needs %s; make the same call once "
                                          "on this stream before capturing", what);
    return IMBA_OK;
}

template <int BLOCK, int STACK_LDS, bool ANY, bool MASKED, bool STATS, int TREE>
int launch_one(const TraceArgs& args_in, int max_stack, int blocks_per_cu, int refill, hipStream_t s) {
    TraceArgs a = args_in;
    const DeviceProps* p = device_props();
    if (!p) return fail(IMBA_ERR_NO_DEVICE, "no HIP device");
    const bool dyn = refill > 0;
    void (*kern)(TraceArgs) = k_trace<BLOCK, STACK_LDS, ANY, MASKED, STATS, TREE>;
    void (*kern_dyn)(TraceArgs, int) = k_trace_dyn<BLOCK, STACK_LDS, ANY, MASKED, STATS, TREE>;
    const size_t tl_bytes = TREE == 2 ? (size_t)a.tl_nodes * 64 : 0;  // dynamic LDS: the treelet copy
    if (TREE != 2) a.tl_nodes = 0;
    int resident = 0;
    if (dyn)
        IMBA_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, kern_dyn, BLOCK, tl_bytes));
    else
        IMBA_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, kern, BLOCK, tl_bytes));
    if (resident <= 0) resident = 1;
    if (blocks_per_cu > 0) resident = std::min(resident, blocks_per_cu);
    const int needed = ceil_div(a.count, BLOCK);
    // One thread per ray, so that the dispatcher starts a new block wherever one finishes (the
    // rays of a wave end at very different times); a grid capped at the resident blocks, each
    // thread striding over several rays, left the SIMDs a quarter occupied on average and was
    // 4% slower on the path-tracing queues.  The persistent variant and an explicit
    // blocks_per_cu keep the resident-sized grid.
    const int grid = (dyn || blocks_per_cu > 0) ? std::max(1, std::min(needed, resident * p->cu_count))
                                                : std::max(1, needed);
    a.spill = nullptr;
    a.spill_depth = 0;
    if (max_stack > STACK_LDS) {
        a.spill_depth = max_stack - STACK_LDS;
        std::lock_guard<std::mutex> lock(g_scratch_mutex);
        Scratch& sc = g_spill[std::make_pair(p->device, s)];
        const size_t need = (size_t)a.spill_depth * grid * BLOCK * sizeof(int);
        if (sc.bytes < need) {
            int r = not_capturing(s, "a larger spill stack");
            if (r) return r;
        }
        // growing frees the old buffer after a device-wide synchronisation (grow_buffer)
        int r = grow_buffer(sc, need);
        if (r) return r;
        a.spill = (int*)sc.ptr;
    }
    if (dyn) {
        if (!g_counters[p->device & 63]) {
            int r = not_capturing(s, "the persistent kernel's work counters");
            if (r) return r;
        }
        int r = counter_slot(p->device & 63, is_capturing(s), &a.work_counter);
        if (r) return r;
        hipLaunchKernelGGL(kern_dyn, dim3(grid), dim3(BLOCK), tl_bytes, s, a, std::min(refill, 64));
    } else {
        hipLaunchKernelGGL(kern, dim3(grid), dim3(BLOCK), tl_bytes, s, a);
    }
    IMBA_HIP_TRY(hipGetLastError());
    return IMBA_OK;
}

// Default launch: 256 threads, 16 LDS stack entries per lane, one ray per thread (measured
// fastest on both the 65K-triangle and the 21M-triangle workloads).
constexpr int kDefaultBlock = 256;
constexpr int kBatchTreeletMax = 64;  // batches: nodes per wave's LDS treelet copy (larger requests clamp)
constexpr int kDefaultStack = 16;
constexpr int kDefaultWidth = 2;

// Tree a configuration asks for: 0 the BVH2, 1 the 4-wide tree, 2 its quantized form; -1 if
// the configuration is invalid.
int tree_code(const ImbaTraceConfig* cfg) {
    const int w = cfg && cfg->bvh_width ? cfg->bvh_width : kDefaultWidth;
    if (w == 2) return (cfg && cfg->quantized) ? -1 : 0;
    if (w == 4) return (cfg && cfg->quantized) ? 2 : 1;
    return -1;
}
int width_of(const ImbaTraceConfig* cfg) {
    const int t = tree_code(cfg);
    return t < 0 ? 0 : (t == 0 ? 2 : 4);
}

template <bool ANY, bool MASKED, bool STATS, int TREE>
int dispatch_shape(const TraceArgs& a, int max_stack, const ImbaTraceConfig* cfg, hipStream_t s) {
    const int block = cfg && cfg->block_size ? cfg->block_size : kDefaultBlock;
    const int stack = cfg && cfg->stack_lds ? cfg->stack_lds : kDefaultStack;
    const int bpc = cfg ? cfg->blocks_per_cu : 0;
    const int refill = cfg ? cfg->persistent : 0;
    if (STATS) {  // visit counts do not depend on the launch shape
        return launch_one<kDefaultBlock, kDefaultStack, ANY, MASKED, STATS, TREE>(a, max_stack, 0, 0, s);
    } else {
#define IMBA_SHAPE(B, S) \
    if (block == B && stack == S) return launch_one<B, S, ANY, MASKED, STATS, TREE>(a, max_stack, bpc, refill, s);
        IMBA_SHAPE(256, 16)
        IMBA_SHAPE(256, 8)
        IMBA_SHAPE(256, 32)
        IMBA_SHAPE(128, 16)
        IMBA_SHAPE(128, 32)
        IMBA_SHAPE(64, 32)
        IMBA_SHAPE(64, 64)
#undef IMBA_SHAPE
    }
    return fail(IMBA_ERR_INVALID_ARG, "unsupported trace shape block=%d stack_lds=%d "
                "(supported: 256/16, 256/8, 256/32, 128/16, 128/32, 64/32, 64/64)", block, stack);
}

template <bool STATS, int TREE>
int dispatch_flags(const TraceArgs& a, int max_stack, bool any, bool masked, const ImbaTraceConfig* cfg,
                   hipStream_t s) {
    return any ? (masked ? dispatch_shape<true, true, STATS, TREE>(a, max_stack, cfg, s)
                         : dispatch_shape<true, false, STATS, TREE>(a, max_stack, cfg, s))
               : (masked ? dispatch_shape<false, true, STATS, TREE>(a, max_stack, cfg, s)
                         : dispatch_shape<false, false, STATS, TREE>(a, max_stack, cfg, s));
}

// Scene arrays of one tree: the BVH2 as given (0), the derived 4-wide tree (1) or its quantized
// form (2).
struct TreeView {
    const ImbaSceneDesc* d;
    const SceneInfo* info;
    int tree;
};

int launch_trace(const TreeView& t, int* d_err, const ImbaRay* rays, ImbaHit* hits, uint32_t* counts, int count,
                 bool any, bool masked, const ImbaTraceConfig* cfg, void* stream, const int* count_dev = nullptr,
                 uint32_t* depth_hist = nullptr) {
    if (count < 0) return fail(IMBA_ERR_INVALID_ARG, "count < 0");
    if (count == 0) return IMBA_OK;
    if (!rays || !hits) return fail(IMBA_ERR_INVALID_ARG, "rays or hits is NULL");
    const ImbaSceneDesc& d = *t.d;
    TraceArgs a{};
    a.count_dev = count_dev;
    a.nodes = (const float4*)d.nodes;       a.n_nodes = d.node_count;
    a.inst = (const float4*)d.instances;    a.n_inst = d.instance_count;
    a.tris = (const float4*)d.tris;         a.n_tris4 = d.tri_vec4_count;
    a.indices = (const int4*)d.indices;     a.n_tri_ids = d.index_count / 4;
    a.texcoords = (const float2*)d.texcoords; a.n_texcoords = d.texcoord_count;
    a.masks = (const int*)d.masks;          a.n_masks = d.mask_count;
    a.mask_buf = d.mask_buffer;             a.mask_buf_size = d.mask_buffer_size;
    a.rays = (const float4*)rays;
    a.hits = (int4*)hits;
    a.count = count;
    a.root = d.root;
    a.counts = counts;
    a.depth_hist = depth_hist;
    a.err = d_err;
    const bool stats = counts || depth_hist;
    if (cfg && cfg->treelet) {
        if (cfg->treelet < 0 || cfg->treelet > kTreeletMaxNodes)
            return fail(IMBA_ERR_INVALID_ARG, "treelet must be 0..%d nodes", kTreeletMaxNodes);
        if (t.tree != 2) return fail(IMBA_ERR_INVALID_ARG, "treelet needs the quantized 4-wide tree");
        if (!stats) {  // visit counts do not depend on where a node is read from
            a.tl_base = t.info->treelet_base;
            a.tl_nodes = std::min(cfg->treelet, t.info->treelet_nodes);
        }
    }
    bool sync;
    hipStream_t s = resolve_stream(stream, &sync);
    const int ms = t.info->max_stack;
    int r;
    if (stats) {
        r = t.tree == 2 ? dispatch_flags<true, 2>(a, ms, any, masked, nullptr, s)
          : t.tree == 1 ? dispatch_flags<true, 1>(a, ms, any, masked, nullptr, s)
                        : dispatch_flags<true, 0>(a, ms, any, masked, nullptr, s);
    } else {
        r = t.tree == 2 ? dispatch_flags<false, 2>(a, ms, any, masked, cfg, s)
          : t.tree == 1 ? dispatch_flags<false, 1>(a, ms, any, masked, cfg, s)
                        : dispatch_flags<false, 0>(a, ms, any, masked, cfg, s);
    }
    if (r) return r;
    if (sync) IMBA_HIP_TRY(hipStreamSynchronize(s));
    return IMBA_OK;
}

// The tree of `s` that `cfg` asks for.
int pick_tree(const ImbaTraceConfig* cfg, const ImbaSceneDesc* d2, const SceneInfo* i2, const ImbaSceneDesc* d4,
              const ImbaSceneDesc* d4q, const SceneInfo* i4, TreeView* t) {
    const int code = tree_code(cfg);
    if (code < 0) return fail(IMBA_ERR_INVALID_ARG, "bvh_width must be 0, 2 or 4 (quantized only with 4)");
    if (code == 2 && !d4q->nodes) return fail(IMBA_ERR_INVALID_ARG, "the scene has no quantized 4-wide tree");
    *t = code == 0 ? TreeView{d2, i2, 0} : code == 1 ? TreeView{d4, i4, 1} : TreeView{d4q, i4, 2};
    return IMBA_OK;
}

int tree_of(const ImbaScene* s, const ImbaTraceConfig* cfg, TreeView* t) {
    return pick_tree(cfg, &s->dev, &s->info, &s->dev4, &s->dev4q, &s->info4, t);
}

// ---------------------------------------------------------------------------------------
// Reference FFI: the caller passes only device pointers and the ray count, and may rewrite or
// reallocate the arrays between calls (Scene::upload_top_level_accel copies a new top level
// into the arrays it has, render/scene.cpp:166-188).  So nothing derived from the arrays'
// contents is kept from one call to the next: each call takes the array sizes from the
// allocations that hold the pointers at that moment, traverses the BVH2 as given with the mask
// test on, and gives every ray a stack of kFfiStack entries (kDefaultStack in LDS, the rest in
// the per-stream spill buffer).  The kernel bounds-checks every index against those sizes, so a
// call is memory safe whatever the arrays hold; a ray that needs a deeper stack is counted as
// an error (imba_ffi_error_count).
constexpr int kFfiStack = 64;
std::mutex g_ffi_mutex;
int* g_ffi_err[64] = {nullptr};  // per device: rays of FFI calls that met a malformed structure

size_t bytes_after(const void* p) {
    if (!p) return 0;
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)p) != hipSuccess || !base) return 0;
    return (size_t)((const char*)base + size - (const char*)p);
}

int ffi_err_counter(int** out) {
    int dev = 0;
    IMBA_HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_ffi_mutex);
    int*& c = g_ffi_err[dev & 63];
    if (!c) {
        IMBA_HIP_TRY(hipMalloc((void**)&c, sizeof(int)));
        IMBA_HIP_TRY(hipMemset(c, 0, sizeof(int)));
        IMBA_HIP_TRY(hipDeviceSynchronize());
    }
    *out = c;
    return IMBA_OK;
}

// `given` holds device pointers and, where known, element counts (0 = the count the
// allocation holding the pointer has room for, from the pointer to the allocation's end).
int ffi_launch(const ImbaSceneDesc& given, const ImbaRay* rays, ImbaHit* hits, int count, bool any, void* stream) {
    auto count_of = [](int given_count, const void* p, size_t rec) {
        if (given_count > 0) return given_count;
        return (int)std::min<size_t>(bytes_after(p) / rec, INT32_MAX);
    };
    ImbaSceneDesc d = given;
    d.node_count = count_of(given.node_count, given.nodes, sizeof(ImbaNode));
    d.instance_count = count_of(given.instance_count, given.instances, sizeof(ImbaInstanceNode));
    d.tri_vec4_count = count_of(given.tri_vec4_count, given.tris, sizeof(ImbaVec4));
    d.index_count = count_of(given.index_count, given.indices, sizeof(int)) / 4 * 4;
    d.texcoord_count = count_of(given.texcoord_count, given.texcoords, sizeof(ImbaVec2));
    d.mask_count = count_of(given.mask_count, given.masks, sizeof(ImbaTransparencyMask));
    d.mask_buffer_size = count_of(given.mask_buffer_size, given.mask_buffer, 1);
    if (d.node_count == 0 || d.instance_count == 0 || d.tri_vec4_count == 0)
        return fail(IMBA_ERR_INVALID_ARG, "scene arrays are not device allocations");
    if (d.root < 0 || d.root >= d.node_count) return fail(IMBA_ERR_INVALID_ARG, "root %d outside the node array", d.root);
    int* err = nullptr;
    int r = ffi_err_counter(&err);
    if (r) return r;
    SceneInfo info;
    info.max_stack = kFfiStack;
    info.all_opaque = false;
    const TreeView t{&d, &info, 0};
    // the mask test runs unless the scene has no mask data at all (then every hit is opaque)
    const bool masked = d.index_count > 0 && d.mask_count > 0 && d.mask_buffer_size > 0;
    return launch_trace(t, err, rays, hits, nullptr, count, any, masked, nullptr, stream);
}

void ffi_trace(bool any, int root, void* nodes, void* instances, void* tris, void* rays, void* hits, int* indices,
               void* texcoords, void* masks, char* mask_buffer, int count) {
    ImbaSceneDesc given{};
    given.root = root;
    given.nodes = (const ImbaNode*)nodes;
    given.instances = (const ImbaInstanceNode*)instances;
    given.tris = (const ImbaVec4*)tris;
    given.indices = indices;
    given.texcoords = (const ImbaVec2*)texcoords;
    given.masks = (const ImbaTransparencyMask*)masks;
    given.mask_buffer = (const uint8_t*)mask_buffer;
    const int r = ffi_launch(given, (const ImbaRay*)rays, (ImbaHit*)hits, count, any, nullptr);
    if (r) fprintf(stderr, "imbatracer_amd: %s_gpu_masked_instanced failed: %s\n", any ? "occluded" : "intersect",
                   imba_last_error());
}

TraceArgs scene_args(const ImbaSceneDesc& d, int* d_err) {
    TraceArgs a{};
    a.nodes = (const float4*)d.nodes;       a.n_nodes = d.node_count;
    a.inst = (const float4*)d.instances;    a.n_inst = d.instance_count;
    a.tris = (const float4*)d.tris;         a.n_tris4 = d.tri_vec4_count;
    a.indices = (const int4*)d.indices;     a.n_tri_ids = d.index_count / 4;
    a.texcoords = (const float2*)d.texcoords; a.n_texcoords = d.texcoord_count;
    a.masks = (const int*)d.masks;          a.n_masks = d.mask_count;
    a.mask_buf = d.mask_buffer;             a.mask_buf_size = d.mask_buffer_size;
    a.root = d.root;
    a.err = d_err;
    return a;
}

}  // namespace
}  // namespace imba

// A batch of ray queues traversed together (imba_trace_batch_*): the queue table lives in
// device memory, with per-group work counters and spill stacks of its own, so that a launch
// allocates nothing and can be captured into a graph.
struct ImbaTraceBatch {
    int device = 0;
    int n = 0;
    imba::BatchGroups groups{};
    int max_stack[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};  // per tree width (2, 4) and group
    long long capacity_chunks[4] = {0, 0, 0, 0};
    imba::BatchItem* d_items[3] = {nullptr, nullptr, nullptr};  // queue tables per tree (BVH2, 4-wide, quantized)
    bool quantized_ok = true;                              // every queue's scene has the quantized tree
    int cap = 0;                                           // queues the tables hold room for
    imba::BatchItem* staging[2] = {nullptr, nullptr};      // pinned host copies of both tables
    hipEvent_t staged[2] = {nullptr, nullptr};             // last copy out of each staging buffer
    int next_staging = 0;
    int* d_group_chunks = nullptr;  // 4
    int* d_counters = nullptr;      // 2 per group
    imba::Scratch spill[4];
    bool captured = false;  // launched into a graph: the tables may no longer move, and the
                            // queues per launch group (kernel arguments) may no longer change
    imba::BatchGroups captured_groups{};
};

namespace imba {
namespace {

template <int BLOCK, int STACK_LDS, bool ANY, bool MASKED, int TREE>
int launch_batch_group(ImbaTraceBatch* b, int g, int blocks_per_cu, int tl_max, hipStream_t s) {
    const DeviceProps* p = device_props();
    if (!p) return fail(IMBA_ERR_NO_DEVICE, "no HIP device");
    auto kern = k_trace_batch<BLOCK, STACK_LDS, ANY, MASKED, TREE, false>;
    if constexpr (TREE == 2) {
        if (tl_max > 0) kern = k_trace_batch<BLOCK, STACK_LDS, ANY, MASKED, TREE, true>;
    } else {
        tl_max = 0;
    }
    const int max_stack = b->max_stack[TREE ? 1 : 0][g];
    const size_t tl_bytes = (size_t)(BLOCK / 64) * tl_max * 64;  // one treelet copy per wave
    int resident = 0;
    IMBA_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, kern, BLOCK, tl_bytes));
    if (resident <= 0) resident = 1;
    if (blocks_per_cu > 0) resident = std::min(resident, blocks_per_cu);
    const long long needed = (b->capacity_chunks[g] * 64 + BLOCK - 1) / BLOCK;
    const int grid = (int)std::max<long long>(1, std::min<long long>(needed, (long long)resident * p->cu_count));
    int* spill = nullptr;
    int spill_depth = 0;
    if (max_stack > STACK_LDS) {
        spill_depth = max_stack - STACK_LDS;
        const size_t need = (size_t)spill_depth * grid * BLOCK * sizeof(int);
        if (b->spill[g].bytes < need) {
            int r = not_capturing(s, "a larger spill stack");
            if (r) return r;
            r = grow_buffer(b->spill[g], need);
            if (r) return r;
        }
        spill = (int*)b->spill[g].ptr;
    }
    hipLaunchKernelGGL(kern, dim3(grid), dim3(BLOCK), tl_bytes, s, b->d_items[TREE], b->groups.first[g],
                       b->groups.first[g + 1] - b->groups.first[g], b->d_group_chunks + g, b->d_counters + 2 * g,
                       spill, spill_depth, tl_max);
    IMBA_HIP_TRY(hipGetLastError());
    return IMBA_OK;
}

template <bool ANY, bool MASKED, int TREE>
int dispatch_batch(ImbaTraceBatch* b, int g, const ImbaTraceConfig* cfg, hipStream_t s) {
    const int block = cfg && cfg->block_size ? cfg->block_size : kDefaultBlock;
    const int stack = cfg && cfg->stack_lds ? cfg->stack_lds : kDefaultStack;
    const int bpc = cfg ? cfg->blocks_per_cu : 0;
    const int tl = cfg ? std::min(cfg->treelet, kBatchTreeletMax) : 0;
#define IMBA_SHAPE(B, S) \
    if (block == B && stack == S) return launch_batch_group<B, S, ANY, MASKED, TREE>(b, g, bpc, tl, s);
    IMBA_SHAPE(256, 16)
    IMBA_SHAPE(256, 8)
    IMBA_SHAPE(256, 32)
    IMBA_SHAPE(128, 16)
    IMBA_SHAPE(128, 32)
    IMBA_SHAPE(64, 32)
    IMBA_SHAPE(64, 64)
#undef IMBA_SHAPE
    return fail(IMBA_ERR_INVALID_ARG, "unsupported trace shape block=%d stack_lds=%d "
                "(supported: 256/16, 256/8, 256/32, 128/16, 128/32, 64/32, 64/64)", block, stack);
}

void free_batch(ImbaTraceBatch* b) {
    if (!b) return;
    for (auto* t : b->d_items) hipFree(t);
    hipFree(b->d_group_chunks);
    hipFree(b->d_counters);
    for (auto& sp : b->spill) hipFree(sp.ptr);
    for (int k = 0; k < 2; k++) {
        if (b->staging[k]) hipHostFree(b->staging[k]);
        if (b->staged[k]) hipEventDestroy(b->staged[k]);
    }
    delete b;
}

// Builds the group-contiguous queue tables (one per tree width) and puts them in device memory:
// with a stream, through pinned staging buffers and copies on that stream (two staging buffers
// used in turn, each reused only after its previous copy has completed); without, synchronously.
int set_tables(ImbaTraceBatch* b, const ImbaTraceItem* items, int n_items, hipStream_t stream) {
    std::vector<BatchItem> table, table4, table4q;
    table.reserve(n_items);
    table4.reserve(n_items);
    table4q.reserve(n_items);
    bool quantized_ok = true;
    BatchGroups groups{};
    int max_stack[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
    long long chunks[4] = {0, 0, 0, 0};
    for (int g = 0; g < 4; g++) {
        groups.first[g] = (int)table.size();
        for (int i = 0; i < n_items; i++) {
            const ImbaTraceItem& it = items[i];
            if (!it.scene || it.count < 0 || (it.count > 0 && (!it.rays || !it.hits)))
                return fail(IMBA_ERR_INVALID_ARG, "queue %d: scene NULL, count < 0 or NULL rays/hits", i);
            if (it.scene->device != b->device)
                return fail(IMBA_ERR_INVALID_ARG, "queue %d: scene lives on device %d, not the batch's device %d", i,
                            it.scene->device, b->device);
            const int masked = !it.scene->info.all_opaque;
            if (2 * (it.any_hit != 0) + masked != g) continue;
            BatchItem e{};
            e.a = scene_args(it.scene->dev, it.scene->d_error);
            e.a.rays = (const float4*)it.rays;
            e.a.hits = (int4*)it.hits;
            e.a.count = it.count;
            e.count_dev = it.count_dev;
            e.capacity = it.count;
            e.group = g;
            max_stack[0][g] = std::max(max_stack[0][g], it.scene->info.max_stack);
            max_stack[1][g] = std::max(max_stack[1][g], it.scene->info4.max_stack);
            chunks[g] += ceil_div(it.count, 64);
            table.push_back(e);
            const TraceArgs a4 = scene_args(it.scene->dev4, it.scene->d_error);
            e.a.nodes = a4.nodes;
            e.a.n_nodes = a4.n_nodes;
            e.a.inst = a4.inst;
            e.a.n_inst = a4.n_inst;
            e.a.root = a4.root;
            table4.push_back(e);
            if (it.scene->dev4q.nodes) e.a.nodes = (const float4*)it.scene->dev4q.nodes;
            else quantized_ok = false;
            e.a.tl_base = it.scene->info4.treelet_base;
            e.a.tl_nodes = it.scene->info4.treelet_nodes;
            table4q.push_back(e);
        }
    }
    groups.first[4] = (int)table.size();
    const int n = (int)table.size();
    if (b->captured && std::memcmp(&groups, &b->captured_groups, sizeof(groups)) != 0)
        return fail(IMBA_ERR_INVALID_ARG, "batch was launched into a graph: the number of queues per kind "
                    "(closest / any hit, masked / opaque) must stay as captured");
    const size_t bytes = sizeof(BatchItem) * (size_t)n;
    if (n > b->cap) {
        // grow: device tables and staging buffers (synchronous).  A graph replays its launches
        // with the table addresses they were recorded with, so a captured batch cannot grow.
        if (b->captured)
            return fail(IMBA_ERR_CAPACITY, "batch was launched into a graph with room for %d queues; %d do not fit "
                        "(create the batch with its largest queue set before capturing)", b->cap, n);
        IMBA_HIP_TRY(hipDeviceSynchronize());
        for (int k = 0; k < 3; k++) {
            if (b->d_items[k]) {
                int r = retire_buffer(b->d_items[k]);
                if (r) return r;
            }
            b->d_items[k] = nullptr;
        }
        for (int k = 0; k < 2; k++) {
            if (b->staging[k]) IMBA_HIP_TRY(hipHostFree(b->staging[k]));
            b->staging[k] = nullptr;
        }
        const int cap = std::max(n, 2 * b->cap);
        for (int k = 0; k < 3; k++) IMBA_HIP_TRY(hipMalloc((void**)&b->d_items[k], sizeof(BatchItem) * (size_t)cap));
        for (int k = 0; k < 2; k++) {
            IMBA_HIP_TRY(hipHostMalloc((void**)&b->staging[k], 3 * sizeof(BatchItem) * (size_t)cap));
            if (!b->staged[k]) IMBA_HIP_TRY(hipEventCreateWithFlags(&b->staged[k], hipEventDisableTiming));
        }
        b->cap = cap;
    }
    if (n > 0) {
        if (stream) {
            const int k = b->next_staging;
            b->next_staging ^= 1;
            IMBA_HIP_TRY(hipEventSynchronize(b->staged[k]));
            BatchItem* st = b->staging[k];
            std::memcpy(st, table.data(), bytes);
            std::memcpy(st + n, table4.data(), bytes);
            std::memcpy(st + 2 * n, table4q.data(), bytes);
            IMBA_HIP_TRY(hipMemcpyAsync(b->d_items[0], st, bytes, hipMemcpyHostToDevice, stream));
            IMBA_HIP_TRY(hipMemcpyAsync(b->d_items[1], st + n, bytes, hipMemcpyHostToDevice, stream));
            IMBA_HIP_TRY(hipMemcpyAsync(b->d_items[2], st + 2 * n, bytes, hipMemcpyHostToDevice, stream));
            IMBA_HIP_TRY(hipEventRecord(b->staged[k], stream));
        } else {
            IMBA_HIP_TRY(hipMemcpy(b->d_items[0], table.data(), bytes, hipMemcpyHostToDevice));
            IMBA_HIP_TRY(hipMemcpy(b->d_items[1], table4.data(), bytes, hipMemcpyHostToDevice));
            IMBA_HIP_TRY(hipMemcpy(b->d_items[2], table4q.data(), bytes, hipMemcpyHostToDevice));
        }
    }
    b->groups = groups;
    b->n = n;
    b->quantized_ok = quantized_ok;
    std::memcpy(b->max_stack, max_stack, sizeof(max_stack));
    std::memcpy(b->capacity_chunks, chunks, sizeof(chunks));
    return IMBA_OK;
}

}  // namespace
}  // namespace imba

using namespace imba;

extern "C" {

int imba_trace_batch_create(const ImbaTraceItem* items, int n_items, ImbaTraceBatch** out) {
    IMBA_CHECK_ARG(out && (items || n_items == 0), "NULL argument");
    IMBA_CHECK_ARG(n_items >= 0, "n_items < 0");
    *out = nullptr;
    auto* b = new ImbaTraceBatch();
    hipError_t e = hipGetDevice(&b->device);
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_group_chunks, sizeof(int) * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_counters, sizeof(int) * 8);
    if (e == hipSuccess) e = hipMemset(b->d_counters, 0, sizeof(int) * 8);
    if (e == hipSuccess) e = hipMemset(b->d_group_chunks, 0, sizeof(int) * 4);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        free_batch(b);
        return fail(e == hipErrorOutOfMemory ? IMBA_ERR_OUT_OF_MEMORY : IMBA_ERR_HIP, "batch setup failed: %s",
                    hipGetErrorString(e));
    }
    int r = set_tables(b, items, n_items, nullptr);
    if (r) {
        free_batch(b);
        return r;
    }
    *out = b;
    return IMBA_OK;
}

int imba_trace_batch_update(ImbaTraceBatch* b, const ImbaTraceItem* items, int n_items, void* stream) {
    IMBA_CHECK_ARG(b && (items || n_items == 0), "NULL argument");
    IMBA_CHECK_ARG(n_items >= 0, "n_items < 0");
    int dev = 0;
    IMBA_HIP_TRY(hipGetDevice(&dev));
    if (dev != b->device) return fail(IMBA_ERR_INVALID_ARG, "batch belongs to device %d, current is %d", b->device, dev);
    bool sync;
    hipStream_t s = resolve_stream(stream, &sync);
    int r = set_tables(b, items, n_items, s);
    if (r) return r;
    if (sync) IMBA_HIP_TRY(hipStreamSynchronize(s));
    return IMBA_OK;
}

int imba_trace_batch_launch(ImbaTraceBatch* b, const ImbaTraceConfig* cfg, void* stream) {
    IMBA_CHECK_ARG(b, "batch is NULL");
    int dev = 0;
    IMBA_HIP_TRY(hipGetDevice(&dev));
    if (dev != b->device) return fail(IMBA_ERR_INVALID_ARG, "batch belongs to device %d, current is %d", b->device, dev);
    if (b->n == 0) return IMBA_OK;
    const int tree = tree_code(cfg);
    if (tree < 0) return fail(IMBA_ERR_INVALID_ARG, "bvh_width must be 0, 2 or 4 (quantized only with 4)");
    if (tree == 2 && !b->quantized_ok) return fail(IMBA_ERR_INVALID_ARG, "a queue's scene has no quantized tree");
    if (cfg && cfg->treelet) {
        if (cfg->treelet < 0 || cfg->treelet > kTreeletMaxNodes)
            return fail(IMBA_ERR_INVALID_ARG, "treelet must be 0..%d nodes", kTreeletMaxNodes);
        if (tree != 2) return fail(IMBA_ERR_INVALID_ARG, "treelet needs the quantized 4-wide tree");
    }
    bool sync;
    hipStream_t s = resolve_stream(stream, &sync);
    if (is_capturing(s)) {
        b->captured = true;
        b->captured_groups = b->groups;
    }
    hipLaunchKernelGGL(k_batch_prefix, dim3(1), dim3(1024), 0, s, b->d_items[tree], b->groups, b->d_group_chunks);
    IMBA_HIP_TRY(hipGetLastError());
    for (int g = 0; g < 4; g++) {
        if (b->groups.first[g + 1] == b->groups.first[g] || b->capacity_chunks[g] == 0) continue;
        int r;
        switch (g + 4 * tree) {
            case 0: r = dispatch_batch<false, false, 0>(b, g, cfg, s); break;
            case 1: r = dispatch_batch<false, true, 0>(b, g, cfg, s); break;
            case 2: r = dispatch_batch<true, false, 0>(b, g, cfg, s); break;
            case 3: r = dispatch_batch<true, true, 0>(b, g, cfg, s); break;
            case 4: r = dispatch_batch<false, false, 1>(b, g, cfg, s); break;
            case 5: r = dispatch_batch<false, true, 1>(b, g, cfg, s); break;
            case 6: r = dispatch_batch<true, false, 1>(b, g, cfg, s); break;
            case 7: r = dispatch_batch<true, true, 1>(b, g, cfg, s); break;
            case 8: r = dispatch_batch<false, false, 2>(b, g, cfg, s); break;
            case 9: r = dispatch_batch<false, true, 2>(b, g, cfg, s); break;
            case 10: r = dispatch_batch<true, false, 2>(b, g, cfg, s); break;
            default: r = dispatch_batch<true, true, 2>(b, g, cfg, s); break;
        }
        if (r) return r;
    }
    if (sync) IMBA_HIP_TRY(hipStreamSynchronize(s));
    return IMBA_OK;
}

int imba_autotune(const ImbaScene* sc, const ImbaRay* rays, ImbaHit* hits, int count, int any_hit, int reps,
                  ImbaTraceConfig* best, float* best_ms, void* stream) {
    IMBA_CHECK_ARG(sc && rays && hits && best, "NULL argument");
    IMBA_CHECK_ARG(count > 0 && reps > 0, "count and reps must be positive");
    static const int kShapes[][2] = {{256, 16}, {256, 8}, {256, 32}, {128, 16}, {128, 32}, {64, 32}, {64, 64}};
    static const int kRefill[] = {0, 32};
    // bvh_width, quantized, treelet nodes staged in LDS
    static const int kTree[][3] = {{2, 0, 0}, {4, 0, 0}, {4, 1, 0}, {4, 1, 64}, {4, 1, 256}};
    bool sync;
    hipStream_t s = resolve_stream(stream, &sync);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    IMBA_HIP_TRY(hipEventCreate(&e0));
    IMBA_HIP_TRY(hipEventCreate(&e1));
    float best_t = -1.0f;
    int r = IMBA_OK;
    for (const auto& tree : kTree)
    for (const auto& shape : kShapes) {
        for (int refill : kRefill) {
            ImbaTraceConfig cfg{shape[0], shape[1], refill, 0, tree[0], tree[1], tree[2]};
            if (tree[1] && !sc->dev4q.nodes) continue;
            if (tree[2] && sc->info4.treelet_nodes <= 0) continue;
            TreeView t{};
            r = tree_of(sc, &cfg, &t);
            if (!r)
                r = launch_trace(t, sc->d_error, rays, hits, nullptr, count, any_hit != 0, !sc->info.all_opaque, &cfg,
                                 s);
            if (r) break;
            hipEventRecord(e0, s);
            for (int k = 0; k < reps && !r; k++)
                r = launch_trace(t, sc->d_error, rays, hits, nullptr, count, any_hit != 0, !sc->info.all_opaque, &cfg,
                                 s);
            if (r) break;
            hipEventRecord(e1, s);
            float ms = 0.0f;
            if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                r = fail(IMBA_ERR_HIP, "event timing failed");
                break;
            }
            ms /= (float)reps;
            if (best_t < 0.0f || ms < best_t) {
                best_t = ms;
                *best = cfg;
            }
        }
        if (r) break;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (r) return r;
    if (best_ms) *best_ms = best_t;
    return IMBA_OK;
}

int imba_trace_batch_release(ImbaTraceBatch* b) {
    if (b) {
        hipDeviceSynchronize();
        free_batch(b);
    }
    return IMBA_OK;
}

static int trace_entry(const ImbaScene* s, const ImbaRay* rays, ImbaHit* hits, uint32_t* counts, int count, bool any,
                       const ImbaTraceConfig* cfg, void* stream, const int* count_dev,
                       uint32_t* depth_hist = nullptr) {
    IMBA_CHECK_ARG(s, "scene is NULL");
    TreeView t{};
    int r = tree_of(s, cfg, &t);
    if (r) return r;
    return launch_trace(t, s->d_error, rays, hits, counts, count, any, !s->info.all_opaque, cfg, stream, count_dev,
                        depth_hist);
}

int imba_intersect(const ImbaScene* s, const ImbaRay* rays, ImbaHit* hits, int count, const ImbaTraceConfig* cfg,
                   void* stream) {
    return trace_entry(s, rays, hits, nullptr, count, false, cfg, stream, nullptr);
}

int imba_occluded(const ImbaScene* s, const ImbaRay* rays, ImbaHit* hits, int count, const ImbaTraceConfig* cfg,
                  void* stream) {
    return trace_entry(s, rays, hits, nullptr, count, true, cfg, stream, nullptr);
}

int imba_intersect_dev(const ImbaScene* s, const ImbaRay* rays, ImbaHit* hits, int capacity, const int* count_dev,
                       const ImbaTraceConfig* cfg, void* stream) {
    IMBA_CHECK_ARG(count_dev, "count_dev is NULL");
    return trace_entry(s, rays, hits, nullptr, capacity, false, cfg, stream, count_dev);
}

int imba_occluded_dev(const ImbaScene* s, const ImbaRay* rays, ImbaHit* hits, int capacity, const int* count_dev,
                      const ImbaTraceConfig* cfg, void* stream) {
    IMBA_CHECK_ARG(count_dev, "count_dev is NULL");
    return trace_entry(s, rays, hits, nullptr, capacity, true, cfg, stream, count_dev);
}

int imba_intersect_stats(const ImbaScene* s, const ImbaRay* rays, ImbaHit* hits, uint32_t* counts, int count,
                         void* stream) {
    IMBA_CHECK_ARG(counts, "counts is NULL");
    const ImbaTraceConfig bvh2{0, 0, 0, 0, 2, 0, 0};
    return trace_entry(s, rays, hits, counts, count, false, &bvh2, stream, nullptr);
}

int imba_occluded_stats(const ImbaScene* s, const ImbaRay* rays, ImbaHit* hits, uint32_t* counts, int count,
                        void* stream) {
    IMBA_CHECK_ARG(counts, "counts is NULL");
    const ImbaTraceConfig bvh2{0, 0, 0, 0, 2, 0, 0};
    return trace_entry(s, rays, hits, counts, count, true, &bvh2, stream, nullptr);
}

int imba_trace_stats(const ImbaScene* s, const ImbaRay* rays, ImbaHit* hits, uint32_t* counts, int count,
                     int any_hit, const ImbaTraceConfig* cfg, void* stream) {
    IMBA_CHECK_ARG(counts, "counts is NULL");
    return trace_entry(s, rays, hits, counts, count, any_hit != 0, cfg, stream, nullptr);
}

int imba_trace_stack_depths(const ImbaScene* s, const ImbaRay* rays, ImbaHit* hits, int count, int any_hit,
                            const ImbaTraceConfig* cfg, uint32_t* depth_hist, void* stream) {
    IMBA_CHECK_ARG(depth_hist, "depth_hist is NULL");
    return trace_entry(s, rays, hits, nullptr, count, any_hit != 0, cfg, stream, nullptr, depth_hist);
}

int imba_trace_desc(const ImbaSceneDesc* dev, const ImbaRay* rays, ImbaHit* hits, int count, int any_hit,
                    void* stream) {
    IMBA_CHECK_ARG(dev, "desc is NULL");
    return ffi_launch(*dev, rays, hits, count, any_hit != 0, stream);
}

int imba_ffi_error_count(int* count) {
    IMBA_CHECK_ARG(count, "count is NULL");
    int* err = nullptr;
    int r = ffi_err_counter(&err);
    if (r) return r;
    IMBA_HIP_TRY(hipMemcpy(count, err, sizeof(int), hipMemcpyDeviceToHost));
    return IMBA_OK;
}

// Reference FFI (render/scheduling/ray_queue.h:297-308, 347-358): synchronous, device pointers.
void intersect_gpu_masked_instanced(int root, void* nodes, void* instances, void* tris, void* rays, void* hits,
                                    int* indices, void* texcoords, void* masks, char* mask_buffer, int count) {
    ffi_trace(false, root, nodes, instances, tris, rays, hits, indices, texcoords, masks, mask_buffer, count);
}

void occluded_gpu_masked_instanced(int root, void* nodes, void* instances, void* tris, void* rays, void* hits,
                                   int* indices, void* texcoords, void* masks, char* mask_buffer, int count) {
    ffi_trace(true, root, nodes, instances, tris, rays, hits, indices, texcoords, masks, mask_buffer, count);
}

}  // extern "C"
