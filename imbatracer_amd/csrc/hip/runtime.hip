// Device runtime: memory, streams, events, device properties and resident scenes.
//
// Replaces the AnyDSL runtime the reference uses for its GPU buffers (anydsl::Array on
// Platform::Cuda, anydsl::copy: render/scheduling/ray_queue.h:77-80,277,292,327,342) and the
// scene upload (render/scene.cpp:119-188).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "common.h"
#include "scene.h"
#include "../host/wide_bvh.h"

namespace imba {

static std::mutex g_props_mutex;
static DeviceProps g_props[64];

const DeviceProps* device_props() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    std::lock_guard<std::mutex> lock(g_props_mutex);
    DeviceProps& p = g_props[dev];
    if (p.device != dev) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return nullptr;
        p.cu_count = prop.multiProcessorCount;
        p.lds_per_block = (int)prop.sharedMemPerBlock;
        p.l2_bytes = prop.l2CacheSize;
        std::strncpy(p.name, prop.gcnArchName, sizeof(p.name) - 1);
        p.device = dev;
    }
    return &p;
}

namespace {
struct ThreadStreams {
    hipStream_t s[64] = {nullptr};
    ~ThreadStreams() {
        // Streams are left to the runtime's teardown: destroying them from a thread-exit
        // handler can run after the runtime has shut down.
    }
};
thread_local ThreadStreams t_streams;
}  // namespace

hipStream_t thread_stream() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    if (!t_streams.s[dev]) {
        hipStream_t s = nullptr;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return nullptr;
        t_streams.s[dev] = s;
    }
    return t_streams.s[dev];
}

// Device buffers replaced while graphs exist are kept until the last graph is released: a
// captured launch keeps the address it was recorded with, so the buffer it names must outlive
// it.  Without graphs they are freed at once.
namespace {
std::mutex g_retired_mutex;
std::vector<void*> g_retired;
int g_live_graphs = 0;
}  // namespace

int retire_buffer(void* p) {
    if (!p) return IMBA_OK;
    std::lock_guard<std::mutex> lock(g_retired_mutex);
    if (g_live_graphs > 0) {
        g_retired.push_back(p);
        return IMBA_OK;
    }
    IMBA_HIP_TRY(hipFree(p));
    return IMBA_OK;
}

int grow_buffer(Scratch& s, size_t bytes) {
    if (s.bytes >= bytes && s.ptr) return IMBA_OK;
    if (s.ptr) {
        IMBA_HIP_TRY(hipDeviceSynchronize());
        int r = retire_buffer(s.ptr);
        if (r) return r;
        s.ptr = nullptr;
        s.bytes = 0;
    }
    const size_t rounded = std::max<size_t>(bytes, 4096);
    IMBA_HIP_TRY(hipMalloc(&s.ptr, rounded));
    s.bytes = rounded;
    return IMBA_OK;
}

// ---------------------------------------------------------------------------------------
// The 4-wide tree of a scene (derived on the host, uploaded next to the BVH2).
int upload_wide(const ImbaSceneDesc& host, const ImbaSceneDesc& dev, const SceneInfo& info, ImbaSceneDesc* dev4,
                ImbaSceneDesc* dev4q, SceneInfo* info4, size_t* bytes) {
    WideScene w;
    int r = collapse_scene4(host, w);
    if (r) return r;
    *dev4 = dev;
    dev4->nodes = nullptr;
    dev4->instances = nullptr;
    *dev4q = *dev4;
    void* pn = nullptr;
    void* pi = nullptr;
    void* pq = nullptr;
    const size_t nb = sizeof(ImbaNode4) * w.nodes.size(), ib = sizeof(ImbaInstanceNode) * w.instances.size();
    const size_t qb = sizeof(QNode4) * w.qnodes.size();
    auto cleanup = [&]() {
        if (pn) hipFree(pn);
        if (pi) hipFree(pi);
        if (pq) hipFree(pq);
    };
    if (hipMalloc(&pn, nb) != hipSuccess || hipMalloc(&pi, ib) != hipSuccess || (qb && hipMalloc(&pq, qb) != hipSuccess)) {
        cleanup();
        return fail(IMBA_ERR_OUT_OF_MEMORY, "4-wide tree");
    }
    if (hipMemcpy(pn, w.nodes.data(), nb, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(pi, w.instances.data(), ib, hipMemcpyHostToDevice) != hipSuccess ||
        (qb && hipMemcpy(pq, w.qnodes.data(), qb, hipMemcpyHostToDevice) != hipSuccess)) {
        cleanup();
        return fail(IMBA_ERR_HIP, "upload of the 4-wide tree failed");
    }
    dev4->nodes = (const ImbaNode*)pn;
    dev4->node_count = (int)w.nodes.size();
    dev4->instances = (const ImbaInstanceNode*)pi;
    dev4->instance_count = (int)w.instances.size();
    dev4->root = w.root;
    *dev4q = *dev4;
    dev4q->nodes = (const ImbaNode*)pq;  // NULL when the boxes could not be quantized
    *info4 = info;
    info4->max_stack = w.max_stack;
    info4->treelet_base = w.treelet_base;
    info4->treelet_nodes = w.treelet_nodes;
    *bytes = nb + ib + qb;
    return IMBA_OK;
}

void free_wide(ImbaSceneDesc& dev4, ImbaSceneDesc& dev4q) {
    if (dev4.nodes) (void)hipFree(const_cast<ImbaNode*>(dev4.nodes));
    if (dev4.instances) (void)hipFree(const_cast<ImbaInstanceNode*>(dev4.instances));
    if (dev4q.nodes) (void)hipFree(const_cast<ImbaNode*>(dev4q.nodes));
    dev4.nodes = nullptr;
    dev4.instances = nullptr;
    dev4q.nodes = nullptr;
    dev4q.instances = nullptr;
}

// ---------------------------------------------------------------------------------------
// Scene validation: walks the top level from the root, then every instanced mesh BVH.
// Checks indices, the parent < child order of the depth-first layout, leaf sentinels, and
// measures the deepest traversal stack a ray can need.
int validate_scene(const ImbaSceneDesc& d, SceneInfo& info) {
    info = SceneInfo{};
    const int nn = d.node_count, ni = d.instance_count, nt = d.tri_vec4_count;
    if (nn <= 0 || !d.nodes) return fail(IMBA_ERR_BAD_BVH, "scene has no nodes");
    if (d.root < 0 || d.root >= nn) return fail(IMBA_ERR_BAD_BVH, "root %d out of range [0,%d)", d.root, nn);
    if (ni <= 0 || !d.instances) return fail(IMBA_ERR_BAD_BVH, "scene has no instances");
    if (nt <= 0 || !d.tris) return fail(IMBA_ERR_BAD_BVH, "scene has no triangles");

    struct Item { int node, depth; };
    std::vector<Item> work;
    std::vector<int> mesh_roots;
    int top_depth = 0;

    // Checks one child link of node `parent`; leaves are checked by the caller.
    auto check_child = [&](int parent, int c, int slot) -> int {
        if (c >= 0) {
            if (slot == 1 && c == IMBA_DUMMY_CHILD) return IMBA_OK;
            if (c <= parent || c >= nn)
                return fail(IMBA_ERR_BAD_BVH, "node %d: child %d breaks the depth-first order or is out of range", parent, c);
        }
        return IMBA_OK;
    };

    // Top level.
    std::vector<uint8_t> inst_seen(ni, 0);
    work.push_back({d.root, 1});
    while (!work.empty()) {
        const Item it = work.back();
        work.pop_back();
        top_depth = std::max(top_depth, it.depth);
        const ImbaNode& n = d.nodes[it.node];
        const int ch[2] = {n.left, n.right};
        for (int s = 0; s < 2; s++) {
            int r = check_child(it.node, ch[s], s);
            if (r) return r;
            if (ch[s] >= 0) {
                if (!(s == 1 && ch[s] == IMBA_DUMMY_CHILD)) work.push_back({ch[s], it.depth + 1});
                continue;
            }
            // Instance list.
            int i = ~ch[s];
            for (;; i++) {
                if (i < 0 || i >= ni) return fail(IMBA_ERR_BAD_BVH, "instance list at %d runs past %d instances", ~ch[s], ni);
                const ImbaInstanceNode& in = d.instances[i];
                if (in.next < 0 || in.next >= nn) return fail(IMBA_ERR_BAD_BVH, "instance %d: root %d out of range", i, in.next);
                if (!inst_seen[i]) { inst_seen[i] = 1; mesh_roots.push_back(in.next); }
                if (in.pad[0] == -1) break;
            }
        }
    }

    // Mesh levels.
    std::sort(mesh_roots.begin(), mesh_roots.end());
    mesh_roots.erase(std::unique(mesh_roots.begin(), mesh_roots.end()), mesh_roots.end());
    int mesh_depth = 0;
    for (int root : mesh_roots) {
        work.push_back({root, 1});
        while (!work.empty()) {
            const Item it = work.back();
            work.pop_back();
            mesh_depth = std::max(mesh_depth, it.depth);
            const ImbaNode& n = d.nodes[it.node];
            const int ch[2] = {n.left, n.right};
            for (int s = 0; s < 2; s++) {
                int r = check_child(it.node, ch[s], s);
                if (r) return r;
                if (ch[s] >= 0) {
                    if (!(s == 1 && ch[s] == IMBA_DUMMY_CHILD)) work.push_back({ch[s], it.depth + 1});
                    continue;
                }
                int t = ~ch[s];
                for (;; t += 3) {
                    if (t < 0 || t + 2 >= nt) return fail(IMBA_ERR_BAD_BVH, "triangle list at %d runs past %d records", ~ch[s], nt);
                    uint32_t bits;
                    std::memcpy(&bits, &d.tris[t + 2].w, 4);
                    if (bits == IMBA_LEAF_END_BITS) break;
                }
            }
        }
    }
    // A ray's stack holds at most one entry per level of each tree it is inside.
    info.max_stack = top_depth + mesh_depth + 2;

    // Masks: all-opaque scenes skip the texel test.
    info.all_opaque = true;
    if (d.mask_count > 0 && d.masks && d.mask_buffer) {
        for (int m = 0; m < d.mask_count; m++) {
            const ImbaTransparencyMask& mk = d.masks[m];
            if (mk.width <= 0 || mk.height <= 0 || mk.offset < 0 ||
                (int64_t)mk.offset + (int64_t)mk.width * mk.height > d.mask_buffer_size)
                return fail(IMBA_ERR_BAD_BVH, "mask %d out of the mask buffer", m);
            for (int64_t k = 0; k < (int64_t)mk.width * mk.height; k++)
                if (d.mask_buffer[mk.offset + k] == 0) { info.all_opaque = false; break; }
        }
    }
    return IMBA_OK;
}

}  // namespace imba

using namespace imba;

struct ImbaGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" {

int imba_device_count(int* count) {
    IMBA_CHECK_ARG(count, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return IMBA_OK;
}

int imba_set_device(int device) {
    IMBA_HIP_TRY(hipSetDevice(device));
    return IMBA_OK;
}

int imba_get_device(int* device) {
    IMBA_CHECK_ARG(device, "device is NULL");
    IMBA_HIP_TRY(hipGetDevice(device));
    return IMBA_OK;
}

int imba_device_info(int* cu_count, int* lds_bytes_per_block, int* l2_bytes, char* name, int name_len) {
    const DeviceProps* p = device_props();
    if (!p) return fail(IMBA_ERR_NO_DEVICE, "no HIP device");
    if (cu_count) *cu_count = p->cu_count;
    if (lds_bytes_per_block) *lds_bytes_per_block = p->lds_per_block;
    if (l2_bytes) *l2_bytes = p->l2_bytes;
    if (name && name_len > 0) {
        std::strncpy(name, p->name, name_len - 1);
        name[name_len - 1] = 0;
    }
    return IMBA_OK;
}

int imba_malloc(void** ptr, size_t bytes) {
    IMBA_CHECK_ARG(ptr, "ptr is NULL");
    *ptr = nullptr;
    if (bytes == 0) return IMBA_OK;
    IMBA_HIP_TRY(hipMalloc(ptr, bytes));
    return IMBA_OK;
}

int imba_free(void* ptr) {
    if (ptr) IMBA_HIP_TRY(hipFree(ptr));
    return IMBA_OK;
}

int imba_host_alloc(void** ptr, size_t bytes) {
    IMBA_CHECK_ARG(ptr, "ptr is NULL");
    IMBA_HIP_TRY(hipHostMalloc(ptr, std::max<size_t>(bytes, 1), hipHostMallocDefault));
    return IMBA_OK;
}

int imba_host_free(void* ptr) {
    if (ptr) IMBA_HIP_TRY(hipHostFree(ptr));
    return IMBA_OK;
}

static int copy_impl(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, void* stream) {
    if (bytes == 0) return IMBA_OK;
    IMBA_CHECK_ARG(dst && src, "NULL pointer");
    bool sync;
    hipStream_t s = resolve_stream(stream, &sync);
    IMBA_HIP_TRY(hipMemcpyAsync(dst, src, bytes, kind, s));
    if (sync) IMBA_HIP_TRY(hipStreamSynchronize(s));
    return IMBA_OK;
}

int imba_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
    return copy_impl(dst, src, bytes, hipMemcpyHostToDevice, stream);
}
int imba_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream) {
    return copy_impl(dst, src, bytes, hipMemcpyDeviceToHost, stream);
}
int imba_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream) {
    return copy_impl(dst, src, bytes, hipMemcpyDeviceToDevice, stream);
}

int imba_memset(void* dst, int value, size_t bytes, void* stream) {
    if (bytes == 0) return IMBA_OK;
    IMBA_CHECK_ARG(dst, "NULL pointer");
    bool sync;
    hipStream_t s = resolve_stream(stream, &sync);
    IMBA_HIP_TRY(hipMemsetAsync(dst, value, bytes, s));
    if (sync) IMBA_HIP_TRY(hipStreamSynchronize(s));
    return IMBA_OK;
}

int imba_stream_create(void** stream) {
    IMBA_CHECK_ARG(stream, "stream is NULL");
    hipStream_t s;
    IMBA_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return IMBA_OK;
}
int imba_stream_destroy(void* stream) {
    if (stream) IMBA_HIP_TRY(hipStreamDestroy((hipStream_t)stream));
    return IMBA_OK;
}
int imba_stream_sync(void* stream) {
    hipStream_t s = stream ? (hipStream_t)stream : thread_stream();
    IMBA_HIP_TRY(hipStreamSynchronize(s));
    return IMBA_OK;
}
int imba_event_create(void** event) {
    IMBA_CHECK_ARG(event, "event is NULL");
    hipEvent_t e;
    IMBA_HIP_TRY(hipEventCreate(&e));
    *event = e;
    return IMBA_OK;
}
int imba_event_destroy(void* event) {
    if (event) IMBA_HIP_TRY(hipEventDestroy((hipEvent_t)event));
    return IMBA_OK;
}
int imba_event_record(void* event, void* stream) {
    IMBA_CHECK_ARG(event, "event is NULL");
    hipStream_t s = stream ? (hipStream_t)stream : thread_stream();
    IMBA_HIP_TRY(hipEventRecord((hipEvent_t)event, s));
    return IMBA_OK;
}
int imba_event_sync(void* event) {
    IMBA_CHECK_ARG(event, "event is NULL");
    IMBA_HIP_TRY(hipEventSynchronize((hipEvent_t)event));
    return IMBA_OK;
}

int imba_stream_wait_event(void* stream, void* event) {
    IMBA_CHECK_ARG(stream && event, "NULL argument");
    IMBA_HIP_TRY(hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)event, 0));
    return IMBA_OK;
}

int imba_event_elapsed_ms(float* ms, void* start, void* end) {
    IMBA_CHECK_ARG(ms && start && end, "NULL argument");
    IMBA_HIP_TRY(hipEventSynchronize((hipEvent_t)end));
    IMBA_HIP_TRY(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)end));
    return IMBA_OK;
}

// ---------------------------------------------------------------------------------------
int imba_scene_validate(const ImbaSceneDesc* h, int* max_stack, int* all_opaque) {
    IMBA_CHECK_ARG(h, "desc is NULL");
    SceneInfo info;
    int r = validate_scene(*h, info);
    if (r) return r;
    if (max_stack) *max_stack = info.max_stack;
    if (all_opaque) *all_opaque = info.all_opaque ? 1 : 0;
    return IMBA_OK;
}

int imba_scene_upload(const ImbaSceneDesc* h, ImbaScene** out) {
    IMBA_CHECK_ARG(h && out, "NULL argument");
    *out = nullptr;
    SceneInfo info;
    int r = validate_scene(*h, info);
    if (r) return r;
    ImbaScene* s = new ImbaScene();
    s->info = info;
    s->counts = *h;
    if (hipGetDevice(&s->device) != hipSuccess) s->device = 0;
    auto up = [&](const void* src, size_t bytes, void** dst) -> int {
        *dst = nullptr;
        if (!src || bytes == 0) return IMBA_OK;
        IMBA_HIP_TRY(hipMalloc(dst, bytes));
        s->footprint += bytes;
        IMBA_HIP_TRY(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
        return IMBA_OK;
    };
    ImbaSceneDesc& dd = s->dev;
    dd = *h;
    void* p;
    if ((r = up(h->nodes, sizeof(ImbaNode) * (size_t)h->node_count, &p))) goto err;
    dd.nodes = (const ImbaNode*)p;
    if ((r = up(h->instances, sizeof(ImbaInstanceNode) * (size_t)h->instance_count, &p))) goto err;
    dd.instances = (const ImbaInstanceNode*)p;
    if ((r = up(h->tris, sizeof(ImbaVec4) * (size_t)h->tri_vec4_count, &p))) goto err;
    dd.tris = (const ImbaVec4*)p;
    if ((r = up(h->indices, sizeof(int) * (size_t)h->index_count, &p))) goto err;
    dd.indices = (const int*)p;
    if ((r = up(h->texcoords, sizeof(ImbaVec2) * (size_t)h->texcoord_count, &p))) goto err;
    dd.texcoords = (const ImbaVec2*)p;
    if ((r = up(h->masks, sizeof(ImbaTransparencyMask) * (size_t)h->mask_count, &p))) goto err;
    dd.masks = (const ImbaTransparencyMask*)p;
    if ((r = up(h->mask_buffer, (size_t)h->mask_buffer_size, &p))) goto err;
    dd.mask_buffer = (const uint8_t*)p;
    {
        size_t wb = 0;
        if ((r = upload_wide(*h, dd, info, &s->dev4, &s->dev4q, &s->info4, &wb))) goto err;
        s->footprint += wb;
    }
    if (hipMalloc((void**)&s->d_error, sizeof(int)) != hipSuccess) { r = fail(IMBA_ERR_OUT_OF_MEMORY, "error counter"); goto err; }
    if (hipMemset(s->d_error, 0, sizeof(int)) != hipSuccess) { r = fail(IMBA_ERR_HIP, "error counter"); goto err; }
    s->footprint += sizeof(int);
    *out = s;
    return IMBA_OK;
err:
    imba_scene_release(s);
    return r;
}

int imba_scene_release(ImbaScene* s) {
    if (!s) return IMBA_OK;
    const void* owned[] = {s->dev.nodes, s->dev.instances, s->dev.tris, s->dev.indices, s->dev.texcoords,
                           s->dev.masks, s->dev.mask_buffer, s->d_error};
    for (const void* p : owned)
        if (p) (void)hipFree(const_cast<void*>(p));
    free_wide(s->dev4, s->dev4q);
    delete s;
    return IMBA_OK;
}

int imba_scene_device_desc(const ImbaScene* s, ImbaSceneDesc* out) {
    IMBA_CHECK_ARG(s && out, "NULL argument");
    *out = s->dev;
    return IMBA_OK;
}

int imba_scene_footprint(const ImbaScene* s, size_t* bytes) {
    IMBA_CHECK_ARG(s && bytes, "NULL argument");
    *bytes = s->footprint;
    return IMBA_OK;
}

int imba_trace_error_count(const ImbaScene* s, int* count) {
    IMBA_CHECK_ARG(s && count, "NULL argument");
    IMBA_HIP_TRY(hipMemcpy(count, s->d_error, sizeof(int), hipMemcpyDeviceToHost));
    return IMBA_OK;
}

// ---------------------------------------------------------------------------------------
// Graphs: stream capture of the library's asynchronous calls.
int imba_graph_begin(void* stream) {
    IMBA_CHECK_ARG(stream, "capture needs a created stream, not NULL");
    IMBA_HIP_TRY(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
    return IMBA_OK;
}

int imba_graph_end(void* stream, ImbaGraph** out) {
    IMBA_CHECK_ARG(stream && out, "NULL argument");
    *out = nullptr;
    hipGraph_t g = nullptr;
    IMBA_HIP_TRY(hipStreamEndCapture((hipStream_t)stream, &g));
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        hipGraphDestroy(g);
        return fail(IMBA_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
    }
    auto* r = new ImbaGraph();
    r->graph = g;
    r->exec = exec;
    {
        std::lock_guard<std::mutex> lock(g_retired_mutex);
        g_live_graphs++;
    }
    *out = r;
    return IMBA_OK;
}

int imba_graph_launch(ImbaGraph* g, void* stream) {
    IMBA_CHECK_ARG(g, "graph is NULL");
    bool sync;
    hipStream_t s = resolve_stream(stream, &sync);
    IMBA_HIP_TRY(hipGraphLaunch(g->exec, s));
    if (sync) IMBA_HIP_TRY(hipStreamSynchronize(s));
    return IMBA_OK;
}

int imba_graph_release(ImbaGraph* g) {
    if (!g) return IMBA_OK;
    IMBA_HIP_TRY(hipDeviceSynchronize());
    hipGraphExecDestroy(g->exec);
    hipGraphDestroy(g->graph);
    delete g;
    std::vector<void*> to_free;
    {
        std::lock_guard<std::mutex> lock(g_retired_mutex);
        if (g_live_graphs > 0) g_live_graphs--;
        if (g_live_graphs == 0) to_free.swap(g_retired);
    }
    for (void* p : to_free) IMBA_HIP_TRY(hipFree(p));
    return IMBA_OK;
}

}  // extern "C"
