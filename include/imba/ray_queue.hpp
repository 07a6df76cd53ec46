// imba/ray_queue.hpp — the reference's RayQueue / TraversalData / HashGrid interfaces on top of
// the C ABI of libimba_hip.so (header only, C++11).
//
//   TraversalData<Node>  <- render/scheduling/ray_queue.h:42-54
//   RayQueue<StateType>  <- render/scheduling/ray_queue.h:57-360
//   HashGrid<Iter, Ph>   <- rangesearch/rangesearch.h:21-158
//
// Same member names, argument meaning and return values, so the integrators compile against
// it unchanged.  Differences, all internal:
//   - device buffers are imba_malloc'd (instead of anydsl::Array on Platform::Cuda);
//   - compact_hits / compact_rays run on the GPU (the rays and hits are already there after a
//     traversal; states are uploaded), keep the hits in order and put the misses after them
//     in order (the reference permutes the misses, which callers drop with shrink());
//   - HashGrid builds and queries on the GPU; query() is one GPU call per position (kept for
//     source compatibility; reentrant, as the reference's shading threads need), query_batch()
//     is the form to use from a bounce loop;
//   - traverse_cpu / traverse_occluded_cpu are declared but throw: the library has no CPU
//     traversal.
// Errors throw imba::Error (the reference asserts or does nothing).
#ifndef IMBA_RAY_QUEUE_HPP
#define IMBA_RAY_QUEUE_HPP

#include <algorithm>
#include <atomic>
#include <cassert>
#include <cstring>
#include <functional>
#include <iterator>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "imba_hip.h"

namespace imba {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& where) : std::runtime_error(where + ": " + imba_last_error()), code(c) {}

protected:
    struct Plain {};
    Error(int c, const std::string& msg, Plain) : std::runtime_error(msg), code(c) {}
};

// A member of the reference interface this library does not provide.
struct Unsupported : Error {
    explicit Unsupported(const std::string& msg) : Error(IMBA_ERR_INVALID_ARG, msg, Plain()) {}
};

inline void check(int code, const char* where) {
    if (code != IMBA_OK) throw Error(code, where);
}

// Device array (the role of anydsl::Array on the GPU platform).
template <typename T>
class DeviceArray {
public:
    DeviceArray() {}
    explicit DeviceArray(size_t n) { resize(n); }
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    DeviceArray(DeviceArray&& o) : ptr_(o.ptr_), n_(o.n_) { o.ptr_ = nullptr; o.n_ = 0; }
    DeviceArray& operator=(DeviceArray&& o) {
        std::swap(ptr_, o.ptr_);
        std::swap(n_, o.n_);
        return *this;
    }
    ~DeviceArray() { if (ptr_) imba_free(ptr_); }

    void resize(size_t n) {
        if (ptr_) imba_free(ptr_);
        ptr_ = nullptr;
        n_ = n;
        void* p = nullptr;
        check(imba_malloc(&p, std::max<size_t>(n, 1) * sizeof(T)), "imba_malloc");
        ptr_ = static_cast<T*>(p);
    }
    void upload(const T* src, size_t n, void* stream = nullptr) {
        assert(n <= n_);
        check(imba_memcpy_h2d(ptr_, src, n * sizeof(T), stream), "imba_memcpy_h2d");
    }
    void download(T* dst, size_t n, void* stream = nullptr) const {
        assert(n <= n_);
        check(imba_memcpy_d2h(dst, ptr_, n * sizeof(T), stream), "imba_memcpy_d2h");
    }
    T* data() { return ptr_; }
    const T* data() const { return ptr_; }
    size_t size() const { return n_; }

private:
    T* ptr_ = nullptr;
    size_t n_ = 0;
};

// Device-resident traversal data of a scene (TraversalData<traversal_gpu::Node>).
template <typename Node = ImbaNode>
struct TraversalData {
    int root = 0;
    DeviceArray<Node> nodes;
    DeviceArray<ImbaInstanceNode> instances;
    DeviceArray<ImbaVec4> tris;
    DeviceArray<ImbaVec2> texcoords;
    DeviceArray<int> indices;
    DeviceArray<ImbaTransparencyMask> masks;
    DeviceArray<char> mask_buffer;

    // Uploads host arrays (Scene::upload_mesh_accels / upload_top_level_accel / upload_mask_buffer).
    void upload(const ImbaSceneDesc& d) {
        static_assert(sizeof(Node) == sizeof(ImbaNode), "node layout");
        root = d.root;
        nodes.resize(d.node_count);
        nodes.upload(reinterpret_cast<const Node*>(d.nodes), d.node_count);
        instances.resize(d.instance_count);
        instances.upload(d.instances, d.instance_count);
        tris.resize(d.tri_vec4_count);
        tris.upload(d.tris, d.tri_vec4_count);
        texcoords.resize(d.texcoord_count);
        if (d.texcoord_count) texcoords.upload(d.texcoords, d.texcoord_count);
        indices.resize(d.index_count);
        if (d.index_count) indices.upload(d.indices, d.index_count);
        masks.resize(d.mask_count);
        if (d.mask_count) masks.upload(d.masks, d.mask_count);
        mask_buffer.resize(d.mask_buffer_size);
        if (d.mask_buffer_size) mask_buffer.upload(reinterpret_cast<const char*>(d.mask_buffer), d.mask_buffer_size);
    }
};

// A queue of rays with their states and hits (host arrays, mirrored on the device).
// StateType must start with `int pixel_id` (RayState, ray_queue.h:20-33).
template <typename StateType>
class RayQueue {
    static int align_cpu(int v) { return v % 8 == 0 ? v : v + 8 - v % 8; }
    static int align_gpu(int v) { return v % 64 == 0 ? v : v + 64 - v % 64; }
    static int align(int v) { return std::max(align_cpu(v), align_gpu(v)); }

public:
    RayQueue() : last_(-1), gpu_buffers_(false) {}
    RayQueue(int capacity, bool gpu_buffers)
        : ray_buffer_(align(capacity)), hit_buffer_(align(capacity)), state_buffer_(align(capacity)),
          sorted_indices_(align(capacity)), last_(-1), gpu_buffers_(gpu_buffers) {
        static_assert(sizeof(StateType) % 4 == 0, "state size must be a multiple of 4");
        std::memset(ray_buffer_.data(), 0, sizeof(ImbaRay) * ray_buffer_.size());
        if (gpu_buffers) {
            const int n = align_gpu(capacity);
            dev_rays_.resize(n);
            dev_hits_.resize(n);
        }
    }
    RayQueue(const RayQueue&) = delete;
    RayQueue& operator=(const RayQueue&) = delete;

    // ray_queue.h:84-102 (the device buffers move with the queue as well).
    RayQueue(RayQueue&& rhs)
        : ray_buffer_(std::move(rhs.ray_buffer_)), hit_buffer_(std::move(rhs.hit_buffer_)),
          state_buffer_(std::move(rhs.state_buffer_)), sorted_indices_(std::move(rhs.sorted_indices_)),
          last_(rhs.last_.load()), gpu_buffers_(rhs.gpu_buffers_), dev_rays_(std::move(rhs.dev_rays_)),
          dev_rays2_(std::move(rhs.dev_rays2_)), dev_hits_(std::move(rhs.dev_hits_)),
          dev_hits2_(std::move(rhs.dev_hits2_)), dev_states_(std::move(rhs.dev_states_)),
          dev_states2_(std::move(rhs.dev_states2_)) {
        rhs.last_ = -1;
    }
    RayQueue& operator=(RayQueue&& rhs) {
        ray_buffer_ = std::move(rhs.ray_buffer_);
        hit_buffer_ = std::move(rhs.hit_buffer_);
        state_buffer_ = std::move(rhs.state_buffer_);
        sorted_indices_ = std::move(rhs.sorted_indices_);
        last_ = rhs.last_.load();
        gpu_buffers_ = rhs.gpu_buffers_;
        dev_rays_ = std::move(rhs.dev_rays_);
        dev_rays2_ = std::move(rhs.dev_rays2_);
        dev_hits_ = std::move(rhs.dev_hits_);
        dev_hits2_ = std::move(rhs.dev_hits2_);
        dev_states_ = std::move(rhs.dev_states_);
        dev_states2_ = std::move(rhs.dev_states2_);
        rhs.last_ = -1;
        return *this;
    }

    int size() const { return last_ + 1; }
    int capacity() const { return (int)state_buffer_.size(); }
    void shrink(int size) { last_ = size - 1; }

    ImbaRay* rays() { return ray_buffer_.data(); }
    StateType* states() { return state_buffer_.data(); }
    ImbaHit* hits() { return hit_buffer_.data(); }

    ImbaRay& ray(int idx) { return ray_buffer_[sorted_indices_[idx]]; }
    ImbaHit& hit(int idx) { return hit_buffer_[sorted_indices_[idx]]; }
    StateType& state(int idx) { return state_buffer_[sorted_indices_[idx]]; }

    void clear() { last_ = -1; }

    // Thread-safe append of one ray (atomic slot).
    void push(const ImbaRay& ray, const StateType& state) {
        const int id = ++last_;
        assert(id < (int)ray_buffer_.size() && "ray queue full");
        ray_buffer_[id] = ray;
        state_buffer_[id] = state;
    }

    template <typename RayIter, typename StateIter>
    void push(RayIter rays_begin, RayIter rays_end, StateIter states_begin, StateIter) {
        const int count = (int)(rays_end - rays_begin);
        const int end_idx = last_ += count;
        const int start_idx = end_idx - (count - 1);
        assert(end_idx < (int)ray_buffer_.size() && "ray queue full");
        std::copy(rays_begin, rays_end, ray_buffer_.begin() + start_idx);
        std::copy(states_begin, states_begin + count, state_buffer_.begin() + start_idx);
    }

    void append(const RayQueue& other) {
        const int count = other.size();
        const int end_idx = last_ += count;
        const int start_idx = end_idx - (count - 1);
        assert(end_idx < (int)ray_buffer_.size() && "ray queue full");
        std::copy(other.ray_buffer_.begin(), other.ray_buffer_.begin() + count, ray_buffer_.begin() + start_idx);
        std::copy(other.state_buffer_.begin(), other.state_buffer_.begin() + count, state_buffer_.begin() + start_idx);
    }

    // Moves rays that hit something (with their states and hits) to the front, in order;
    // returns their number.  Resets the sorted order to the identity.
    int compact_hits() {
        const int n = size();
        int hits = 0;
        if (n > 0) {
            stage(n, true);
            check(imba_compact_hits(dev_rays_.data(), dev_hits_.data(), dev_states_.data(), (int)sizeof(StateType), n,
                                    dev_rays2_.data(), dev_hits2_.data(), dev_states2_.data(), &hits, nullptr),
                  "imba_compact_hits");
            dev_rays2_.download(ray_buffer_.data(), n);
            dev_hits2_.download(hit_buffer_.data(), n);
            dev_states2_.download(reinterpret_cast<char*>(state_buffer_.data()), n * sizeof(StateType));
        }
        for (int i = 0; i <= last_; ++i) sorted_indices_[i] = i;
        return hits;
    }

    // Moves continued rays (state.pixel_id >= 0) to the front in order; hits do not move.
    // The size changes only if at least one ray was terminated.
    void compact_rays() {
        const int n = size();
        if (n <= 0) return;
        stage(n, false);
        int kept = n;
        check(imba_compact_rays(dev_rays_.data(), dev_states_.data(), (int)sizeof(StateType), n, dev_rays2_.data(),
                                dev_states2_.data(), &kept, nullptr),
              "imba_compact_rays");
        dev_rays2_.download(ray_buffer_.data(), kept);
        dev_states2_.download(reinterpret_cast<char*>(state_buffer_.data()), kept * sizeof(StateType));
        if (kept != n) shrink(kept);
    }

    typedef std::function<int(const ImbaHit&)> GetMatIDFn;

    // Counting sort of the first `count` hits by material (ray_queue.h:211-244): the stored
    // order is the order of the hits within each material.
    void sort_by_material(GetMatIDFn get_mat_id, int num_mats, int count) {
        std::vector<int> start(num_mats + 1, 0);
        std::vector<int> mat(count);
        for (int i = 0; i < count; ++i) {
            mat[i] = get_mat_id(hit_buffer_[i]);
            float f;
            std::memcpy(&f, &mat[i], 4);
            ray_buffer_[i].dir.w = f;
            start[mat[i] + 1]++;
        }
        for (int m = 0; m < num_mats; ++m) start[m + 1] += start[m];
        for (int i = 0; i < count; ++i) sorted_indices_[start[mat[i]]++] = i;
    }

    // The same sort on the GPU with the material lookup of the integrators (pt.cpp:85-91):
    // the material of a hit is data.indices[4 tri_id + 3].  Hits without a valid material
    // come last.
    template <typename Node>
    void sort_by_material_gpu(const TraversalData<Node>& data, int num_mats, int count) {
        if (count <= 0) return;
        stage(count, true);
        DeviceArray<int> order((size_t)count);
        check(imba_sort_by_material(data.indices.data(), (int)(data.indices.size() / 4), dev_rays_.data(),
                                    dev_hits_.data(), count, num_mats, order.data(), nullptr, nullptr),
              "imba_sort_by_material");
        dev_rays_.download(ray_buffer_.data(), count);
        order.download(sorted_indices_.data(), count);
    }

    // The reference's CPU traversal (ray_queue.h:247-266, 296-316: the AnyDSL CPU kernels) is
    // not part of this library, which traverses on the GPU only.  The members are declared so
    // that a caller written against the reference compiles; calling them throws
    // imba::Unsupported (an imba::Error, code IMBA_ERR_INVALID_ARG).
    template <typename Node>
    void traverse_cpu(const TraversalData<Node>&) {
        throw Unsupported("RayQueue::traverse_cpu is not supported (GPU traversal only; use traverse_gpu)");
    }
    template <typename Node>
    void traverse_occluded_cpu(const TraversalData<Node>&) {
        throw Unsupported("RayQueue::traverse_occluded_cpu is not supported (GPU traversal only; use traverse_occluded_gpu)");
    }

    // Traverses all rays of the queue on the GPU (closest hit).
    template <typename Node>
    void traverse_gpu(const TraversalData<Node>& c_data) {
        assert(size() != 0);
        assert(gpu_buffers_);
        traverse(c_data, false);
    }

    // Traverses all rays of the queue on the GPU (shadow rays: any hit).
    template <typename Node>
    void traverse_occluded_gpu(const TraversalData<Node>& c_data) {
        assert(size() != 0);
        assert(gpu_buffers_);
        traverse(c_data, true);
    }

private:
    template <typename Node>
    void traverse(const TraversalData<Node>& c_data, bool any) {
        const TraversalData<Node>& d = c_data;
        ImbaSceneDesc desc;
        desc.root = d.root;
        desc.nodes = reinterpret_cast<const ImbaNode*>(d.nodes.data());
        desc.node_count = (int)d.nodes.size();
        desc.instances = d.instances.data();
        desc.instance_count = (int)d.instances.size();
        desc.tris = d.tris.data();
        desc.tri_vec4_count = (int)d.tris.size();
        desc.indices = d.indices.data();
        desc.index_count = (int)d.indices.size();
        desc.texcoords = d.texcoords.data();
        desc.texcoord_count = (int)d.texcoords.size();
        desc.masks = d.masks.data();
        desc.mask_count = (int)d.masks.size();
        desc.mask_buffer = reinterpret_cast<const uint8_t*>(d.mask_buffer.data());
        desc.mask_buffer_size = (int)d.mask_buffer.size();
        dev_rays_.upload(ray_buffer_.data(), size());
        // The reference traverses align_gpu(size()) rays (the padding is never read back).
        check(imba_trace_desc(&desc, dev_rays_.data(), dev_hits_.data(), size(), any ? 1 : 0, nullptr),
              "imba_trace_desc");
        dev_hits_.download(hit_buffer_.data(), size());
    }

    // Device copies of the queue for a compaction (rays are uploaded again: the host may have
    // rewritten them since the traversal).
    void stage(int n, bool with_hits) {
        if (dev_rays_.size() < (size_t)n) dev_rays_.resize(align_gpu(capacity()));
        if (dev_hits_.size() < (size_t)n) dev_hits_.resize(align_gpu(capacity()));
        if (dev_states_.size() < (size_t)n * sizeof(StateType)) {
            dev_states_.resize(align_gpu(capacity()) * sizeof(StateType));
            dev_states2_.resize(align_gpu(capacity()) * sizeof(StateType));
            dev_rays2_.resize(align_gpu(capacity()));
            dev_hits2_.resize(align_gpu(capacity()));
        }
        dev_rays_.upload(ray_buffer_.data(), n);
        if (with_hits) dev_hits_.upload(hit_buffer_.data(), n);
        dev_states_.upload(reinterpret_cast<const char*>(state_buffer_.data()), n * sizeof(StateType));
    }

    std::vector<ImbaRay> ray_buffer_;
    std::vector<ImbaHit> hit_buffer_;
    std::vector<StateType> state_buffer_;
    std::vector<int> sorted_indices_;
    std::atomic<int> last_;
    bool gpu_buffers_;
    DeviceArray<ImbaRay> dev_rays_, dev_rays2_;
    DeviceArray<ImbaHit> dev_hits_, dev_hits2_;
    DeviceArray<char> dev_states_, dev_states2_;
};

namespace detail {

// Device and host buffers of one host thread's grid queries.  query() is const and the
// integrators call it from many shading threads at once (vcm.cpp:370 -> vertex_merging ->
// get_merge), so every thread queries through buffers of its own, on its own stream (the
// library's per-thread stream), and nothing of the grid object is written.  The buffers are
// allocated on a thread's first query on a device (or when a query asks for more results than
// they hold), not on every call, and freed when the thread ends.
struct QueryScratch {
    size_t cap = 0;          // positions the buffers hold
    int k_cap = 0;           // results per position they hold
    float* dq = nullptr;     // device: 3 floats per position
    int* di = nullptr;       // device: k_cap ints per position
    int* dc = nullptr;       // device: 1 int per position
    std::vector<int> host_idx;
    std::vector<int> host_cnt;

    QueryScratch() {}
    QueryScratch(const QueryScratch&) = delete;
    QueryScratch& operator=(const QueryScratch&) = delete;
    ~QueryScratch() { release(); }

    void release() {
        if (dq) imba_free(dq);
        if (di) imba_free(di);
        if (dc) imba_free(dc);
        dq = nullptr;
        di = nullptr;
        dc = nullptr;
        cap = 0;
        k_cap = 0;
    }

    // Buffers for nq positions with k results each (on the device this scratch belongs to,
    // which is current).
    void reserve(size_t nq, int k) {
        if (nq <= cap && k <= k_cap) return;
        const size_t n = std::max(nq, cap);
        const int kk = std::max(k, k_cap);
        release();
        void* p = nullptr;
        check(imba_malloc(&p, 3 * sizeof(float) * n), "imba_malloc");
        dq = static_cast<float*>(p);
        check(imba_malloc(&p, sizeof(int) * n * (size_t)kk), "imba_malloc");
        di = static_cast<int*>(p);
        check(imba_malloc(&p, sizeof(int) * n), "imba_malloc");
        dc = static_cast<int*>(p);
        cap = n;
        k_cap = kk;
    }
};

// The calling thread's scratch on the current device.
inline QueryScratch& query_scratch() {
    thread_local std::vector<std::unique_ptr<QueryScratch>> per_device;
    int dev = 0;
    check(imba_get_device(&dev), "imba_get_device");
    if ((size_t)dev >= per_device.size()) per_device.resize((size_t)dev + 1);
    if (!per_device[dev]) per_device[dev].reset(new QueryScratch());
    return *per_device[dev];
}

}  // namespace detail

// Photon hash grid (rangesearch.h:21-158).  Photon must be trivially copyable with a public
// `float3`-like position of three floats at offset 0 as `position` (VCMPhoton) or returned
// by `position()` (LightPathVertex) — the grid reads 3 floats at the start of each record.
// build() must not overlap with queries; queries may run concurrently from any number of
// host threads (each uses buffers of its own, see detail::QueryScratch).
template <typename Iter, typename Photon>
class HashGrid {
public:
    HashGrid() { check(imba_hashgrid_create(&grid_), "imba_hashgrid_create"); }
    ~HashGrid() { imba_hashgrid_release(grid_); }
    HashGrid(const HashGrid&) = delete;
    HashGrid& operator=(const HashGrid&) = delete;

    void build(const Iter& begin, const Iter& end, float radius) {
        typedef typename std::iterator_traits<Iter>::value_type Src;
        const int n = (int)(end - begin);
        src_.assign(begin, end);
        photons_.resize(n);
        for (int i = 0; i < n; i++) photons_[i] = Photon(src_[i]);
        dev_src_.resize(std::max(n, 1) * sizeof(Src));
        if (n) dev_src_.upload(reinterpret_cast<const char*>(src_.data()), n * sizeof(Src));
        check(imba_hashgrid_build(grid_, dev_src_.data(), (int)sizeof(Src), 0, n, radius, nullptr),
              "imba_hashgrid_build");
    }

    // Up to k photons within the radius of query_pos, nearest first (out[i] = const Photon*).
    // Reentrant: concurrent calls from several threads are allowed.
    template <typename Container, typename Float3>
    int query(const Float3& query_pos, Container& out, int k) const {
        if (k <= 0) return 0;
        const float q[3] = {query_pos.x, query_pos.y, query_pos.z};
        detail::QueryScratch& sc = detail::query_scratch();
        run_queries(sc, q, 1, k);
        const int n = sc.host_cnt[0];
        for (int i = 0; i < n; i++) out[i] = &photons_[sc.host_idx[i]];
        return n;
    }

    // nq positions (3 floats each, host) -> per query up to k photon indices (into the build
    // input) in idx[q * k ...] and the number found in counts[q].  Reentrant.
    void query_batch(const float* positions, int nq, int k, std::vector<int>& idx, std::vector<int>& counts) const {
        idx.assign((size_t)std::max(nq, 0) * std::max(k, 0), -1);
        counts.assign((size_t)std::max(nq, 0), 0);
        if (nq <= 0 || k <= 0) return;
        detail::QueryScratch& sc = detail::query_scratch();
        run_queries(sc, positions, nq, k);
        for (int q = 0; q < nq; q++) {
            counts[q] = sc.host_cnt[q];
            std::copy(sc.host_idx.begin() + (size_t)q * k, sc.host_idx.begin() + (size_t)q * k + counts[q],
                      idx.begin() + (size_t)q * k);
        }
    }

private:
    // Runs nq queries on the calling thread's stream through its scratch; results end in
    // sc.host_idx (k per query) and sc.host_cnt.
    void run_queries(detail::QueryScratch& sc, const float* positions, int nq, int k) const {
        sc.reserve((size_t)nq, k);
        check(imba_memcpy_h2d(sc.dq, positions, 3 * sizeof(float) * (size_t)nq, nullptr), "imba_memcpy_h2d");
        check(imba_hashgrid_query(grid_, sc.dq, 12, 0, nq, k, sc.di, nullptr, sc.dc, nullptr), "imba_hashgrid_query");
        sc.host_idx.resize((size_t)nq * k);
        sc.host_cnt.resize((size_t)nq);
        check(imba_memcpy_d2h(sc.host_idx.data(), sc.di, sizeof(int) * (size_t)nq * k, nullptr), "imba_memcpy_d2h");
        check(imba_memcpy_d2h(sc.host_cnt.data(), sc.dc, sizeof(int) * (size_t)nq, nullptr), "imba_memcpy_d2h");
    }

    typedef typename std::iterator_traits<Iter>::value_type SrcT;
    ImbaHashGrid* grid_ = nullptr;
    std::vector<SrcT> src_;
    std::vector<Photon> photons_;
    DeviceArray<char> dev_src_;
};

}  // namespace imba

#endif  // IMBA_RAY_QUEUE_HPP
