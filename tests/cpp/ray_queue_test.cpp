// Drives the reference-shaped C++ interfaces (include/imba/ray_queue.hpp) the way the
// reference's schedulers and integrators do (tile_scheduler.h:100-144, pt.cpp:82-160,
// light_vertices.h:70-100) and checks every result against the CPU oracle.
// Built by __graft_entry__.build() into tests/cpp/ray_queue_test.bin; run by
// tests/test_gpu_cpp_interface.py on the GPU.  Exit status 0 = all checks passed.
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#include "imba/ray_queue.hpp"
#include "traversal_gpu.h"
#include "../../oracle/oracle.h"

namespace {

struct State {  // RayState (ray_queue.h:20-33): pixel_id first
    int pixel_id;
    int sample_id;
    unsigned long long rng;
};

struct PhotonSrc {  // a light-vertex-like record: position first
    float pos[3];
    float payload[5];
};
struct Photon {
    float position[3];
    Photon() {}
    explicit Photon(const PhotonSrc& s) { std::memcpy(position, s.pos, sizeof(position)); }
};
struct F3 {
    float x, y, z;
};

int failures = 0;
#define EXPECT(cond, ...)                 \
    do {                                  \
        if (!(cond)) {                    \
            std::printf("FAIL: " __VA_ARGS__); \
            std::printf("\n");            \
            failures++;                   \
        }                                 \
    } while (0)

bool same_hit(const ImbaHit& a, const ImbaHit& b) { return std::memcmp(&a, &b, sizeof(ImbaHit)) == 0; }


}  // namespace

int main() {
    int ndev = 0;
    imba_device_count(&ndev);
    if (ndev < 1) {
        std::printf("no HIP device\n");
        return 2;
    }
    // scene: 20,000 random triangles, SBVH
    const int ntri = 20000;
    std::vector<float> verts(12 * ntri);
    std::vector<uint32_t> idx(4 * ntri);
    imba::check(imba_gen_random_triangles(0x1234, ntri, verts.data(), idx.data()), "gen");
    for (int t = 0; t < ntri; t++) idx[4 * t + 3] = t % 7;  // seven materials
    ImbaSceneBuild* b = nullptr;
    imba::check(imba_build_create(&b), "build_create");
    ImbaMeshDesc md{verts.data(), 3 * ntri, idx.data(), ntri, nullptr};
    imba::check(imba_build_add_mesh(b, &md, nullptr), "add_mesh");
    imba::check(imba_build_run(b, IMBA_BUILDER_SBVH, 2), "build_run");
    ImbaSceneDesc host;
    imba::check(imba_build_get_desc(b, &host), "get_desc");
    imba::TraversalData<> data;
    data.upload(host);

    // camera rays into a queue, as PixelRayGen::fill_queue + push does
    const int w = 160, h = 120;
    std::vector<ImbaRay> cam(w * h);
    const float pos[3] = {0.5f, 0.5f, -1.5f}, dir[3] = {0, 0, 1}, up[3] = {0, 1, 0};
    imba::check(imba_gen_camera_rays(pos, dir, up, 60.0f, w, h, cam.data()), "camera");
    imba::RayQueue<State> q(w * h, true);
    for (int i = 0; i < w * h; i++) q.push(cam[i], State{i, 0, (unsigned long long)i});
    EXPECT(q.size() == w * h, "queue size %d", q.size());

    q.traverse_gpu(data);
    std::vector<ImbaHit> want(w * h);
    oracle_trace(&host, cam.data(), w * h, 0, want.data(), nullptr);
    int hits = 0;
    for (int i = 0; i < w * h; i++) {
        EXPECT(same_hit(q.hits()[i], want[i]), "closest hit of ray %d differs", i);
        hits += want[i].tri_id >= 0;
    }
    EXPECT(hits > 100, "too few hits (%d)", hits);

    // compact_hits: the hit prefix equals the reference loop's
    std::vector<ImbaRay> r_ref(q.rays(), q.rays() + q.size());
    std::vector<ImbaHit> h_ref(q.hits(), q.hits() + q.size());
    std::vector<State> s_ref(q.states(), q.states() + q.size());
    const int n_ref = oracle_compact_hits(r_ref.data(), h_ref.data(), s_ref.data(), sizeof(State), q.size());
    const int n = q.compact_hits();
    EXPECT(n == n_ref && n == hits, "compact_hits returned %d, oracle %d", n, n_ref);
    for (int i = 0; i < n; i++) {
        EXPECT(same_hit(q.hits()[i], h_ref[i]) && q.states()[i].pixel_id == s_ref[i].pixel_id &&
                   std::memcmp(&q.rays()[i], &r_ref[i], sizeof(ImbaRay)) == 0,
               "compact_hits entry %d differs", i);
        EXPECT(q.hit(i).tri_id >= 0, "sorted index %d not reset", i);
    }
    q.shrink(n);

    // sort_by_material on the GPU (the material lookup of pt.cpp:85-91) against the serial
    // restatement: same keys in dir.w, same grouped order
    {
        std::vector<ImbaRay> rs(q.rays(), q.rays() + n);
        std::vector<ImbaHit> hs(q.hits(), q.hits() + n);
        std::vector<int> order_ref(n);
        oracle_sort_by_material(host.indices, host.index_count / 4, rs.data(), hs.data(), n, 7, order_ref.data(),
                                nullptr);
        q.sort_by_material_gpu(data, 7, n);
        for (int i = 0; i < n; i++) {
            EXPECT(&q.hit(i) - q.hits() == order_ref[i], "sort_by_material position %d differs", i);
            EXPECT(std::memcmp(&q.rays()[i], &rs[i], sizeof(ImbaRay)) == 0, "sort_by_material key of %d differs", i);
        }
    }

    // terminate every third path and compact the rays (pt.cpp:150 / compact_rays)
    for (int i = 0; i < q.size(); i += 3) q.states()[i].pixel_id = -1;
    std::vector<ImbaRay> r2(q.rays(), q.rays() + q.size());
    std::vector<State> s2(q.states(), q.states() + q.size());
    const int kept_ref = oracle_compact_rays(r2.data(), s2.data(), sizeof(State), q.size());
    q.compact_rays();
    EXPECT(q.size() == kept_ref, "compact_rays size %d, oracle %d", q.size(), kept_ref);
    for (int i = 0; i < q.size(); i++)
        EXPECT(q.states()[i].pixel_id == s2[i].pixel_id && std::memcmp(&q.rays()[i], &r2[i], sizeof(ImbaRay)) == 0,
               "compact_rays entry %d differs", i);

    // shadow rays towards a point above the cloud (traverse_occluded_gpu)
    imba::RayQueue<State> sq(w * h, true);
    std::mt19937 rng(7);
    std::uniform_real_distribution<float> U(0.0f, 1.0f);
    std::vector<ImbaRay> shadow;
    for (int i = 0; i < 4000; i++) {
        const float o[3] = {U(rng), U(rng), U(rng)};
        float d[3] = {0.5f - o[0], 2.0f - o[1], 0.5f - o[2]};
        const float len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        for (float& c : d) c /= len;
        ImbaRay r{{o[0], o[1], o[2], 1e-3f * len}, {d[0], d[1], d[2], len - 1e-3f * len}};
        shadow.push_back(r);
        sq.push(r, State{i, 0, 0});
    }
    sq.traverse_occluded_gpu(data);
    std::vector<ImbaHit> want_o(shadow.size());
    oracle_trace(&host, shadow.data(), (int)shadow.size(), 1, want_o.data(), nullptr);
    for (size_t i = 0; i < shadow.size(); i++)
        EXPECT(same_hit(sq.hits()[i], want_o[i]), "occluded hit of ray %zu differs", i);

    // photon hash grid (LightVertices::build / get_merge)
    std::vector<PhotonSrc> photons(20000);
    for (auto& p : photons) {
        for (float& c : p.pos) c = U(rng);
        for (float& c : p.payload) c = U(rng);
    }
    imba::HashGrid<std::vector<PhotonSrc>::iterator, Photon> grid;
    grid.build(photons.begin(), photons.end(), 0.03f);
    OracleGrid og;
    std::memset(&og, 0, sizeof(og));
    oracle_grid_build(&og, &photons[0].pos[0], sizeof(PhotonSrc) / 4, (int)photons.size(), 0.03f);
    const int k = 10, nq = 500;
    std::vector<float> qpos(3 * nq);
    for (float& c : qpos) c = U(rng);
    std::vector<int> got, got_counts;
    grid.query_batch(qpos.data(), nq, k, got, got_counts);
    int nonempty = 0;
    std::vector<int> want_idx((size_t)nq * k, -1), want_cnt(nq);
    for (int qi = 0; qi < nq; qi++) {
        int oi[64];
        float od[64];
        const int m = oracle_grid_query(&og, &qpos[3 * qi], k, oi, od);
        want_cnt[qi] = m;
        for (int j = 0; j < m; j++) want_idx[(size_t)qi * k + j] = oi[j];
        EXPECT(got_counts[qi] == m, "query %d count %d, oracle %d", qi, got_counts[qi], m);
        for (int j = 0; j < m && j < got_counts[qi]; j++) EXPECT(got[qi * k + j] == oi[j], "query %d slot %d", qi, j);
        nonempty += m > 0;
    }
    EXPECT(nonempty > 50, "too few non-empty queries (%d)", nonempty);
    // single-position query() returns pointers to photons with the right positions
    const Photon* out[16];
    const F3 p0{qpos[0], qpos[1], qpos[2]};
    const int m0 = grid.query(p0, out, k);
    EXPECT(m0 == got_counts[0] || nq == 0, "query() count");
    for (int j = 0; j < m0; j++)
        EXPECT(std::memcmp(out[j]->position, photons[got[j]].pos, 12) == 0, "query() photon %d", j);

    // query() from 8 threads at once (vertex_merging runs on the TBB shading threads,
    // vcm.cpp:370-380): every thread's results equal the oracle's.  Results point into the
    // grid's own photon array, whose start follows from the first query's results.
    {
        const Photon* base = m0 > 0 ? out[0] - got[0] : nullptr;
        std::atomic<int> bad(0);
        std::vector<std::thread> pool;
        for (int t = 0; t < 8; t++) {
            pool.emplace_back([&, t]() {
                try {
                    for (int rep = 0; rep < 3; rep++) {
                        for (int qi = t; qi < nq; qi += 8) {
                            const Photon* o[16];
                            const F3 p{qpos[3 * qi], qpos[3 * qi + 1], qpos[3 * qi + 2]};
                            const int m = grid.query(p, o, k);
                            if (m != want_cnt[qi]) {
                                bad++;
                                continue;
                            }
                            for (int j = 0; j < m; j++) {
                                const int want_j = want_idx[(size_t)qi * k + j];
                                if (std::memcmp(o[j]->position, photons[want_j].pos, 12) != 0 ||
                                    (base && o[j] != base + want_j))
                                    bad++;
                            }
                        }
                    }
                } catch (const std::exception& e) {
                    std::printf("thread %d: %s\n", t, e.what());
                    bad++;
                }
            });
        }
        for (auto& th : pool) th.join();
        EXPECT(bad.load() == 0, "concurrent HashGrid::query: %d mismatches", bad.load());
    }
    oracle_grid_free(&og);

    // four RayQueues traversed from four threads at once, as TileScheduler's render threads do
    // (tile_scheduler.h:69-78,120-132): each queue's hits equal the oracle's
    {
        std::atomic<int> bad(0);
        std::vector<std::thread> pool;
        for (int t = 0; t < 4; t++) {
            pool.emplace_back([&, t]() {
                try {
                    imba::RayQueue<State> tq(w * h, true);
                    for (int i = t; i < w * h; i += 4) tq.push(cam[i], State{i, 0, 0});
                    for (int rep = 0; rep < 3; rep++) {
                        tq.traverse_gpu(data);
                        for (int j = 0; j < tq.size(); j++)
                            if (!same_hit(tq.hits()[j], want[tq.states()[j].pixel_id])) bad++;
                        tq.traverse_occluded_gpu(data);
                    }
                } catch (const std::exception& e) {
                    std::printf("thread %d: %s\n", t, e.what());
                    bad++;
                }
            });
        }
        for (auto& th : pool) th.join();
        EXPECT(bad.load() == 0, "concurrent traverse_gpu: %d mismatches", bad.load());
    }

    // a moved queue keeps its rays, states and device buffers
    {
        imba::RayQueue<State> a(256, true);
        for (int i = 0; i < 256; i++) a.push(cam[i], State{i, 0, 0});
        imba::RayQueue<State> moved(std::move(a));
        EXPECT(moved.size() == 256 && a.size() == 0, "move: sizes %d / %d", moved.size(), a.size());
        moved.traverse_gpu(data);
        for (int i = 0; i < 256; i++) EXPECT(same_hit(moved.hits()[i], want[i]), "moved queue hit %d", i);
        imba::RayQueue<State> assigned;
        assigned = std::move(moved);
        assigned.traverse_gpu(data);
        for (int i = 0; i < 256; i++) EXPECT(same_hit(assigned.hits()[i], want[i]), "move-assigned queue hit %d", i);
        bool threw = false;
        try {
            assigned.traverse_cpu(data);
        } catch (const imba::Unsupported&) {
            threw = true;
        }
        EXPECT(threw, "traverse_cpu must throw imba::Unsupported");
    }

    // reference FFI over caller-owned arrays that are rewritten in place and reallocated
    // between calls (Scene::upload_top_level_accel writes into the arrays it has,
    // scene.cpp:166-188): every call's hits equal the oracle's for what the arrays hold then
    {
        std::vector<float> verts2(12 * ntri);
        std::vector<uint32_t> idx2(4 * ntri);
        imba::check(imba_gen_random_triangles(0xBEEF, ntri, verts2.data(), idx2.data()), "gen");
        for (int t = 0; t < ntri; t++) idx2[4 * t + 3] = t % 3;
        ImbaSceneBuild* b2 = nullptr;
        imba::check(imba_build_create(&b2), "build_create");
        ImbaMeshDesc md2{verts2.data(), 3 * ntri, idx2.data(), ntri, nullptr};
        imba::check(imba_build_add_mesh(b2, &md2, nullptr), "add_mesh");
        // a second instance (translated), so the top level differs too
        const float shift[16] = {1, 0, 0, 0.25f, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
        imba::check(imba_build_add_instance(b2, 0, shift, nullptr), "add_instance");
        imba::check(imba_build_run(b2, IMBA_BUILDER_BINNED_SAH, 2), "build_run");
        ImbaSceneDesc host2;
        imba::check(imba_build_get_desc(b2, &host2), "get_desc");
        const ImbaSceneDesc* scenes[2] = {&host, &host2};
        std::vector<ImbaHit> ffi_want[2];
        for (int k2 = 0; k2 < 2; k2++) {
            ffi_want[k2].resize(w * h);
            oracle_trace(scenes[k2], cam.data(), w * h, 0, ffi_want[k2].data(), nullptr);
        }
        struct Arrays {
            void *nodes = nullptr, *inst = nullptr, *tris = nullptr, *indices = nullptr, *tex = nullptr,
                 *masks = nullptr, *mbuf = nullptr;
        };
        auto alloc = [](Arrays& a, const ImbaSceneDesc& d0, const ImbaSceneDesc& d1) {
            auto mx = [](int x, int y) { return (size_t)std::max(std::max(x, y), 1); };
            imba::check(imba_malloc(&a.nodes, sizeof(ImbaNode) * mx(d0.node_count, d1.node_count)), "malloc");
            imba::check(imba_malloc(&a.inst, sizeof(ImbaInstanceNode) * mx(d0.instance_count, d1.instance_count)), "malloc");
            imba::check(imba_malloc(&a.tris, sizeof(ImbaVec4) * mx(d0.tri_vec4_count, d1.tri_vec4_count)), "malloc");
            imba::check(imba_malloc(&a.indices, sizeof(int) * mx(d0.index_count, d1.index_count)), "malloc");
            imba::check(imba_malloc(&a.tex, sizeof(ImbaVec2) * mx(d0.texcoord_count, d1.texcoord_count)), "malloc");
            imba::check(imba_malloc(&a.masks, sizeof(ImbaTransparencyMask) * mx(d0.mask_count, d1.mask_count)), "malloc");
            imba::check(imba_malloc(&a.mbuf, mx(d0.mask_buffer_size, d1.mask_buffer_size)), "malloc");
        };
        auto release = [](Arrays& a) {
            for (void* p : {a.nodes, a.inst, a.tris, a.indices, a.tex, a.masks, a.mbuf}) imba_free(p);
            a = Arrays();
        };
        auto upload = [](Arrays& a, const ImbaSceneDesc& d) {
            imba::check(imba_memcpy_h2d(a.nodes, d.nodes, sizeof(ImbaNode) * d.node_count, nullptr), "h2d");
            imba::check(imba_memcpy_h2d(a.inst, d.instances, sizeof(ImbaInstanceNode) * d.instance_count, nullptr), "h2d");
            imba::check(imba_memcpy_h2d(a.tris, d.tris, sizeof(ImbaVec4) * d.tri_vec4_count, nullptr), "h2d");
            if (d.index_count) imba::check(imba_memcpy_h2d(a.indices, d.indices, sizeof(int) * d.index_count, nullptr), "h2d");
            if (d.texcoord_count)
                imba::check(imba_memcpy_h2d(a.tex, d.texcoords, sizeof(ImbaVec2) * d.texcoord_count, nullptr), "h2d");
            if (d.mask_count)
                imba::check(imba_memcpy_h2d(a.masks, d.masks, sizeof(ImbaTransparencyMask) * d.mask_count, nullptr), "h2d");
            if (d.mask_buffer_size) imba::check(imba_memcpy_h2d(a.mbuf, d.mask_buffer, d.mask_buffer_size, nullptr), "h2d");
        };
        void *drays = nullptr, *dhits = nullptr;
        imba::check(imba_malloc(&drays, sizeof(ImbaRay) * w * h), "malloc");
        imba::check(imba_malloc(&dhits, sizeof(ImbaHit) * w * h), "malloc");
        imba::check(imba_memcpy_h2d(drays, cam.data(), sizeof(ImbaRay) * w * h, nullptr), "h2d");
        std::vector<ImbaHit> got_h(w * h);
        auto ffi_run = [&](const Arrays& a, const ImbaSceneDesc& d, int which, const char* what) {
            intersect_gpu_masked_instanced(d.root, (Node*)a.nodes, (InstanceNode*)a.inst, (Vec4*)a.tris, (Ray*)drays,
                                           (Hit*)dhits, (int*)a.indices, (Vec2*)a.tex, (TransparencyMask*)a.masks,
                                           (char*)a.mbuf, w * h);
            imba::check(imba_memcpy_d2h(got_h.data(), dhits, sizeof(ImbaHit) * w * h, nullptr), "d2h");
            int diff = 0;
            for (int i = 0; i < w * h; i++) diff += !same_hit(got_h[i], ffi_want[which][i]);
            EXPECT(diff == 0, "FFI (%s): %d hits differ from the oracle", what, diff);
        };
        int err0 = 0;
        imba::check(imba_ffi_error_count(&err0), "ffi_error_count");
        Arrays a;
        alloc(a, host, host2);
        upload(a, host);
        ffi_run(a, host, 0, "scene A");
        upload(a, host2);  // same allocations, new contents
        ffi_run(a, host2, 1, "scene B rewritten in place over A");
        upload(a, host);
        ffi_run(a, host, 0, "scene A rewritten in place over B");
        const void* old_nodes = a.nodes;
        release(a);
        alloc(a, host2, host2);  // freed and reallocated (often at the same addresses)
        upload(a, host2);
        ffi_run(a, host2, 1, "scene B after free and reallocation");
        std::printf("FFI reallocation reused the node array address: %s\n", a.nodes == old_nodes ? "yes" : "no");
        release(a);
        int err1 = 0;
        imba::check(imba_ffi_error_count(&err1), "ffi_error_count");
        EXPECT(err1 == err0, "FFI calls reported %d malformed rays", err1 - err0);
        imba_free(drays);
        imba_free(dhits);
        imba_build_release(b2);
    }
    imba_build_release(b);

    if (failures) {
        std::printf("ray_queue_test: %d failures\n", failures);
        return 1;
    }
    std::printf("ray_queue_test OK (%d hits, %d shadow rays, %d grid queries)\n", hits, (int)shadow.size(), nq);
    return 0;
}
