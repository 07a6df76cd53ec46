// Drives the reference-shaped C++ interfaces (include/imba/ray_queue.hpp) the way the
// reference's schedulers and integrators do (tile_scheduler.h:100-144, pt.cpp:82-160,
// light_vertices.h:70-100) and checks every result against the CPU oracle.
// Built by __graft_entry__.build() into tests/cpp/ray_queue_test.bin; run by
// tests/test_gpu_cpp_interface.py on the GPU.  Exit status 0 = all checks passed.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "imba/ray_queue.hpp"
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
    std::vector<int> got;
    grid.query_batch(qpos.data(), nq, k, got);
    int nonempty = 0;
    for (int qi = 0; qi < nq; qi++) {
        int oi[64];
        float od[64];
        const int m = oracle_grid_query(&og, &qpos[3 * qi], k, oi, od);
        EXPECT(grid.counts()[qi] == m, "query %d count %d, oracle %d", qi, grid.counts()[qi], m);
        for (int j = 0; j < m && j < grid.counts()[qi]; j++) EXPECT(got[qi * k + j] == oi[j], "query %d slot %d", qi, j);
        nonempty += m > 0;
    }
    EXPECT(nonempty > 50, "too few non-empty queries (%d)", nonempty);
    // single-position query() returns pointers to photons with the right positions
    const Photon* out[16];
    const F3 p0{qpos[0], qpos[1], qpos[2]};
    const int m0 = grid.query(p0, out, k);
    EXPECT(m0 == grid.counts()[0] || nq == 0, "query() count");
    for (int j = 0; j < m0; j++)
        EXPECT(std::memcmp(out[j]->position, photons[got[j]].pos, 12) == 0, "query() photon %d", j);
    oracle_grid_free(&og);
    imba_build_release(b);

    if (failures) {
        std::printf("ray_queue_test: %d failures\n", failures);
        return 1;
    }
    std::printf("ray_queue_test OK (%d hits, %d shadow rays, %d grid queries)\n", hits, (int)shadow.size(), nq);
    return 0;
}
