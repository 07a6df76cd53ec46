"""Parity at the measured configurations' own sizes (GPU).

  C2  path tracing at 1280x720 over the 32 x 32 sphere grid (20,971,533 triangles, SBVH), on the
      quantized 4-wide tree the benchmark picks: a sample of every bounce's closest-hit queue
      (every k-th ray, about 64K per bounce) bit for bit against the oracle's traversal of the
      same rays, every bounce's shadow rays (same sample) with the oracle's occlusion answers,
      and a graph-captured frame equal to the eager one (counts and every buffer).
  C3  photon mapping over the same grid with 30% mirror / glass spheres and 921,600 light
      paths: the hash grid over the frame's ~4.2M photons (table size, bounds, cell ends, slot
      order) equal to the oracle grid's, a 20,000-query sample bit for bit, and a sample of the
      first light queue's hits against the oracle; the VCM iteration on the same scene: the
      light vertex cache, its grid and merge records, 20,000 grid queries and a light-queue
      sample, each against its restatement or the oracle.
  C4  256 queues (1K..1M rays, log-uniform, ~37M rays; half pinhole, half cosine-hemisphere)
      over 8 random-triangle scenes of 10K..4M triangles in one batched launch: a 2,048-ray
      sample of every queue against the oracle.
  C5  requests drawn from the serving mix over the 8 sphere-grid scenes (20K..21M triangles):
      every request's counts and image alone and in the mix.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle
from imbatracer_amd import native as N
from imbatracer_amd import scene as S
from imbatracer_amd import serve as V
from imbatracer_amd import workloads as W

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))
QUANTIZED = N.TraceConfig(256, 16, 0, 0, 4, 1, 0)


def _same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.int32), np.ascontiguousarray(b).view(np.int32))


def _every_kth(n, want):
    return np.arange(0, n, max(1, n // want))


@pytest.fixture(scope="module")
def grid21m(gpu):
    scene = W.SphereGridScene(side=32, level=5, builder=N.BUILDER_SBVH, threads=THREADS)
    assert scene.tri_count == 20_971_533
    return scene, scene.build.arrays()


def test_c2_pt_frame_full_size(grid21m):
    scene, arrays = grid21m
    print("C2: 1280x720 PT frame, %d triangles" % scene.tri_count, flush=True)
    pt = W.PathTracer(scene, 1280, 720, spp=1, max_path_len=8, rr_start=3, trace_cfg=QUANTIZED,
                      shadow_cfg=QUANTIZED)
    counts, _, queues = pt.frame(keep_queues=8)
    pt.stream.sync()
    assert counts[0] == 1280 * 720 and len(counts) >= 6
    checked = 0
    for b, rays in enumerate(queues):
        idx = _every_kth(len(rays), 65536)
        want, bad = oracle.trace(arrays, rays[idx], threads=THREADS)
        assert bad == 0
        assert _same(pt.kept_hits[b][idx], want), b
        srays, sst, shits = pt.kept_shadow[b]
        sidx = idx[sst["pixel"][idx] >= 0]
        want_o, bad = oracle.trace(arrays, srays[sidx], any_hit=True, threads=THREADS)
        assert bad == 0
        got = shits[sidx]
        assert np.array_equal(got["tri_id"] < 0, want_o["tri_id"] < 0), b
        hit = got["tri_id"] >= 0
        # an occluding hit lies inside the shadow ray's interval
        assert np.all((got["tmax"][hit] >= srays["org"][sidx][hit, 3]) & (got["tmax"][hit] <= srays["dir"][sidx][hit, 3]))
        checked += len(idx) + len(sidx)
    assert checked > 500_000
    assert pt.gpu_scene.error_count() == 0

    # the graph-captured frame writes what the eager frame writes
    def buffers():
        return [x.download(np.uint8, x.nbytes) for x in pt.rays + pt.states +
                [pt.hits, pt.shadow_rays, pt.shadow_states, pt.shadow_hits]]

    pt.build_graph()
    pt.frame()
    pt.stream.sync()
    eager = buffers()
    assert pt.frame_graph() == counts
    for k, (x, y) in enumerate(zip(buffers(), eager)):
        assert np.array_equal(x, y), k


def test_c3_photon_frame_full_size(grid21m):
    scene, arrays = grid21m
    c3 = scene.with_specular(0.3)
    print("C3: photon mapping, 921,600 light paths", flush=True)
    pm = W.PhotonMapper(c3, 1280, 720, light_paths=921_600, max_depth=6, rr_start=3, radius_factor=0.005, k=10,
                        trace_cfg=QUANTIZED)
    light, cam, n_ph = pm.frame()
    assert 3_000_000 < n_ph < 6_000_000
    ph_rays, _ = pm.photons_host()
    pos = np.ascontiguousarray(ph_rays["pos"])
    og = oracle.Grid()
    og.build(pos, pm.radius)
    info = pm.grid.info()
    assert info.photon_count == n_ph == og.g.photon_count
    assert info.table_size == og.g.table_size
    assert list(info.bbox_min) == list(og.g.bbox_min) and list(info.bbox_max) == list(og.g.bbox_max)
    cell_ends = np.empty(info.table_size, np.int32)
    order = np.empty(n_ph, np.int32)
    N.check(N.lib.imba_memcpy_d2h(N.ptr(cell_ends), info.cell_ends, cell_ends.nbytes, None), "d2h")
    N.check(N.lib.imba_memcpy_d2h(N.ptr(order), info.order, order.nbytes, None), "d2h")
    assert np.array_equal(cell_ends, og.cell_ends())
    assert np.array_equal(order, og.order())
    # a query sample against the oracle grid
    q = pm.queries_host()
    idx_all, dist_all, cnt_all = pm.results_host()
    valid = np.nonzero(q["pixel"] >= 0)[0]
    sample = np.sort(np.random.default_rng(3).choice(valid, size=min(len(valid), 20_000), replace=False))
    o_idx, o_dist, o_cnt = og.query_many(np.ascontiguousarray(q["pos"][sample]), pm.k)
    for r, j in enumerate(sample):
        c = cnt_all[j]
        assert c == o_cnt[r]
        assert np.array_equal(idx_all[j, :c], o_idx[r, :c])
        assert np.array_equal(dist_all[j, :c].view(np.int32), o_dist[r, :c].view(np.int32))
    assert (o_cnt > 0).sum() > 1000
    # the frame's first light queue (the light rays of imba_ppm_light) traversed on the quantized
    # tree, against the oracle
    lr, ls = N.DeviceBuffer(pm.P * 32), N.DeviceBuffer(pm.P * W.PPM_STATE_BYTES)
    N.check(N.lib.imba_ppm_light(N.ptr(pm.light_tri), N.ptr(pm.light_int), pm.light_seed, 0, pm.P, lr.ptr, ls.ptr,
                                 None), "ppm_light")
    rays0 = lr.download(N.RAY_DTYPE, pm.P)
    idx = _every_kth(pm.P, 65536)
    want, bad = oracle.trace(arrays, rays0[idx], threads=THREADS)
    assert bad == 0
    assert _same(pm.gpu_scene.trace_host(rays0[idx], cfg=QUANTIZED), want)
    assert pm.gpu_scene.error_count() == 0


def test_c3_vcm_iteration_full_size(grid21m):
    """One VCM iteration at the configuration's size: the light vertex cache is the stable
    compaction of every light bounce's vertex slots, the grid over it (table, bounds, cell ends,
    slot order) is the oracle grid's, the merge records are the cache in that order, 20,000 grid
    queries around cached vertices return the oracle's lists bit for bit, and the first light
    queue's hits are the oracle's on a sample."""
    from imbatracer_amd import vcm as VC

    scene, arrays = grid21m
    c3 = scene.with_specular(0.3)
    print("C3: VCM iteration, 921,600 light paths", flush=True)
    vr = VC.VcmRenderer(c3, 1280, 720, light_paths=921_600, max_path_len=6, light_max_depth=6, rr_start=0,
                        radius_factor=0.005, k=10, trace_cfg=QUANTIZED, shadow_cfg=QUANTIZED)
    light, cam, nv = vr.frame()
    assert light[0] == 921_600 and cam[0] == 921_600
    assert 3_000_000 < nv < 6_000_000
    a, b = vr.cache_host()
    # the cache is the occupied slots in (bounce, path) order
    sa = vr.slot_a.download(VC.VTX_A_DTYPE, vr.slots)
    sb = vr.slot_b.download(VC.VTX_B_DTYPE, vr.slots)
    keep = sb["key"] >= 0
    assert keep.sum() == nv
    assert np.array_equal(a.view(np.uint8), sa[keep].view(np.uint8))
    assert np.array_equal(b.view(np.uint8), sb[keep].view(np.uint8))
    del sa, sb
    # grid and merge records
    og = oracle.Grid()
    og.build(np.ascontiguousarray(a["pos"]), vr.radius)
    info = vr.grid.info()
    assert info.photon_count == nv and info.table_size == og.g.table_size
    assert list(info.bbox_min) == list(og.g.bbox_min) and list(info.bbox_max) == list(og.g.bbox_max)
    cell_ends = np.empty(info.table_size, np.int32)
    order = np.empty(nv, np.int32)
    N.check(N.lib.imba_memcpy_d2h(N.ptr(cell_ends), info.cell_ends, cell_ends.nbytes, None), "d2h")
    N.check(N.lib.imba_memcpy_d2h(N.ptr(order), info.order, order.nbytes, None), "d2h")
    assert np.array_equal(cell_ends, og.cell_ends())
    assert np.array_equal(order, og.order())
    rec = vr.records_host()
    assert np.array_equal(rec["out"], a["out"][order]) and np.array_equal(rec["tp"], b["tp"][order])
    assert np.array_equal(rec["dVCM"].view(np.int32), b["dVCM"][order].view(np.int32))
    assert np.array_equal(rec["dVM"].view(np.int32), b["dVM"][order].view(np.int32))
    # grid queries around cached vertices (merge queries sit on surfaces among them)
    rng = np.random.default_rng(5)
    pick = rng.choice(nv, size=20_000, replace=False)
    qpos = (a["pos"][pick] + rng.uniform(-0.5, 0.5, (20_000, 3)).astype(np.float32) * np.float32(vr.radius))
    qpos = np.ascontiguousarray(qpos, np.float32)
    qb = N.DeviceBuffer.from_array(qpos)
    oi, od, oc = N.DeviceBuffer(20_000 * 40), N.DeviceBuffer(20_000 * 40), N.DeviceBuffer(20_000 * 4)
    vr.grid.query(qb, 12, 0, 20_000, vr.k, oi, oc, od)
    g_idx = oi.download(np.int32, 20_000 * vr.k).reshape(-1, vr.k)
    g_dist = od.download(np.float32, 20_000 * vr.k).reshape(-1, vr.k)
    g_cnt = oc.download(np.int32, 20_000)
    o_idx, o_dist, o_cnt = og.query_many(qpos, vr.k)
    assert np.array_equal(g_cnt, o_cnt) and (o_cnt == vr.k).mean() > 0.2
    for r in range(20_000):
        c = g_cnt[r]
        assert np.array_equal(g_idx[r, :c], o_idx[r, :c]), r
        assert np.array_equal(g_dist[r, :c].view(np.int32), o_dist[r, :c].view(np.int32)), r
    # the image and the first light queue
    img = vr.image_host()
    assert np.all(np.isfinite(img)) and np.all(img >= 0) and (img.sum(-1) > 0).mean() > 0.5
    lr, ls = N.DeviceBuffer(vr.P * 32), N.DeviceBuffer(vr.P * VC.VCM_STATE_DTYPE.itemsize)
    N.check(N.lib.imba_vcm_light(C.byref(vr.frame_args), vr.light_seed, 0, vr.P, lr.ptr, ls.ptr, None), "vcm_light")
    rays0 = lr.download(N.RAY_DTYPE, vr.P)
    idx = _every_kth(vr.P, 65536)
    want, bad = oracle.trace(arrays, rays0[idx], threads=THREADS)
    assert bad == 0
    assert _same(vr.gpu_scene.trace_host(rays0[idx], cfg=QUANTIZED), want)
    assert vr.gpu_scene.error_count() == 0


def test_c4_batched_queues_full_size(gpu):
    print("C4: 256 queues over 8 scenes of 10K..4M triangles", flush=True)
    builds, scenes, geometry = [], [], []
    for k, n in enumerate(W.C4_SCENE_TRIS):
        b, v, i = W.random_tri_scene(n, N.BUILDER_SBVH if n <= 250_000 else N.BUILDER_BINNED_SAH, seed=0x1234 + k,
                                     threads=THREADS)
        builds.append(b)
        geometry.append((v, i))
        scenes.append(S.Scene(b.desc()))
    sizes = W.c4_queue_sizes(256)
    assert 30_000_000 < sizes.sum() < 45_000_000
    items, samples = [], []
    for q in range(256):
        n = int(sizes[q])
        v, i = geometry[q % 8]
        rays = W.c4_rays(v, i, n, coherent=q % 2 == 0, seed=5000 + q)
        hb = N.DeviceBuffer(n * 16)
        items.append((scenes[q % 8], N.DeviceBuffer.from_array(rays), hb, n, None, False))
        idx = _every_kth(n, 2048)
        samples.append((idx, rays[idx]))
    batch = S.TraceBatch(items)
    batch.launch(QUANTIZED)
    arrays = [b.arrays() for b in builds]
    for q, (sc, rb, hb, n, _, _) in enumerate(items):
        idx, rays = samples[q]
        got = hb.download(N.HIT_DTYPE, n)[idx]
        want, bad = oracle.trace(arrays[q % 8], rays, threads=THREADS)
        assert bad == 0 and _same(got, want), q
    for sc in scenes:
        assert sc.error_count() == 0


def test_c5_requests_from_the_serving_mix(grid21m):
    print("C5: requests over the 8-scene serving mix", flush=True)
    big, _ = grid21m
    scenes, centres = [], []
    for side in (1, 2, 4, 6, 8, 12, 20, 32):
        sg = big if side == 32 else W.SphereGridScene(side=side, level=5, threads=THREADS)
        scenes.append(V.ServedScene(sg))
        centres.append(W.served_view(side))
    reqs = V.poisson_requests(10, 200.0, 16, len(scenes), centres, seed=11, integrators=("pt", "ppm"),
                              resolutions=((256, 256), (640, 480)), ceiling=W.SERVED_CEILING)
    # every scene size class is represented: force the first requests onto the largest scenes
    for r, sc in zip(reqs, (7, 6, 5)):
        r.scene = sc
        cx, cy, cz, radius = centres[sc]
        r.cam_pos = (cx, min(cy + 0.6 * radius, W.SERVED_CEILING), cz - radius)
        r.cam_dir = (0.0, cy - r.cam_pos[1], radius)

    def run(batched):
        srv = V.Server(scenes, 24 << 30, max_active=16, batched=batched)
        return srv.run([V.Request(**{**r.__dict__}) for r in reqs], keep_images=True)

    done_mix, img_mix = run(True)
    done_alone, img_alone = run(False)
    assert len(done_mix) == len(done_alone) == len(reqs)
    by_mix = {r.rid: r for r in done_mix}
    for r in done_alone:
        m = by_mix[r.rid]
        assert m.counts == r.counts and m.rays == r.rays
        a, b = img_alone[r.rid], img_mix[r.rid]
        assert np.allclose(a, b, rtol=1e-5, atol=1e-6)
        on = np.abs(a) > 1e-6
        if r.integrator == "pt":
            assert on.mean() > 0.05, r.rid  # next-event estimation lights the visible surfaces
        if on.any():
            assert np.mean(np.abs(a[on] - b[on]) / np.abs(a[on])) <= 1e-4
