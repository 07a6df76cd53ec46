"""Vertex connection and merging (VCMIntegrator<algo>, render/integrators/vcm.cpp) on the
GPU over a small sphere grid with mirror and glass spheres.

  * the light vertex cache is the stable compaction of every light bounce's vertex slots;
  * the merge records are the cache's entries in the grid's slot order;
  * every camera bounce's merging (the fused grid query + merge) equals vertex_merging
    (vcm.cpp:572-611) restated in numpy over the oracle grid's k-nearest lists, within 1e-5;
  * the frame's image is finite, non-negative and lit, and a second frame repeats the first;
  * every light and camera bounce is restated by oracle/vcm.py on the GPU's own queues, for
    VCM and for the sub-algorithms BPT, LT and the VCM integrator's PT.
"""
import numpy as np
import pytest

import oracle
from imbatracer_amd import native as N
from imbatracer_amd import vcm as V
from imbatracer_amd import workloads as W

pytestmark = pytest.mark.gpu

F = np.float32


@pytest.fixture(scope="module")
def scene(gpu):
    return W.SphereGridScene(side=4, level=2, builder=N.BUILDER_SBVH, specular_fraction=0.3)


def _render(scene, algo):
    r = V.VcmRenderer(scene, 48, 32, light_paths=48 * 32, spp=1, max_path_len=6, light_max_depth=6, rr_start=0,
                      radius_factor=0.01, k=10, algo=algo)
    r.keep_queues = True
    print("%s frame 48x32, %d triangles" % (algo.upper(), scene.tri_count), flush=True)
    r.counts_frame = r.frame()
    return r


@pytest.fixture(scope="module")
def vr(scene):
    return _render(scene, "vcm")


@pytest.fixture(scope="module", params=["bpt", "lt", "pt"])
def mode(scene, request):
    return _render(scene, request.param)


def test_frame_runs(vr):
    light, cam, nv = vr.counts_frame
    assert light[0] == vr.P and cam[0] == vr.n_cam
    assert 0 < nv <= vr.L * vr.P
    img = vr.image_host()
    assert np.all(np.isfinite(img)) and np.all(img >= 0)
    assert (img.sum(-1) > 0).mean() > 0.5
    assert vr.gpu_scene.error_count() == 0


def test_cache_is_the_compacted_slots(vr):
    a, b = vr.cache_host()
    want_a, want_b = [], []
    for k in vr.kept["light"]:
        keep = k["vtx_b"]["key"] >= 0
        want_a.append(k["vtx_a"][keep])
        want_b.append(k["vtx_b"][keep])
    want_a, want_b = np.concatenate(want_a), np.concatenate(want_b)
    assert len(a) == len(want_a) == vr.vertex_count
    assert np.array_equal(a.view(np.uint8), want_a.view(np.uint8))
    assert np.array_equal(b.view(np.uint8), want_b.view(np.uint8))
    # vertices are stored on Lambertian hits only, with path lengths 2 .. light_max_depth + 1
    assert np.all(vr.s.kinds[b["mat"]] == 0)
    assert a["path_length"].min() >= 2 and a["path_length"].max() <= vr.L + 1


def test_merge_records_in_grid_order(vr):
    a, b = vr.cache_host()
    rec = vr.records_host()
    order = np.empty(vr.vertex_count, np.int32)
    N.check(N.lib.imba_memcpy_d2h(N.ptr(order), vr.grid.info().order, order.nbytes, None), "d2h")
    og = oracle.Grid()
    og.build(np.ascontiguousarray(a["pos"]), vr.radius)
    assert np.array_equal(order, og.order())
    assert np.array_equal(rec["out"], a["out"][order])
    assert np.array_equal(rec["tp"], b["tp"][order])
    assert np.array_equal(rec["dVCM"].view(np.int32), b["dVCM"][order].view(np.int32))
    assert np.array_equal(rec["dVM"].view(np.int32), b["dVM"][order].view(np.int32))


def _dot(x, y):
    return F(F(F(x[0] * y[0]) + F(x[1] * y[1])) + F(x[2] * y[2]))


def _merge_restated(vr, q, og, a, b):
    """vertex_merging for every query of q, in float32 with the reference's operation order;
    returns the per-pixel sums (float64 accumulation of the per-query contributions)."""
    img = np.zeros((vr.w * vr.h, 3))
    valid = np.nonzero(q["pixel"] >= 0)[0]
    if len(valid) == 0:
        return img, 0
    idx, dist, cnt = og.query_many(np.ascontiguousarray(q["pos"][valid]), vr.k)
    pi = F(3.14159265359)
    r = F(vr.radius)
    eta_vcm = pi * (r * r) * F(vr.P)
    eta_vc = F(1.0) / eta_vcm
    albedo = np.asarray(vr.s.albedo, F)
    merged = 0
    for row, j in enumerate(valid):
        c = cnt[row]
        if c == 0:
            continue
        n, out = q["n"][j], q["out"][j]
        rs = dist[row, vr.k - 1] if c == vr.k else r * r
        alb = albedo[q["mat"][j]] * (F(1.0) / pi)
        acc = np.zeros(3, F)
        for l in range(c):
            p = idx[row, l]
            pin = a["out"][p]
            cn_in, cn_out = _dot(n, pin), _dot(n, out)
            if not (cn_in * cn_out > 0):
                continue
            pdf_dir = F(abs(cn_in)) / pi
            pdf_rev = F(abs(cn_out)) / pi
            wl = b["dVCM"][p] * eta_vc + b["dVM"][p] * pdf_dir
            wc = q["dVCM"][j] * eta_vc + q["dVM"][j] * pdf_rev
            mis = F(1.0) / (wl + F(1.0) + wc)
            kern = F(1.0) - dist[row, l] / rs
            acc = acc + mis * alb * kern * b["tp"][p]
        scale = F(2.0) / (pi * rs * F(vr.P))
        img[q["pixel"][j]] += q["tp"][j] * (acc * scale)
        merged += 1
    return img, merged


def test_merging_matches_restatement(vr):
    a, b = vr.cache_host()
    og = oracle.Grid()
    og.build(np.ascontiguousarray(a["pos"]), vr.radius)
    total = 0
    for bounce, k in enumerate(vr.kept["camera"]):
        got = (k["image_merge"].astype(np.float64) - k["image_shadow"]).reshape(-1, 3)
        want, merged = _merge_restated(vr, k["queries"], og, a, b)
        total += merged
        scale = max(np.abs(want).max(), 1e-30)
        assert np.allclose(got, want, rtol=1e-4, atol=1e-5 * scale), bounce
    assert total > 100


# ---- every step restated (oracle/vcm.py) on the GPU's own queues ----

def _params(vr):
    from oracle import vcm as OV

    f = vr.frame_args
    return OV.Params(vr.s, vr.r2w, vr.w2r, vr.cam_pos, vr.fwd, vr.ipd, vr.w, vr.h, vr.spp, vr.P, vr.radius,
                     f.max_path_len, f.light_max_depth, f.rr_start, f.num_connections, algo=vr.algo)


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


def _close(a, b, rtol=1e-5):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.allclose(a, b, rtol=rtol, atol=rtol * max(1e-30, np.abs(b).max(initial=0.0)))


def _check_states(got_rays, got_st, want, exact_continuation):
    """The bounce's outcome: which paths continue, their random-number states and path lengths
    and ray origins bit for bit; directions, throughputs and partial weights within 1e-5 (the
    direction's sin / cos are the device's)."""
    alive = want["alive"]
    assert np.array_equal(got_st["pixel"] >= 0, alive)
    ws = want["states"]
    assert np.array_equal(got_st["rng"], ws["rng"])
    assert np.array_equal(got_st["path_length"][alive], ws["path_length"][alive])
    assert np.array_equal(_bits(got_rays["org"][alive]), _bits(want["rays"]["org"][alive]))
    assert np.allclose(got_rays["dir"][alive, :3], want["rays"]["dir"][alive, :3], atol=1e-5)
    for f in ("dVC", "dVM", "dVCM"):
        assert _close(got_st[f][alive], ws[f][alive]), f
    assert _close(got_st["tp"][alive], ws["tp"][alive])
    if exact_continuation is not None:
        assert np.array_equal(_bits(got_st["dVCM"][exact_continuation]), _bits(ws["dVCM"][exact_continuation]))


def test_light_start_restated(vr):
    from oracle import vcm as OV

    p = _params(vr)
    rays, st = OV.light_start(p, 0, vr.P, vr.light_seed)
    got = vr.kept["light"][0]
    assert np.array_equal(_bits(got["rays"]["org"]), _bits(rays["org"]))
    assert np.allclose(got["rays"]["dir"][:, :3], rays["dir"][:, :3], atol=1e-5)
    for f in ("pixel", "rng", "path_length"):
        assert np.array_equal(got["states"][f], st[f]), f
    for f in ("tp", "dVCM", "dVC", "dVM"):
        assert np.array_equal(_bits(got["states"][f]), _bits(st[f])), f


def test_light_bounces_restated(vr):
    _light_bounces_restated(vr)


def _light_bounces_restated(vr):
    from oracle import vcm as OV

    p = _params(vr)
    conns = 0
    for b, k in enumerate(vr.kept["light"]):
        w = OV.light_bounce(p, k["rays"], k["hits"], k["states"])
        # the light vertex and the camera connection depend only on the step's inputs
        # (an empty slot holds only its key)
        assert np.array_equal(k["vtx_b"]["key"], w["vtx_b"]["key"]), b
        live = w["vtx_b"]["key"] >= 0
        assert np.array_equal(_bits(k["vtx_b"][live]), _bits(w["vtx_b"][live])), b
        assert np.array_equal(_bits(k["vtx_a"][live]), _bits(w["vtx_a"][live])), b
        assert np.array_equal(_bits(k["shadow_states"]), _bits(w["shadow_states"])), b
        assert np.array_equal(_bits(k["shadow_rays"]), _bits(w["shadow_rays"])), b
        conns += int((w["shadow_states"]["pixel"] >= 0).sum())
        _check_states(k["out_rays"], k["out_states"], w, None)
        # the camera connections' occlusion answers are the oracle's
        occ = w["shadow_states"]["pixel"] >= 0
        want, bad = oracle.trace(vr.s.build.arrays(), w["shadow_rays"][occ], any_hit=True)
        assert bad == 0
        assert np.array_equal(k["shadow_hits"]["tri_id"][occ] < 0, want["tri_id"] < 0), b
    assert conns > 100
    return len(vr.kept["light"])


def test_camera_start_restated(vr):
    from oracle import vcm as OV

    p = _params(vr)
    rays, st = OV.camera_start(p, 0, vr.n_cam, vr.seed_base)
    got = vr.kept["camera"][0]
    assert np.array_equal(_bits(got["rays"]), _bits(rays))
    for f in ("pixel", "sample", "rng", "path_length"):
        assert np.array_equal(got["states"][f], st[f]), f
    for f in ("tp", "dVCM", "dVC", "dVM"):
        assert np.array_equal(_bits(got["states"][f]), _bits(st[f])), f


def test_camera_bounces_restated(vr):
    counts = _camera_bounces_restated(vr)
    assert all(v > 0 for v in counts.values()), counts


def _camera_bounces_restated(vr):
    from oracle import vcm as OV

    p = _params(vr)
    a, b_ = vr.cache_host()
    counts = {"nee": 0, "conn": 0, "query": 0, "emit": 0}
    C = vr.C
    for b, k in enumerate(vr.kept["camera"]):
        w = OV.camera_bounce(p, k["rays"], k["hits"], k["states"], a, b_)
        assert np.array_equal(_bits(k["shadow_states"]), _bits(w["shadow_states"])), b
        assert np.array_equal(_bits(k["shadow_rays"]), _bits(w["shadow_rays"])), b
        # (an empty query slot holds only its position, +inf, and pixel -1)
        q, wq = k["queries"], w["queries"]
        assert np.array_equal(_bits(q["pos"]), _bits(wq["pos"])) and np.array_equal(q["pixel"], wq["pixel"]), b
        used = wq["pixel"] >= 0
        assert np.array_equal(_bits(q[used]), _bits(wq[used])), b
        _check_states(k["out_rays"], k["out_states"], w, None)
        ss = w["shadow_states"]["pixel"]
        counts["nee"] += int((ss[0::1 + C] >= 0).sum())
        counts["conn"] += int((ss.reshape(-1, 1 + C)[:, 1:] >= 0).sum())
        counts["query"] += int((w["queries"]["pixel"] >= 0).sum())
        counts["emit"] += len(w["emit_pixel"])
        # emitter contributions of this step
        want = np.zeros((vr.w * vr.h, 3))
        np.add.at(want, w["emit_pixel"], w["emit_rgb"].astype(np.float64))
        got = (k["image_emit"].astype(np.float64) - k["image_before"]).reshape(-1, 3)
        assert np.allclose(got, want, rtol=1e-4, atol=1e-6 * max(1e-30, np.abs(want).max(initial=0))), b
        # occlusion answers of the shadow rays are the oracle's; unoccluded ones add their share
        occ = ss >= 0
        wt, bad = oracle.trace(vr.s.build.arrays(), w["shadow_rays"][occ], any_hit=True)
        assert bad == 0
        assert np.array_equal(k["shadow_hits"]["tri_id"][occ] < 0, wt["tri_id"] < 0), b
        vis = np.nonzero(occ)[0][wt["tri_id"] < 0]
        want = np.zeros((vr.w * vr.h, 3))
        np.add.at(want, ss[vis], w["shadow_states"]["tp"][vis].astype(np.float64))
        got = (k["image_shadow"].astype(np.float64) - k["image_emit"]).reshape(-1, 3)
        assert np.allclose(got, want, rtol=1e-4, atol=1e-5 * max(1e-30, np.abs(want).max(initial=0))), b
    return counts


def test_frame_repeats(vr):
    img = vr.image_host().copy()
    a0 = vr.cache_host()[0].copy()
    vr.keep_queues = False
    vr.clear_image()
    vr.frame()
    a1 = vr.cache_host()[0]
    assert np.array_equal(a0.view(np.uint8), a1.view(np.uint8))
    img2 = vr.image_host()
    on = img > 1e-6
    assert np.mean(np.abs(img2[on] - img[on]) / img[on]) <= 1e-4


# ---- the sub-algorithms: the same steps with techniques switched off ----

def test_mode_frame(mode):
    light, cam, nv = mode.counts_frame
    assert (len(light) == 0) == (mode.algo == "pt")
    assert (len(cam) == 0) == (mode.algo == "lt")
    assert (nv > 0) == (mode.algo == "bpt")  # the cache is built for connections only
    img = mode.image_host()
    assert np.all(np.isfinite(img)) and np.all(img >= 0)
    assert (img.sum(-1) > 0).mean() > 0.3


def test_mode_light_bounces_restated(mode):
    if mode.algo == "pt":
        assert mode.kept["light"] == []
        return
    assert _light_bounces_restated(mode) > 0
    if mode.algo == "lt":  # light tracing stores no vertices
        assert all(np.all(k["vtx_b"]["key"] == -1) for k in mode.kept["light"])


def test_mode_camera_bounces_restated(mode):
    if mode.algo == "lt":
        assert mode.kept["camera"] == []
        return
    counts = _camera_bounces_restated(mode)
    assert counts["nee"] > 0 and counts["emit"] > 0 and counts["query"] == 0
    assert (counts["conn"] > 0) == (mode.algo == "bpt"), counts
