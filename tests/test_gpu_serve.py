"""The request loop (imbatracer_amd.serve): a request's result is the same alone and in a mix
of concurrent requests sharing batched launches (GPU)."""
import numpy as np
import pytest

from imbatracer_amd import serve as V
from imbatracer_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scenes():
    out = []
    for side, level in ((2, 2), (3, 3)):
        out.append(V.ServedScene(W.SphereGridScene(side=side, level=level, threads=4)))
    return out


def _requests():
    res = [(64, 48), (32, 32), (80, 60), (48, 48), (64, 64), (40, 30)]
    reqs = []
    for k, (w, h) in enumerate(res):
        sc = k % 2
        side = 2 + sc
        c = (side * 0.5, 0.0, side * 0.5)
        pos = (c[0] + 0.3 * k, 3.0 + 0.2 * k, c[2] - 6.0)
        reqs.append(V.Request(k, k % 3, sc, w, h, 1 + k % 2, pos, (c[0] - pos[0], -pos[1], c[2] - pos[2]), 1000 + 17 * k,
                              0.0))
    return reqs


def test_alone_and_in_a_mix_give_the_same_frames(gpu, scenes):
    mix = V.Server(scenes, 256 << 20, max_active=8, batched=True)
    done_mix, img_mix = mix.run(_requests(), keep_images=True)
    alone = V.Server(scenes, 256 << 20, batched=False)
    done_alone, img_alone = alone.run(_requests(), keep_images=True)
    assert len(done_mix) == len(done_alone) == 6
    by_mix = {r.rid: r for r in done_mix}
    for r in done_alone:
        m = by_mix[r.rid]
        assert m.counts == r.counts and m.rays == r.rays and r.counts[0] == r.samples
        a, b = img_alone[r.rid], img_mix[r.rid]
        assert np.allclose(a, b, rtol=1e-5, atol=1e-6)
        lit = np.abs(a) > 1e-6
        if lit.any():
            assert np.mean(np.abs(a[lit] - b[lit]) / np.abs(a[lit])) <= 1e-4
    # the requests really shared steps: fewer steps than one request at a time
    assert mix.steps < alone.steps
    assert mix.arena.free == [(0, mix.arena.size)] and alone.arena.free == [(0, alone.arena.size)]


def test_admission_waits_for_memory(gpu, scenes):
    reqs = _requests()
    small = V.Server(scenes, max(V.footprint(r) for r in reqs) + 1024, max_active=8, batched=True)
    done = small.run(reqs)
    assert len(done) == 6
    # with room for about one request at a time, admissions never overlapped in memory
    spans = sorted((r.admitted, r.done) for r in done)
    for (a0, d0), (a1, d1) in zip(spans, spans[1:]):
        assert a1 >= d0 - 1e-9 or V.footprint(reqs[0]) * 2 <= small.arena.size


def _mixed_requests():
    reqs = _requests()
    for k, r in enumerate(reqs):
        r.integrator = "ppm" if k % 2 == 0 else "pt"
    return reqs


def test_photon_mapping_requests_alone_and_in_a_mix(gpu, scenes):
    """Photon-mapping (PPM) requests served beside PT requests: each request's counts and image
    are the same alone and in the mix."""
    mix = V.Server(scenes, 512 << 20, max_active=8, batched=True)
    done_mix, img_mix = mix.run(_mixed_requests(), keep_images=True)
    alone = V.Server(scenes, 512 << 20, batched=False)
    done_alone, img_alone = alone.run(_mixed_requests(), keep_images=True)
    assert len(done_mix) == len(done_alone) == 6
    by_mix = {r.rid: r for r in done_mix}
    n_ppm_lit = 0
    for r in done_alone:
        m = by_mix[r.rid]
        assert m.counts == r.counts and m.rays == r.rays
        first = r.width * r.height if r.integrator == "ppm" else r.samples
        assert r.counts[0] == first
        a, b = img_alone[r.rid], img_mix[r.rid]
        assert np.allclose(a, b, rtol=1e-5, atol=1e-6)
        lit = np.abs(a) > 1e-6
        if lit.any():
            assert np.mean(np.abs(a[lit] - b[lit]) / np.abs(a[lit])) <= 1e-4
            n_ppm_lit += r.integrator == "ppm"
    assert n_ppm_lit >= 2  # the photon-mapping images received photons
    assert mix.arena.free == [(0, mix.arena.size)]


def _vcm_requests():
    reqs = _requests()
    for k, r in enumerate(reqs):
        r.integrator = "vcm" if k % 2 == 0 else "pt"
    return reqs


def test_vcm_requests_alone_and_in_a_mix(gpu, scenes):
    """VCM requests served beside PT requests: each request's counts and image are the same
    alone and in the mix."""
    mix = V.Server(scenes, 512 << 20, max_active=8, batched=True)
    done_mix, img_mix = mix.run(_vcm_requests(), keep_images=True)
    alone = V.Server(scenes, 512 << 20, batched=False)
    done_alone, img_alone = alone.run(_vcm_requests(), keep_images=True)
    assert len(done_mix) == len(done_alone) == 6
    by_mix = {r.rid: r for r in done_mix}
    n_vcm_lit = 0
    for r in done_alone:
        m = by_mix[r.rid]
        assert m.counts == r.counts and m.rays == r.rays
        first = r.width * r.height if r.integrator == "vcm" else r.samples
        assert r.counts[0] == first
        a, b = img_alone[r.rid], img_mix[r.rid]
        assert np.allclose(a, b, rtol=1e-5, atol=1e-6)
        lit = np.abs(a) > 1e-6
        if lit.any():
            assert np.mean(np.abs(a[lit] - b[lit]) / np.abs(a[lit])) <= 1e-4
            n_vcm_lit += r.integrator == "vcm"
    assert n_vcm_lit >= 2
    assert mix.arena.free == [(0, mix.arena.size)]


def test_served_vcm_request_is_the_vcm_iteration(gpu, scenes):
    """A served VCM request renders the iteration imbatracer_amd.vcm.VcmRenderer renders with
    the same camera, seeds and settings (whose steps tests/test_gpu_vcm.py restates)."""
    from imbatracer_amd import vcm as VC

    req = _vcm_requests()[2]
    sc = scenes[req.scene]
    srv = V.Server(scenes, 512 << 20, batched=False)
    done, img = srv.run([req], keep_images=True)
    vr = VC.VcmRenderer(sc.s, req.width, req.height, light_paths=req.width * req.height, spp=req.spp,
                        max_path_len=V.VCM_DEPTH, light_max_depth=V.VCM_DEPTH, rr_start=V.VCM_RR_START,
                        radius_factor=0.005, k=V.VCM_K, num_connections=V.VCM_C, seed_base=req.seed,
                        light_seed=(req.seed ^ 0x5A5A5A5A) & 0xFFFFFFFF, gpu_scene=sc.gpu,
                        camera=(req.cam_pos, req.cam_dir, (0.0, 1.0, 0.0), 60.0))
    assert vr.radius == sc.radius
    light, cam, nv = vr.frame()
    assert done[0].counts[:len(light)] == light
    want = vr.image_host().reshape(-1)
    got = img[req.rid]
    lit = np.abs(want) > 1e-6
    assert lit.mean() > 0.2
    assert np.mean(np.abs(got[lit] - want[lit]) / np.abs(want[lit])) <= 1e-4


def test_served_pt_and_ppm_requests_are_the_renderers_frames(gpu, scenes):
    """A served PT request renders the frame workloads.PathTracer renders, and a served
    photon-mapping request the frame workloads.PhotonMapper renders, with the same camera,
    seeds and settings (both renderers' steps are restated against the oracle in
    tests/test_gpu_pt_image.py and tests/test_gpu_photon_bounces.py)."""
    reqs = _mixed_requests()
    pt_req = next(r for r in reqs if r.integrator == "pt")
    ppm_req = next(r for r in reqs if r.integrator == "ppm")
    srv = V.Server(scenes, 512 << 20, batched=False)
    done, img = srv.run([pt_req, ppm_req], keep_images=True)
    assert len(done) == 2
    for req in (pt_req, ppm_req):
        sc = scenes[req.scene]
        cam = (req.cam_pos, req.cam_dir, (0.0, 1.0, 0.0), 60.0)
        if req.integrator == "pt":
            r = W.PathTracer(sc.s, req.width, req.height, spp=req.spp, max_path_len=srv.max_path_len,
                             rr_start=srv.rr_start, seed_base=req.seed, gpu_scene=sc.gpu, camera=cam)
        else:
            r = W.PhotonMapper(sc.s, req.width, req.height, light_paths=req.width * req.height, max_depth=V.PPM_DEPTH,
                               rr_start=srv.rr_start, radius_factor=0.005, k=V.PPM_K, spp=req.spp, seed_base=req.seed,
                               light_seed=(req.seed ^ 0x5A5A5A5A) & 0xFFFFFFFF, camera=cam)
            assert r.radius == sc.radius
        r.frame()
        r.stream.sync()
        want = r.image.download(np.float32, req.width * req.height * 3)
        got = img[req.rid]
        lit = np.abs(want) > 1e-6
        assert lit.sum() > 50, req.integrator
        assert np.mean(np.abs(got[lit] - want[lit]) / np.abs(want[lit])) <= 1e-4, req.integrator
