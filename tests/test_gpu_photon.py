"""Photon-mapping frame on the GPU (light paths -> photons -> hash grid -> camera queries ->
merging) against CPU restatements on a small sphere-grid scene with mirror and glass spheres.

Pins, bit for bit: the photons deposited by the first hit of every light path (traversal by the
oracle, deposit rule and slot order restated in numpy), the merge queries written at the first
hit of camera paths, the grid built over the GPU's photons and the k-nearest results of every
query (oracle grid).  The merge kernel's image against a numpy restatement within 1e-5
relative.  The light sampling (seeding of LightRayGen, render/ray_gen/ray_gen.h:138-145, and
TriangleLight::sample_emit, render/light.h:109-141) within 1e-5, since sqrt and sincos are the
device's.
"""
import numpy as np
import pytest

import oracle
from imbatracer_amd import native as N
from imbatracer_amd import workloads as W

pytestmark = pytest.mark.gpu

F32_INF_BITS = 0x7F800000


@pytest.fixture(scope="module")
def pm(gpu):
    scene = W.SphereGridScene(side=4, level=2, builder=N.BUILDER_SBVH, specular_fraction=0.5)
    assert (scene.kinds == 1).any() and (scene.kinds == 2).any()
    m = W.PhotonMapper(scene, 64, 48, light_paths=30000, max_depth=6, rr_start=2, radius_factor=0.02, k=10)
    m.clear_image()
    light, cam, n_ph = m.frame()
    assert n_ph > 0 and light[0] == m.P and cam[0] == m.n_cam
    return m


def _mwc(seed_base, i, pixel, draws):
    seed = seed_base & 0xFFFFFFFF
    for _ in range(4):
        seed = ((33 * seed) & 0xFFFFFFFF) ^ i
    sseed = seed - (1 << 32) if seed >= 1 << 31 else seed
    s = sseed & 0xFFFFFFFFFFFFFFFF
    skip = int(np.fmod(sseed, 5)) + 16 + pixel % 5
    out = []
    for k in range(skip + draws):
        c, x = s >> 32, s & 0xFFFFFFFF
        s = (x * 4294883355 + c) & 0xFFFFFFFFFFFFFFFF
        if k >= skip:
            out.append(np.float32(np.float32(x ^ c) / np.float32(4294967296.0)))
    return out


def _first_rays(m, light: bool):
    n = m.P if light else m.n_cam
    rays, states = N.DeviceBuffer(n * 32), N.DeviceBuffer(n * 32)
    if light:
        N.check(N.lib.imba_ppm_light(N.ptr(m.light_tri), N.ptr(m.light_int), m.light_seed, 0, n, rays.ptr,
                                     states.ptr, None), "ppm_light")
    else:
        q = N.DeviceBuffer(n * 64)
        N.check(N.lib.imba_ppm_camera(N.ptr(m.r2w), N.ptr(m.cam_pos), m.w, m.h, m.spp, m.seed_base, 0, n, rays.ptr,
                                      states.ptr, q.ptr, None), "ppm_camera")
        qh = q.download(W.QUERY_DTYPE, n)
        assert np.all(qh["pixel"] == -1) and np.all(np.isinf(qh["pos"]))
    st = np.dtype([("pixel", "<i4"), ("sample", "<i4"), ("rng", "<u8"), ("tp", "<f4", 3), ("bounces", "<i4")])
    return rays.download(N.RAY_DTYPE, n), states.download(st, n)


def test_light_sampling(pm):
    rays, states = _first_rays(pm, True)
    v = pm.light_tri.reshape(3, 3).astype(np.float64)
    e1, e2 = v[1] - v[0], v[2] - v[0]
    nrm = np.cross(e1, e2)
    area = 0.5 * np.linalg.norm(nrm)
    nrm /= np.linalg.norm(nrm)
    assert np.array_equal(states["pixel"], np.arange(pm.P))
    for i in [0, 1, 7, 123, pm.P - 1]:
        r1, r2, u1, u2 = _mwc(pm.light_seed, i, i, 4)
        su = np.sqrt(np.float64(r1))
        u, w = 1 - su, r2 * su
        pos = u * v[0] + w * v[1] + (1 - u - w) * v[2]
        assert np.allclose(rays["org"][i, :3], pos, rtol=1e-5, atol=1e-5)
        assert rays["org"][i, 3] == np.float32(1e-3)
        d = rays["dir"][i, :3].astype(np.float64)
        assert abs(np.linalg.norm(d) - 1) < 1e-5
        assert abs(d @ nrm - np.sqrt(u2)) < 1e-5          # cos theta = sqrt(u2)
    tp = np.float32(pm.light_int[0]) * np.float32(np.float32(area) * np.float32(np.pi))
    assert np.allclose(states["tp"][:, 0], tp, rtol=1e-5)
    assert np.all(rays["dir"][:, :3] @ nrm.astype(np.float32) >= 0)


def _hits(m, rays):
    return oracle.trace(m.s.build.arrays(), rays, threads=8)[0]


def test_first_bounce_photons(pm):
    rays, states = _first_rays(pm, True)
    hits = _hits(pm, rays)
    ph_rays, ph_states = pm.photons_host()
    first = ph_rays["length"] == 2
    # bounce order: the first-hit photons come first, by path
    assert np.all(np.diff(ph_rays["length"]) >= 0)
    got_r, got_s = ph_rays[first], ph_states[first]
    s = pm.s
    hit = hits["tri_id"] >= 0
    idx = np.nonzero(hit)[0]
    tri = hits["tri_id"][idx]
    kind = s.kinds[s.tri_materials[tri]]
    t = hits["tmax"][idx]
    o, d = rays["org"][idx, :3], rays["dir"][idx, :3]
    cos_o = -(d[:, 0] * s.tri_normals[tri, 0] + d[:, 1] * s.tri_normals[tri, 1] + d[:, 2] * s.tri_normals[tri, 2])
    keep = (kind == 0) & (cos_o != 0)
    idx, t, o, d = idx[keep], t[keep], o[keep], d[keep]
    assert len(got_r) == len(idx) > 0
    assert np.array_equal(got_s["path"], idx)
    pos = o + t[:, None] * d
    assert np.array_equal(got_r["pos"], pos)
    assert np.array_equal(got_r["out_dir"], -d)
    assert np.array_equal(got_s["tp"], states["tp"][idx])
    # later bounces: path ids increase within each bounce
    for L in np.unique(ph_rays["length"]):
        p = ph_states["path"][ph_rays["length"] == L]
        assert np.all(np.diff(p) > 0)
        assert p.min() >= 0 and p.max() < pm.P


def test_first_hit_queries(pm):
    rays, _ = _first_rays(pm, False)
    hits = _hits(pm, rays)
    q = pm.queries_host()
    s = pm.s
    tri = hits["tri_id"]
    hit = tri >= 0
    mat = np.where(hit, s.tri_materials[np.maximum(tri, 0)], 0)
    emissive = (s.emission[mat] > 0).any(1)
    diffuse = hit & (s.kinds[mat] == 0) & ~emissive
    d = rays["dir"][:, :3]
    nrm = s.tri_normals[np.maximum(tri, 0)]
    cos_o = -(d[:, 0] * nrm[:, 0] + d[:, 1] * nrm[:, 1] + d[:, 2] * nrm[:, 2])
    diffuse &= cos_o != 0
    # misses and emitters end with an empty query
    ended = ~hit | (hit & emissive)
    assert np.all(q["pixel"][ended] == -1)
    assert np.all(np.ascontiguousarray(q["pos"][ended]).view(np.int32) == F32_INF_BITS)
    assert np.all(np.isfinite(rays["dir"][:, :3]))
    i = np.nonzero(diffuse)[0]
    assert len(i) > 0.5 * len(rays)
    assert np.array_equal(q["pixel"][i], i)
    assert np.array_equal(q["pos"][i], rays["org"][i, :3] + hits["tmax"][i, None] * d[i])
    assert np.array_equal(q["normal"][i], nrm[i])
    assert np.array_equal(q["out_dir"][i], -d[i])
    c = np.float32(1.0) / np.float32(3.14159265359)
    assert np.array_equal(q["weight"][i], s.albedo[mat[i]] * c)
    # specular first hits continued: their query is a later diffuse hit or empty
    spec = hit & ~emissive & (s.kinds[mat] != 0)
    assert spec.any()
    assert np.all((q["pixel"][spec] == np.nonzero(spec)[0]) | (q["pixel"][spec] == -1))


def test_grid_and_queries_match_oracle(pm):
    ph_rays, _ = pm.photons_host()
    pos = np.ascontiguousarray(ph_rays["pos"])
    og = oracle.Grid()
    og.build(pos, pm.radius)
    info = pm.grid.info()
    assert info.table_size == og.g.table_size and info.photon_count == len(pos)
    q = pm.queries_host()
    idx, dist, cnt = pm.results_host()
    qpos = np.ascontiguousarray(q["pos"])
    assert (cnt > 0).sum() > 0.5 * (q["pixel"] >= 0).sum()
    for j in range(len(q)):
        oi, od = og.query(qpos[j], pm.k)
        assert cnt[j] == len(oi)
        assert np.array_equal(idx[j, :cnt[j]], oi)
        assert np.array_equal(dist[j, :cnt[j]].view(np.int32), od.view(np.int32))


def test_merge_image(pm):
    ph_rays, ph_states = pm.photons_host()
    q = pm.queries_host()
    idx, dist, cnt = pm.results_host()
    pm.clear_image()
    pm.merge()
    pm.stream.sync()
    got = pm.image.download(np.float32, pm.w * pm.h * 3).reshape(-1, 3)
    want = np.zeros_like(got, dtype=np.float64)
    r2 = np.float32(pm.radius) * np.float32(pm.radius)
    for j in np.nonzero(cnt > 0)[0]:
        c = cnt[j]
        rs = dist[j, pm.k - 1] if c == pm.k else r2
        if not rs > 0:
            continue
        n, o = q["normal"][j].astype(np.float64), q["out_dir"][j].astype(np.float64)
        cos_c = n @ o
        acc = np.zeros(3)
        for l in range(c):
            p = idx[j, l]
            if (n @ ph_rays["out_dir"][p]) * cos_c > 0:
                acc += (1.0 - dist[j, l] / rs) * ph_states["tp"][p]
        want[q["pixel"][j]] += q["weight"][j] * acc * 2.0 / (np.pi * rs * pm.P)
    assert np.count_nonzero(want) > 0
    assert np.allclose(got, want, rtol=1e-5, atol=1e-6 * np.abs(want).max())


def test_fused_merge(pm):
    """imba_ppm_merge_records gathers every photon's direction and state into grid slot order,
    and imba_ppm_merge_grid (query and merge in one launch, reading those records) adds the
    image of imba_hashgrid_query + imba_ppm_merge: the same lists and terms, only the running
    sum's rounding may differ."""
    n = pm.photon_count
    rec = N.DeviceBuffer(n * 32)
    N.check(N.lib.imba_ppm_merge_records(pm.grid.h, pm.photon_rays.ptr, pm.photon_states.ptr, rec.ptr, None),
            "ppm_merge_records")
    order = np.empty(n, np.int32)
    N.check(N.lib.imba_memcpy_d2h(N.ptr(order), pm.grid.info().order, order.nbytes, None), "d2h")
    ph_rays, ph_states = pm.photons_host()
    got = rec.download(np.uint8, n * 32).reshape(n, 32)
    assert np.array_equal(got[:, :16], ph_rays.view(np.uint8).reshape(n, 32)[order, 16:])
    assert np.array_equal(got[:, 16:], ph_states.view(np.uint8).reshape(n, 16)[order])
    pm.clear_image()
    pm.merge()
    lists = pm.image.download(np.float32, pm.w * pm.h * 3, pm.stream).astype(np.float64)
    pm.clear_image()
    N.check(N.lib.imba_ppm_merge_grid(pm.grid.h, rec.ptr, pm.queries.ptr, pm.n_cam, pm.k, pm.radius, pm.P,
                                      pm.image.ptr, pm.stream.handle), "ppm_merge_grid")
    fused = pm.image.download(np.float32, pm.w * pm.h * 3, pm.stream).astype(np.float64)
    assert np.count_nonzero(lists) > 0
    assert np.allclose(fused, lists, rtol=1e-5, atol=1e-6 * np.abs(lists).max())


def test_fused_frame(pm):
    """A frame with merge="fused" renders the image of the query + merge frame."""
    pm.clear_image()
    pm.frame()
    want = pm.image.download(np.float32, pm.w * pm.h * 3, pm.stream).astype(np.float64)
    f = W.PhotonMapper(pm.s, pm.w, pm.h, light_paths=pm.P, max_depth=pm.L, rr_start=pm.rr_start,
                       radius_factor=0.02, k=pm.k, merge="fused")
    assert f.frame()[2] == pm.photon_count
    got = f.image.download(np.float32, f.w * f.h * 3, f.stream).astype(np.float64)
    assert np.allclose(got, want, rtol=1e-5, atol=1e-6 * np.abs(want).max())


def test_frames_repeat(pm):
    """A second frame reproduces the photon set and the query results exactly."""
    a = pm.photons_host()[0].copy()
    r = pm.results_host()
    pm.frame()
    b = pm.photons_host()[0]
    assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    r2 = pm.results_host()
    for x, y in zip(r, r2):
        assert np.array_equal(x, y)
