"""GPU traversal parity: the HIP kernels against the CPU oracle on identical rays and BVHs.

Closest-hit and any-hit results must be identical bit for bit (tri_id, inst_id, t, u): both
sides use the same fixed arithmetic and traversal order.  Visit counts of the stats variant
must equal the oracle's.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from imbatracer_amd import native as N
from imbatracer_amd import scene as S
from tests import scenes as T

pytestmark = pytest.mark.gpu


def _same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.int32), np.ascontiguousarray(b).view(np.int32))


def _report(got, want):
    bad = np.nonzero(got.view(np.int32).reshape(-1, 4).any(1) != want.view(np.int32).reshape(-1, 4).any(1))
    diff = np.nonzero((got.view(np.int32).reshape(-1, 4) != want.view(np.int32).reshape(-1, 4)).any(1))[0]
    return "%d rays differ, first: %s" % (len(diff), [(int(k), got[k], want[k]) for k in diff[:3]])


@pytest.mark.parametrize("builder", [N.BUILDER_BINNED_SAH, N.BUILDER_SBVH])
def test_closest_and_occluded_c1_shape(gpu, builder):
    b, v, i = T.random_tri_scene(65536, builder)
    arrays = b.arrays()
    scene = S.Scene(b.desc())
    prim = T.primary_rays(256, 256)
    got = scene.trace_host(prim)
    want, bad = oracle.trace(arrays, prim, threads=8)
    assert bad == 0
    assert _same(got, want), _report(got, want)
    assert (got["tri_id"] >= 0).sum() > 1000
    shadow = T.shadow_rays(prim, want)
    got_o = scene.trace_host(shadow, any_hit=True)
    want_o, bad = oracle.trace(arrays, shadow, any_hit=True, threads=8)
    assert _same(got_o, want_o), _report(got_o, want_o)
    assert scene.error_count() == 0


def test_c1_full_size(gpu):
    """The CPU-runnable configuration at full size: 65,536 random triangles (binned SAH),
    1024x1024 primary rays and 1,048,576 occlusion rays towards (0.5, 2, 0.5)."""
    b, v, i = T.random_tri_scene(65536, N.BUILDER_BINNED_SAH)
    arrays = b.arrays()
    scene = S.Scene(b.desc())
    prim = T.primary_rays(1024, 1024)
    rb = N.DeviceBuffer.from_array(prim)
    hb = N.DeviceBuffer(len(prim) * 16)
    cb = N.DeviceBuffer(len(prim) * 12)
    scene.intersect_stats(rb, hb, cb, len(prim))
    got = hb.download(N.HIT_DTYPE, len(prim))
    cnt = cb.download(np.uint32, 3 * len(prim)).reshape(-1, 3)
    want, wcnt, bad = oracle.trace(arrays, prim, counts=True)
    assert bad == 0
    assert _same(got, want), _report(got, want)
    assert np.array_equal(cnt, wcnt)
    scene.intersect(rb, hb, len(prim))
    assert _same(hb.download(N.HIT_DTYPE, len(prim)), want)
    shadow = T.shadow_rays(prim, want)
    got_o = scene.trace_host(shadow, any_hit=True)
    want_o, bad = oracle.trace(arrays, shadow, any_hit=True, threads=16)
    assert _same(got_o, want_o), _report(got_o, want_o)
    # the primary hits form a plausible image of the triangle cloud
    assert 0.2 < (want["tri_id"] >= 0).mean() < 0.99


def test_incoherent_rays_and_stats(gpu):
    b, v, i = T.random_tri_scene(65536, N.BUILDER_BINNED_SAH)
    arrays = b.arrays()
    scene = S.Scene(b.desc())
    rays = T.random_rays(20000)
    rb = N.DeviceBuffer.from_array(rays)
    hb = N.DeviceBuffer(len(rays) * 16)
    cb = N.DeviceBuffer(len(rays) * 12)
    for any_hit in (False, True):
        scene.intersect_stats(rb, hb, cb, len(rays), any_hit=any_hit)
        got = hb.download(N.HIT_DTYPE, len(rays))
        cnt = cb.download(np.uint32, 3 * len(rays)).reshape(-1, 3)
        want, wcnt, bad = oracle.trace(arrays, rays, any_hit=any_hit, counts=True)
        assert _same(got, want), _report(got, want)
        assert np.array_equal(cnt, wcnt)
        # the plain kernel gives the same hits as the stats variant
        (scene.occluded if any_hit else scene.intersect)(rb, hb, len(rays))
        assert _same(hb.download(N.HIT_DTYPE, len(rays)), want)


@pytest.mark.parametrize("shape", [(256, 16), (256, 8), (256, 32), (128, 16), (128, 32), (64, 32), (64, 64)])
def test_launch_shapes_agree(gpu, shape):
    b, v, i = T.random_tri_scene(20000, N.BUILDER_SBVH)
    scene = S.Scene(b.desc())
    rays = T.random_rays(30000, seed=8)
    want, _ = oracle.trace(b.arrays(), rays, threads=8)
    cfg = N.TraceConfig(shape[0], shape[1], 0, 0)
    for any_hit in (False, True):
        got = scene.trace_host(rays, any_hit=any_hit, cfg=cfg)
        if not any_hit:
            assert _same(got, want), _report(got, want)
        else:
            assert np.array_equal(got["tri_id"] >= 0, oracle.trace(b.arrays(), rays, any_hit=True)[0]["tri_id"] >= 0)


@pytest.mark.parametrize("shape", [(256, 32, 1), (256, 16, 16), (256, 8, 64), (128, 32, 8), (64, 64, 32)])
def test_persistent_fetch_agrees(gpu, shape):
    """The persistent kernel (rays fetched per wave from a global counter) gives the
    oracle's results, repeatedly (its counter must return to zero after each launch)."""
    b = T.instanced_scene()
    arrays = b.arrays()
    scene = S.Scene(b.desc())
    rays = T.random_rays(50000, seed=12, box=(-2.0, 2.0))
    cfg = N.TraceConfig(shape[0], shape[1], shape[2], 0)
    for any_hit in (False, True):
        want, _ = oracle.trace(arrays, rays, any_hit=any_hit, threads=8)
        for _ in range(3):
            got = scene.trace_host(rays, any_hit=any_hit, cfg=cfg)
            assert _same(got, want), _report(got, want)
    small = rays[:37]
    assert _same(scene.trace_host(small, cfg=cfg), oracle.trace(arrays, small)[0])


def test_instanced_scene(gpu):
    b = T.instanced_scene()
    arrays = b.arrays()
    assert len(arrays["instances"]) == 7
    scene = S.Scene(b.desc())
    rays = T.random_rays(40000, seed=9, box=(-2.0, 2.0))
    for any_hit in (False, True):
        got = scene.trace_host(rays, any_hit=any_hit)
        want, bad = oracle.trace(arrays, rays, any_hit=any_hit, threads=8)
        assert bad == 0
        assert _same(got, want), _report(got, want)
    assert len(np.unique(got["inst_id"][got["tri_id"] >= 0])) >= 1


def test_masked_scene(gpu):
    b = T.masked_scene()
    arrays = b.arrays()
    scene = S.Scene(b.desc())
    rays = T.primary_rays(128, 128)
    for any_hit in (False, True):
        got = scene.trace_host(rays, any_hit=any_hit)
        want, bad = oracle.trace(arrays, rays, any_hit=any_hit)
        assert bad == 0
        assert _same(got, want), _report(got, want)
    # the mask removes hits: fewer hits than with an all-opaque copy of the scene
    opaque = dict(arrays)
    opaque["mask_buffer"] = np.ones_like(arrays["mask_buffer"])
    want_opaque, _ = oracle.trace(opaque, rays)
    assert (want["tri_id"] >= 0).sum() < (want_opaque["tri_id"] >= 0).sum()


def test_reference_ffi_entry_points(gpu):
    """intersect/occluded_gpu_masked_instanced with device pointers, as RayQueue calls them."""
    b, v, i = T.random_tri_scene(8192, N.BUILDER_SBVH)
    a = b.arrays()
    bufs = {k: N.DeviceBuffer.from_array(a[k]) for k in
            ("nodes", "instances", "tris", "indices", "texcoords", "masks", "mask_buffer")}
    rays = T.primary_rays(64, 64)
    rb = N.DeviceBuffer.from_array(rays)
    hb = N.DeviceBuffer(len(rays) * 16)
    for any_hit, fn in ((False, N.lib.intersect_gpu_masked_instanced), (True, N.lib.occluded_gpu_masked_instanced)):
        fn(a["root"], bufs["nodes"].ptr, bufs["instances"].ptr, bufs["tris"].ptr, rb.ptr, hb.ptr, bufs["indices"].ptr,
           bufs["texcoords"].ptr, bufs["masks"].ptr, bufs["mask_buffer"].ptr, len(rays))
        got = hb.download(N.HIT_DTYPE, len(rays))
        want, _ = oracle.trace(a, rays, any_hit=any_hit)
        assert _same(got, want), _report(got, want)


def test_empty_and_tiny_queues(gpu):
    b, v, i = T.random_tri_scene(1000)
    scene = S.Scene(b.desc())
    rb = N.DeviceBuffer(32)
    hb = N.DeviceBuffer(16)
    scene.intersect(rb, hb, 0)
    rays = T.primary_rays(1, 1)
    assert _same(scene.trace_host(rays), oracle.trace(b.arrays(), rays)[0])


def test_single_triangle_and_dummy_parent(gpu):
    """A one-triangle mesh is a single leaf under a dummy parent whose right child must never
    be visited, also for rays through the origin (where the dummy box sits)."""
    v = np.array([[-1, -1, 1, 1], [1, -1, 1, 1], [0, 1, 1, 1]], np.float32)
    i = np.array([[0, 1, 2, 0]], np.uint32)
    b = S.SceneBuild()
    b.add_mesh(v, i)
    b.run(N.BUILDER_SBVH, 1)
    a = b.arrays()
    assert (a["nodes"]["right"] == N.DUMMY_CHILD).sum() == 2  # mesh level and top level
    scene = S.Scene(b.desc())
    o = np.array([[0, 0, -1], [0, 0, -2], [0.2, -0.3, 0]], np.float32)
    d = np.array([[0, 0, 1], [0, 0, 1], [0, 0, 1]], np.float32)
    rays = S.make_rays(o, d, 0.0, 1e30)
    got = scene.trace_host(rays)
    want, bad = oracle.trace(a, rays)
    assert bad == 0 and _same(got, want)
    assert list(got["tri_id"]) == [0, 0, 0]
