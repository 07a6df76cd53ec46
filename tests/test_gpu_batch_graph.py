"""Batched multi-queue traversal, device-count launches, graph-captured frames and the launch
autotuner (GPU).

Every queue's hits are checked against the oracle's traversal of the same rays: closest hits
bit for bit on every tree (the closest hit does not depend on the traversal order), any hits
bit for bit on the BVH2 (the oracle walks the BVH2 in the kernel's order) and, on the 4-wide
trees, whose order differs, the occlusion answer plus bit equality with a single-queue launch
of the same tree.  The device-count variants read the queue size when they run; a graph
replays the very launches of an eager frame (whose every bounce the path-tracing image test
checks against the oracle and the restated integrator).
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from imbatracer_amd import native as N
from imbatracer_amd import scene as S
from imbatracer_amd import workloads as W
from tests import scenes as T

pytestmark = pytest.mark.gpu

SENTINEL = np.array([(-7, -7, 7.0, 7.0)], N.HIT_DTYPE)


def _same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.int32), np.ascontiguousarray(b).view(np.int32))


def _scenes():
    """(device scene, host arrays) of a random-triangle, an instanced and a masked scene."""
    b0, _, _ = T.random_tri_scene(20000, N.BUILDER_SBVH)
    b1 = T.instanced_scene()
    b2 = T.masked_scene()
    return [(S.Scene(b.desc()), b.arrays()) for b in (b0, b1, b2)]


def _check(got, rays, arrays, any_hit, sc, cfg):
    """got = the hits of `rays` under launch configuration cfg, checked against the oracle."""
    want, bad = oracle.trace(arrays, rays, any_hit=any_hit, threads=8)
    assert bad == 0
    wide = cfg is not None and cfg.bvh_width == 4
    if not any_hit or not wide:
        return _same(got, want)
    # any hit on a 4-wide tree: the same occlusion answer, and the single-queue launch's hit
    return (np.array_equal(got["tri_id"] < 0, want["tri_id"] < 0) and
            _same(got, sc.trace_host(rays, any_hit=True, cfg=cfg)))


def _filled_hits(n):
    hb = N.DeviceBuffer(max(n, 1) * 16)
    if n:
        hb.upload(np.repeat(SENTINEL, n))
    return hb


# (block, stack) or (block, stack, bvh_width, quantized, treelet); batches clamp treelets to 64
# nodes per wave, and a treelet never changes a result
CFGS = [None, (128, 32), (64, 64), (256, 16, 4, 1, 0), (256, 16, 4, 1, 64), (128, 32, 4, 1, 256)]


def _cfg(cfg):
    return N.TraceConfig(cfg[0], cfg[1], 0, 0, *cfg[2:]) if cfg else None


@pytest.mark.parametrize("cfg", CFGS)
def test_batch_matches_single_launches(gpu, cfg):
    scenes = _scenes()
    sizes = [0, 1, 63, 64, 65, 1000, 4097, 20000, 333, 7000]
    queues = []
    for k, n in enumerate(sizes):
        sc, arrays = scenes[k % 3]
        any_hit = k % 4 == 3
        rays = T.random_rays(max(n, 1), seed=100 + k, box=(-2.0, 2.0))[:n] if k % 3 == 1 else \
            (T.primary_rays(128, 128)[:n] if k % 3 == 2 else T.random_rays(max(n, 1), seed=100 + k)[:n])
        rb = N.DeviceBuffer.from_array(rays) if n else N.DeviceBuffer(32)
        queues.append((sc, rb, _filled_hits(n), n, None, any_hit, rays, arrays))
    tc = _cfg(cfg)
    single = _cfg(cfg[:4]) if cfg and len(cfg) > 4 else tc  # the same tree without a treelet
    batch = S.TraceBatch([q[:6] for q in queues])
    for rep in range(2):  # the work counters must return to zero after a launch
        for q in queues:
            if q[3]:
                q[2].upload(np.repeat(SENTINEL, q[3]))
        batch.launch(tc)
        for sc, rb, hb, n, _, any_hit, rays, arrays in queues:
            if n == 0:
                continue
            got = hb.download(N.HIT_DTYPE, n)
            assert _check(got, rays, arrays, any_hit, sc, single), (rep, n, any_hit)
    for sc, _ in scenes:
        assert sc.error_count() == 0


@pytest.mark.parametrize("cfg", [None, (256, 16, 4, 1, 64)])
def test_batch_device_counts_and_many_queues(gpu, cfg):
    """256 queues over 8 scenes, sizes read from device memory at launch time; entries past a
    queue's device count are left untouched.  With a treelet every wave reloads its LDS copy
    whenever its next chunk belongs to another scene."""
    bs = [T.random_tri_scene(2000 + 500 * k, N.BUILDER_SBVH if k % 2 else N.BUILDER_BINNED_SAH, seed=40 + k)[0]
          for k in range(8)]
    scenes = [S.Scene(b.desc()) for b in bs]
    rng = np.random.default_rng(7)
    items, expect = [], []
    caps = rng.integers(1, 3000, 256)
    counts = np.array([int(rng.integers(0, c + 1)) for c in caps], np.int32)
    cdev = N.DeviceBuffer.from_array(counts)
    for q in range(256):
        sc = scenes[q % 8]
        rays = T.random_rays(int(caps[q]), seed=1000 + q)
        rb = N.DeviceBuffer.from_array(rays)
        hb = _filled_hits(int(caps[q]))
        items.append((sc, rb, hb, int(caps[q]), cdev.at(4 * q), q % 5 == 0))
        expect.append(rays)
    batch = S.TraceBatch(items)
    batch.launch(_cfg(cfg))
    single = _cfg(cfg[:4]) if cfg else None
    arrays = [b.arrays() for b in bs]
    for q, (sc, rb, hb, cap, _, any_hit) in enumerate(items):
        got = hb.download(N.HIT_DTYPE, cap)
        n = int(counts[q])
        if n:
            assert _check(got[:n], expect[q][:n], arrays[q % 8], any_hit, sc, single), q
        assert _same(got[n:], np.repeat(SENTINEL, cap - n)), q


def test_intersect_dev_and_compact_dev(gpu):
    b, _, _ = T.random_tri_scene(10000, N.BUILDER_SBVH)
    sc = S.Scene(b.desc())
    arrays = b.arrays()
    rays = T.random_rays(5000, seed=3)
    rb = N.DeviceBuffer.from_array(rays)
    for any_hit in (False, True):
        hb = _filled_hits(5000)
        cnt = N.DeviceBuffer.from_array(np.array([3001], np.int32))
        fn = N.lib.imba_occluded_dev if any_hit else N.lib.imba_intersect_dev
        N.check(fn(sc.h, rb.ptr, hb.ptr, 5000, cnt.ptr, None, None), "dev")
        got = hb.download(N.HIT_DTYPE, 5000)
        assert _same(got[:3001], oracle.trace(arrays, rays[:3001], any_hit=any_hit)[0])
        assert _same(got[3001:], np.repeat(SENTINEL, 1999))
    # compaction with device counts equals the host-count version on the first *count entries
    states = np.zeros((5000, 8), np.int32)
    states[:, 0] = np.where(np.random.default_rng(1).random(5000) < 0.4, -1, np.arange(5000))
    states[:, 1:] = np.arange(5000)[:, None]
    sb = N.DeviceBuffer.from_array(states)
    ro, so = N.DeviceBuffer(5000 * 32), N.DeviceBuffer(5000 * 32)
    ro2, so2 = N.DeviceBuffer(5000 * 32), N.DeviceBuffer(5000 * 32)
    for n in (0, 1, 2048, 4000, 5000):
        cin = N.DeviceBuffer.from_array(np.array([n], np.int32))
        cout = N.DeviceBuffer.from_array(np.array([-1], np.int32))
        ws = N.DeviceBuffer(int(N.lib.imba_compact_workspace_size(5000)))
        N.check(N.lib.imba_compact_rays_dev(rb.ptr, sb.ptr, 32, 5000, cin.ptr, ro.ptr, so.ptr, cout.ptr, ws.ptr, None),
                "compact_rays_dev")
        m = S.compact_rays(rb, sb, 32, n, ro2, so2)
        assert int(cout.download(np.int32, 1)[0]) == m
        if m:
            assert _same(ro.download(N.RAY_DTYPE, m), ro2.download(N.RAY_DTYPE, m))
            assert np.array_equal(so.download(np.int32, 8 * m), so2.download(np.int32, 8 * m))


@pytest.fixture(scope="module")
def small_pt():
    scene = W.SphereGridScene(side=4, level=3, threads=4)
    return scene


def _zero_image(pt):
    pt.image.upload(np.zeros(pt.w * pt.h * 3, np.float32))


@pytest.mark.parametrize("use_batch", [False, True])
def test_graph_frame_replays_eager_frame(gpu, small_pt, use_batch):
    pt = W.PathTracer(small_pt, 160, 120)
    _zero_image(pt)
    counts, _, queues = pt.frame(keep_queues=8)
    pt.stream.sync()
    img_eager = pt.image.download(np.float32, pt.w * pt.h * 3)
    assert len(counts) >= 3 and counts[0] == 160 * 120
    pt.build_graph(use_batch=use_batch)   # runs one eager device-count frame, then captures one

    def buffers():
        out = [b.download(np.uint8, b.nbytes) for b in pt.rays + pt.states]
        return out + [b.download(np.uint8, b.nbytes) for b in (pt.hits, pt.shadow_rays, pt.shadow_states,
                                                                pt.shadow_hits)]

    for rep in range(3):
        # eager work and synchronisation between replays
        _zero_image(pt)
        c_eager, _, _ = pt.frame()
        pt.stream.sync()
        assert c_eager == counts
        after_eager = buffers()
        _zero_image(pt)
        c = pt.frame_graph()
        assert c == counts, (rep, c, counts)
        img = pt.image.download(np.float32, pt.w * pt.h * 3)
        # the image is a sum of float atomics: equal up to the order of additions
        assert np.allclose(img, img_eager, rtol=1e-5, atol=1e-6)
        # a graph frame writes what an eager frame writes: queues, states and hits
        for k, (x, y) in enumerate(zip(buffers(), after_eager)):
            assert np.array_equal(x, y), (rep, k)
    assert pt.gpu_scene.error_count() == 0


def test_autotune_returns_a_supported_shape(gpu):
    b, _, _ = T.random_tri_scene(20000, N.BUILDER_SBVH)
    sc = S.Scene(b.desc())
    rays = T.random_rays(50000, seed=4)
    rb = N.DeviceBuffer.from_array(rays)
    hb = N.DeviceBuffer(50000 * 16)
    cfg, ms = sc.autotune(rb, hb, 50000, reps=2)
    assert (cfg.block_size, cfg.stack_lds) in [(256, 16), (256, 8), (256, 32), (128, 16), (128, 32), (64, 32), (64, 64)]
    assert cfg.persistent in (0, 32) and ms > 0
    assert _same(sc.trace_host(rays, cfg=cfg), oracle.trace(b.arrays(), rays, threads=8)[0])
