"""GPU traversal of the 4-wide tree derived at upload (bvh_width = 4) against the CPU oracle's
BVH2 traversal.

The 4-wide tree keeps the BVH2's leaves and leaf boxes and uses the same box and triangle
arithmetic, so closest hits (smallest t, then smaller triangle id, then smaller instance id)
must equal the oracle's bit for bit.  An occlusion query may stop at another triangle than the
BVH2's first hit; its occluded / unoccluded answer must be the same, and a reported hit must be
a real one (t inside the ray's interval).  Staging the quantized tree's top levels in LDS
(treelet) changes only where those nodes are read from, so its results are held to the same bar.
"""
import numpy as np
import pytest

import oracle
from imbatracer_amd import native as N
from imbatracer_amd import scene as S
from tests import scenes as T

pytestmark = pytest.mark.gpu

WIDE = N.TraceConfig(0, 0, 0, 0, 4)
# 4-wide, 4-wide quantized, quantized with its top 256 nodes staged in LDS
TREES = [N.TraceConfig(0, 0, 0, 0, 4, 0), N.TraceConfig(0, 0, 0, 0, 4, 1), N.TraceConfig(0, 0, 0, 0, 4, 1, 256)]
TREE_IDS = ["wide", "quantized", "treelet"]


def _same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.int32), np.ascontiguousarray(b).view(np.int32))


def _check_occlusion(got, want, rays):
    assert np.array_equal(got["tri_id"] >= 0, want["tri_id"] >= 0)
    hit = got["tri_id"] >= 0
    t = got["tmax"][hit]
    assert np.all(t >= rays["org"][hit, 3]) and np.all(t <= rays["dir"][hit, 3])


@pytest.mark.parametrize("tree", TREES, ids=TREE_IDS)
@pytest.mark.parametrize("builder", [N.BUILDER_BINNED_SAH, N.BUILDER_SBVH])
def test_closest_and_occluded(gpu, builder, tree):
    WIDE = tree
    b, _, _ = T.random_tri_scene(65536, builder)
    a = b.arrays()
    sc = S.Scene(b.desc())
    prim = T.primary_rays(512, 512)
    want, _ = oracle.trace(a, prim, threads=8)
    assert _same(sc.trace_host(prim, cfg=WIDE), want)
    shadow = T.shadow_rays(prim, want)
    want_o, _ = oracle.trace(a, shadow, any_hit=True, threads=8)
    _check_occlusion(sc.trace_host(shadow, any_hit=True, cfg=WIDE), want_o, shadow)
    rays = T.random_rays(40000, seed=21)
    want_r, _ = oracle.trace(a, rays, threads=8)
    assert _same(sc.trace_host(rays, cfg=WIDE), want_r)
    assert sc.error_count() == 0


@pytest.mark.parametrize("quantized,treelet", [(0, 0), (1, 0), (1, 64), (1, 256)])
@pytest.mark.parametrize("shape", [(256, 16, 0), (256, 8, 0), (128, 32, 0), (64, 64, 0), (256, 16, 32), (128, 32, 8)])
def test_launch_shapes(gpu, shape, quantized, treelet):
    b, _, _ = T.random_tri_scene(20000, N.BUILDER_SBVH)
    sc = S.Scene(b.desc())
    rays = T.random_rays(30000, seed=8)
    want, _ = oracle.trace(b.arrays(), rays, threads=8)
    cfg = N.TraceConfig(shape[0], shape[1], shape[2], 0, 4, quantized, treelet)
    for _ in range(2):
        assert _same(sc.trace_host(rays, cfg=cfg), want)


@pytest.mark.parametrize("tree", TREES, ids=TREE_IDS)
def test_instanced_and_masked(gpu, tree):
    WIDE = tree
    b = T.instanced_scene()
    sc = S.Scene(b.desc())
    rays = T.random_rays(40000, seed=9, box=(-2.0, 2.0))
    want, _ = oracle.trace(b.arrays(), rays, threads=8)
    assert _same(sc.trace_host(rays, cfg=WIDE), want)
    want_o, _ = oracle.trace(b.arrays(), rays, any_hit=True, threads=8)
    _check_occlusion(sc.trace_host(rays, any_hit=True, cfg=WIDE), want_o, rays)
    m = T.masked_scene()
    msc = S.Scene(m.desc())
    prim = T.primary_rays(128, 128)
    want, _ = oracle.trace(m.arrays(), prim)
    assert _same(msc.trace_host(prim, cfg=WIDE), want)
    want_o, _ = oracle.trace(m.arrays(), prim, any_hit=True)
    _check_occlusion(msc.trace_host(prim, any_hit=True, cfg=WIDE), want_o, prim)


@pytest.mark.parametrize("tree", TREES, ids=TREE_IDS)
def test_single_triangle(gpu, tree):
    WIDE = tree
    v = np.array([[-1, -1, 1, 1], [1, -1, 1, 1], [0, 1, 1, 1]], np.float32)
    i = np.array([[0, 1, 2, 0]], np.uint32)
    b = S.SceneBuild()
    b.add_mesh(v, i)
    b.run(N.BUILDER_SBVH, 1)
    sc = S.Scene(b.desc())
    o = np.array([[0, 0, -1], [0, 0, -2], [0.2, -0.3, 0], [0, 0, 0]], np.float32)
    d = np.array([[0, 0, 1], [0, 0, 1], [0, 0, 1], [1, 0, 0]], np.float32)
    rays = S.make_rays(o, d, 0.0, 1e30)
    got = sc.trace_host(rays, cfg=WIDE)
    assert list(got["tri_id"]) == [0, 0, 0, -1]
    assert _same(got, oracle.trace(b.arrays(), rays)[0])


def test_visit_counts_and_batch(gpu):
    b, _, _ = T.random_tri_scene(65536, N.BUILDER_SBVH)
    sc = S.Scene(b.desc())
    rays = T.random_rays(20000, seed=5)
    rb = N.DeviceBuffer.from_array(rays)
    hb = N.DeviceBuffer(len(rays) * 16)
    c2, c4 = N.DeviceBuffer(len(rays) * 12), N.DeviceBuffer(len(rays) * 12)
    want, want_counts, _ = oracle.trace(b.arrays(), rays, counts=True)
    sc.trace_stats(rb, hb, c2, len(rays), cfg=N.TraceConfig(0, 0, 0, 0, 2))
    h2 = hb.download(N.HIT_DTYPE, len(rays))
    assert _same(h2, want)
    sc.trace_stats(rb, hb, c4, len(rays), cfg=WIDE)
    h4 = hb.download(N.HIT_DTYPE, len(rays))
    assert _same(h4, want)
    n2 = c2.download(np.uint32, 3 * len(rays)).reshape(-1, 3).sum(0)
    n4 = c4.download(np.uint32, 3 * len(rays)).reshape(-1, 3).sum(0)
    assert n4[0] < n2[0]          # fewer (wider) nodes visited
    assert n4[2] == n2[2]         # one instance entered per ray
    # the BVH2 counts are the oracle's
    assert np.array_equal(c2.download(np.uint32, 3 * len(rays)).reshape(-1, 3), want_counts)
    # the batched kernel on the 4-wide trees: the oracle's closest hits
    out = N.DeviceBuffer(len(rays) * 16)
    batch = S.TraceBatch([(sc, rb, out, len(rays), None, False)])
    for tree in TREES:
        batch.launch(tree)
        assert _same(out.download(N.HIT_DTYPE, len(rays)), want)
    # the quantized tree: same hits, about as many node visits as the exact 4-wide tree
    cq = N.DeviceBuffer(len(rays) * 12)
    sc.trace_stats(rb, hb, cq, len(rays), cfg=TREES[1])
    assert _same(hb.download(N.HIT_DTYPE, len(rays)), want)
    nq = cq.download(np.uint32, 3 * len(rays)).reshape(-1, 3).sum(0)
    assert 0.9 * n4[0] < nq[0] < 1.5 * n4[0]


def test_treelet_arguments(gpu):
    """A treelet needs the quantized tree and at most 256 nodes; a larger request than the tree
    has is clamped to the tree."""
    v = np.array([[-1, -1, 1, 1], [1, -1, 1, 1], [0, 1, 1, 1]], np.float32)
    b = S.SceneBuild()
    b.add_mesh(v, np.array([[0, 1, 2, 0]], np.uint32))
    b.run(N.BUILDER_SBVH, 1)
    sc = S.Scene(b.desc())
    rays = S.make_rays(np.array([[0, 0, -1]], np.float32), np.array([[0, 0, 1]], np.float32), 0.0, 1e30)
    rb, hb = N.DeviceBuffer.from_array(rays), N.DeviceBuffer(16)
    for cfg in (N.TraceConfig(0, 0, 0, 0, 4, 1, 257), N.TraceConfig(0, 0, 0, 0, 4, 1, -1),
                N.TraceConfig(0, 0, 0, 0, 2, 0, 64), N.TraceConfig(0, 0, 0, 0, 4, 0, 64)):
        assert N.lib.imba_intersect(sc.h, rb.ptr, hb.ptr, 1, N.C.byref(cfg), None) == 1  # IMBA_ERR_INVALID_ARG
    got = sc.trace_host(rays, cfg=N.TraceConfig(0, 0, 0, 0, 4, 1, 256))
    assert list(got["tri_id"]) == [0]
