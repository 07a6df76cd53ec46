"""Pins of the compaction oracle (oracle/compact.c) against RayQueue::compact_hits /
compact_rays (render/scheduling/ray_queue.h:161-207)."""
import numpy as np

import oracle
from imbatracer_amd import native as N


def _q(tri_ids):
    n = len(tri_ids)
    rays = np.zeros(n, N.RAY_DTYPE)
    rays["org"][:, 0] = np.arange(n)
    hits = np.zeros(n, N.HIT_DTYPE)
    hits["tri_id"] = tri_ids
    states = np.zeros((n, 4), np.int32)
    states[:, 0] = np.arange(n)
    return rays, hits, states


def test_compact_hits_hand_worked():
    # H M M H M H: the reference loop swaps each later hit into the first miss slot:
    # i=3: swap(1,3) -> H0 H3 M2 M1 M4 H5; i=5: swap(2,5) -> H0 H3 H5 M1 M4 M2
    rays, hits, states = _q([0, -1, -1, 3, -1, 5])
    n, r, h, s = oracle.compact_hits(rays, hits, states)
    assert n == 3
    assert s[:, 0].tolist() == [0, 3, 5, 1, 4, 2]
    assert h["tri_id"].tolist() == [0, 3, 5, -1, -1, -1]
    assert r["org"][:, 0].tolist() == [0, 3, 5, 1, 4, 2]


def test_compact_hits_edge_cases():
    for ids, want in (([], 0), ([-1, -1], 0), ([1, 2, 3], 3), ([-1, 7], 1), ([7, -1], 1)):
        rays, hits, states = _q(ids)
        n, r, h, s = oracle.compact_hits(rays, hits, states)
        assert n == want


def test_compact_hits_properties():
    rng = np.random.default_rng(3)
    ids = np.where(rng.random(5000) < 0.4, rng.integers(0, 100, 5000), -1)
    rays, hits, states = _q(ids)
    n, r, h, s = oracle.compact_hits(rays, hits, states)
    hit = ids >= 0
    assert n == hit.sum()
    assert np.array_equal(s[:n, 0], np.nonzero(hit)[0])          # hit prefix stable
    assert np.array_equal(np.sort(s[n:, 0]), np.nonzero(~hit)[0])  # misses: same multiset
    assert np.array_equal(r["org"][:, 0].astype(int), s[:, 0])     # rays moved with their states


def test_compact_rays_hand_worked_and_properties():
    rays, hits, states = _q([0] * 6)
    states[:, 0] = [0, -1, -1, 3, -1, 5]
    n, r, s = oracle.compact_rays(rays, states)
    assert n == 3 and s[:3, 0].tolist() == [0, 3, 5] and r["org"][:3, 0].tolist() == [0, 3, 5]
    states[:, 0] = np.arange(6)
    assert oracle.compact_rays(rays, states)[0] == 6  # nothing removed: size unchanged
    states[:, 0] = -1
    assert oracle.compact_rays(rays, states)[0] == 0
    rng = np.random.default_rng(5)
    rays, hits, states = _q([0] * 4000)
    alive = rng.random(4000) < 0.3
    states[~alive, 0] = -1
    n, r, s = oracle.compact_rays(rays, states)
    assert n == alive.sum() and np.array_equal(s[:n, 0], np.nonzero(alive)[0])


def _mat_queue(tri_ids):
    rays = np.zeros(len(tri_ids), oracle.RAY_DTYPE)
    hits = np.zeros(len(tri_ids), oracle.HIT_DTYPE)
    hits["tri_id"] = tri_ids
    return rays, hits


def test_sort_by_material_hand_worked():
    # triangles 0..4 with materials 2, 0, 2, 1, 7 (7 is out of range for num_mats = 3)
    indices = np.array([[0, 0, 0, 2], [0, 0, 0, 0], [0, 0, 0, 2], [0, 0, 0, 1], [0, 0, 0, 7]], np.int32)
    rays, hits = _mat_queue([2, 0, 3, -1, 1, 2, 4, 9])
    r, order, starts = oracle.sort_by_material(indices, rays, hits, 3)
    # keys per entry: 2, 2, 1, 3(miss), 0, 2, 3(mat 7), 3(tri 9 out of range)
    assert r["dir"][:, 3].view(np.int32).tolist() == [2, 2, 1, 3, 0, 2, 3, 3]
    assert order.tolist() == [4, 2, 0, 1, 5, 3, 6, 7]
    assert starts.tolist() == [0, 1, 2, 5]


def test_sort_by_material_properties():
    rng = np.random.default_rng(11)
    mats = rng.integers(0, 50, 3000)
    indices = np.zeros((3000, 4), np.int32)
    indices[:, 3] = mats
    rays, hits = _mat_queue(np.where(rng.random(20000) < 0.9, rng.integers(0, 3000, 20000), -1))
    r, order, starts = oracle.sort_by_material(indices, rays, hits, 50)
    keys = r["dir"][:, 3].view(np.int32)
    assert sorted(order.tolist()) == list(range(20000))                # a permutation
    assert np.all(np.diff(keys[order]) >= 0)                            # grouped by material
    for k in range(51):                                                 # queue order inside a group
        grp = order[starts[k]:(starts[k + 1] if k < 50 else 20000)]
        assert np.all(np.diff(grp) > 0) and np.all(keys[grp] == k)
    assert (keys == 50).sum() == (hits["tri_id"] < 0).sum()
