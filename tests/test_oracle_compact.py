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
