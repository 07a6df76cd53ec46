"""Pins of the hash-grid oracle (oracle/hashgrid.c) against HashGrid
(rangesearch/rangesearch.h:25-148), independent of the GPU.

- a hand-computed bucket hash;
- the table-size history max(previous, 2 * count);
- geometry: with cell size 2r, the probed cell and its neighbours towards the nearer half
  contain the whole radius ball, so when fewer than k photons are found the distinct
  results are exactly the photons within r of the query (brute force over all photons);
- the list is sorted by distance, all within r, and duplicates occur only for coinciding
  buckets; the k cap keeps the k nearest (ties: the later candidate replaces the k-th);
- queries outside the grown bounds return nothing.
"""
import numpy as np

import oracle

M = 0xFFFFFFFF


def bucket(x, y, z, table):
    return (((x * 73856093) & M) ^ ((y * 19349663) & M) ^ ((z * 83492791) & M)) % table


def test_hash_by_hand():
    # (1*73856093) ^ (2*19349663) ^ (3*83492791) = 73856093 ^ 38699326 ^ 250478373
    assert bucket(1, 2, 3, 1 << 32) == 147163718
    # negative cell coordinates wrap as uint32 (cell -1 = 0xFFFFFFFF)
    assert bucket(0xFFFFFFFF, 0, 0, 1000) == ((0xFFFFFFFF * 73856093) & M) % 1000


def test_oracle_places_photons_in_their_buckets():
    rng = np.random.default_rng(9)
    p = rng.random((3000, 3), dtype=np.float32)
    g = _grid(p, 0.02)
    ends, order = g.cell_ends(), g.order()
    inv = np.float32(g.g.inv_cell_size)
    bmin = np.array(g.g.bbox_min, np.float32)
    slot_of = np.empty(len(p), np.int64)
    slot_of[order] = np.arange(len(p))
    for i in range(len(p)):
        c = np.floor(inv * (p[i] - bmin)).astype(np.int64)
        b = bucket(int(c[0]) & M, int(c[1]) & M, int(c[2]) & M, g.g.table_size)
        lo = 0 if b == 0 else ends[b - 1]
        assert lo <= slot_of[i] < ends[b]
    # within a bucket, photons are in index order
    for b in np.nonzero(np.diff(np.concatenate([[0], ends])) > 1)[0]:
        lo = 0 if b == 0 else ends[b - 1]
        assert np.all(np.diff(order[lo:ends[b]]) > 0)


def _grid(p, r):
    g = oracle.Grid()
    g.build(p, r)
    return g


def test_table_history_and_bounds():
    rng = np.random.default_rng(0)
    g = oracle.Grid()
    for n, want in ((1000, 2000), (10, 2000), (5000, 10000), (0, 10000)):
        p = rng.random((n, 3), dtype=np.float32)
        g.build(p, 0.1)
        assert g.g.table_size == want
    p = rng.random((100, 3), dtype=np.float32)
    g.build(p, 0.1)
    lo, hi = p.min(0), p.max(0)
    e = hi - lo
    assert np.array_equal(np.array(g.g.bbox_min, np.float32), lo - e * np.float32(0.001))
    assert np.array_equal(np.array(g.g.bbox_max, np.float32), hi + e * np.float32(0.001))
    assert g.cell_ends()[-1] == 100 and sorted(g.order().tolist()) == list(range(100))


def test_queries_against_brute_force():
    rng = np.random.default_rng(1)
    n, r, k = 4000, 0.04, 64
    p = rng.random((n, 3), dtype=np.float32)
    g = _grid(p, r)
    qs = np.concatenate([p[:300] + rng.normal(0, r / 2, (300, 3)).astype(np.float32),
                         rng.random((200, 3), dtype=np.float32)]).astype(np.float32)
    r2 = np.float32(r) * np.float32(r)
    checked = 0
    bmin, bmax = np.array(g.g.bbox_min, np.float32), np.array(g.g.bbox_max, np.float32)
    for q in qs:
        idx, dist = g.query(q, k)
        if not (np.all(q >= bmin) and np.all(q <= bmax)):
            assert len(idx) == 0  # early out outside the grown bounds (rangesearch.h:80)
            continue
        d = q[None, :] - p
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        within = np.nonzero(d2 <= r2)[0]
        assert np.all(np.diff(dist) >= 0) and np.all(dist <= r2)
        assert np.array_equal(dist, d2[idx])
        if len(idx) < k:
            assert set(idx.tolist()) == set(within.tolist())
            checked += 1
    assert checked > 400


def test_k_cap_keeps_nearest_and_tie_rule():
    # photons on a line at distances 0.01 * j from the query; k = 5
    q = np.array([0.5, 0.5, 0.5], np.float32)
    p = np.array([[0.5 + 0.01 * j, 0.5, 0.5] for j in (7, 3, 9, 1, 4, 8, 2, 6, 5, 0)], np.float32)
    # a first, large build leaves a table of 200,000 buckets (it only grows), so the 8 probed
    # cells of the small builds below fall into distinct buckets
    g = oracle.Grid()
    g.build(np.random.default_rng(0).random((100_000, 3), dtype=np.float32), 0.2)
    g.build(np.concatenate([p, [[0.0, 0.0, 0.0], [1, 1, 1]]]).astype(np.float32), 0.2)
    assert g.g.table_size == 200_000
    idx, dist = g.query(q, 5)
    assert len(idx) == 5
    assert [round(float(x), 2) for x in (p[idx, 0] - 0.5)] == [0.0, 0.01, 0.02, 0.03, 0.04]
    # many photons at the same point: the first k-1 stay, the last one seen takes slot k
    same = np.tile(q, (20, 1))
    g.build(np.concatenate([same, [[0, 0, 0], [1, 1, 1]]]).astype(np.float32), 0.1)
    idx, dist = g.query(q, 4)
    assert idx.tolist() == [0, 1, 2, 19]


def test_outside_bounds_returns_nothing():
    rng = np.random.default_rng(2)
    p = rng.random((500, 3), dtype=np.float32)
    g = _grid(p, 0.05)
    assert len(g.query(np.array([2.0, 0.5, 0.5], np.float32), 10)[0]) == 0
    assert len(g.query(np.array([0.5, -0.5, 0.5], np.float32), 10)[0]) == 0


def test_duplicates_from_coinciding_buckets():
    """With a table of 2 buckets every probed cell maps to one of two buckets, so the same
    photons are visited several times and appear several times (rangesearch.h:100-126)."""
    p = np.array([[0.5, 0.5, 0.5]], np.float32)
    g = _grid(p, 0.25)  # 1 photon -> table of 2 buckets
    assert g.g.table_size == 2
    idx, dist = g.query(np.array([0.5, 0.5, 0.5], np.float32), 16)
    # the photon's bucket is probed once per probed cell that hashes to it
    pc = np.floor(np.float32(1 / 0.5) * (p[0] - np.array(g.g.bbox_min, np.float32))).astype(np.int64)
    cell = np.float32(1 / 0.5) * (np.array([0.5] * 3, np.float32) - np.array(g.g.bbox_min, np.float32))
    fl = np.floor(cell).astype(np.int64)
    other = fl + np.where(cell - np.floor(cell) < 0.5, -1, 1)
    b_photon = bucket(*(int(c) & M for c in pc), 2)
    probes = 0
    for j in range(8):
        x = other[0] if j & 4 else fl[0]
        y = other[1] if j & 2 else fl[1]
        z = other[2] if j & 1 else fl[2]
        probes += bucket(int(x) & M, int(y) & M, int(z) & M, 2) == b_photon
    assert len(idx) == probes >= 2 and set(idx.tolist()) == {0}


def test_query_many_is_the_query_loop():
    rng = np.random.default_rng(12)
    p = rng.random((4000, 4), dtype=np.float32)
    g = oracle.Grid()
    g.build(p, 0.06)
    q = rng.random((500, 4), dtype=np.float32) * 1.1 - 0.05  # some outside the bounds
    idx, dist, cnt = g.query_many(q, 7)
    for j in range(len(q)):
        a, b = g.query(q[j, :3], 7)
        assert cnt[j] == len(a)
        assert np.array_equal(idx[j, :cnt[j]], a) and np.array_equal(dist[j, :cnt[j]], b)
    assert (cnt == 0).any() and (cnt == 7).any()
