"""GPU photon hash grid against the serial restatement of HashGrid (rangesearch/rangesearch.h).

Build: bucket table size, grown bounds, bucket ends and the slot order (photon indices by
bucket, then by index) equal the oracle's exactly.  Query: the photon index lists (with the
duplicates that coinciding buckets produce, in result order) and the squared distances
equal the oracle's bit for bit, for k up to 64 (list in lanes) and above (list in LDS, up to
4096; the reference sizes its list by k, rangesearch.h:98).
"""
import numpy as np
import pytest

import oracle
from imbatracer_amd import native as N
from imbatracer_amd import scene as S

pytestmark = pytest.mark.gpu


def _grid(photons, radius, stride_words=4, g=None):
    g = g or S.HashGrid()
    pb = N.DeviceBuffer.from_array(photons)
    g.build(pb, 4 * stride_words, 0, len(photons), radius)
    return g, pb


def _query(g, queries, k):
    nq = len(queries)
    qb = N.DeviceBuffer.from_array(queries)
    oi, od, oc = N.DeviceBuffer(max(nq * k * 4, 4)), N.DeviceBuffer(max(nq * k * 4, 4)), N.DeviceBuffer(max(nq * 4, 4))
    g.query(qb, queries.strides[0], 0, nq, k, oi, oc, od)
    return (oi.download(np.int32, nq * k).reshape(nq, k), od.download(np.float32, nq * k).reshape(nq, k),
            oc.download(np.int32, nq))


def _photons(n, seed, clustered=False):
    rng = np.random.default_rng(seed)
    p = rng.random((n, 4), dtype=np.float32)
    if clustered:  # photons on a few spheres, like deposits on the sphere-grid scene
        c = rng.random((8, 3), dtype=np.float32)
        d = rng.normal(size=(n, 3)).astype(np.float32)
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        p[:, :3] = c[rng.integers(0, 8, n)] + 0.05 * d
    p[:, 3] = rng.random(n, dtype=np.float32)  # payload
    return p


def _check_build(g, og, n):
    info = g.info()
    assert info.table_size == og.g.table_size
    assert list(info.bbox_min) == list(og.g.bbox_min) and list(info.bbox_max) == list(og.g.bbox_max)
    import ctypes as C

    e = np.empty(info.table_size, np.int32)
    o = np.empty(n, np.int32)
    N.check(N.lib.imba_memcpy_d2h(N.ptr(e), C.c_void_p(info.cell_ends), e.nbytes, None))
    N.check(N.lib.imba_memcpy_d2h(N.ptr(o), C.c_void_p(info.order), o.nbytes, None))
    assert np.array_equal(e, og.cell_ends())
    assert np.array_equal(o, og.order())


def _check_queries(g, og, queries, k):
    idx, dist, cnt = _query(g, queries, k)
    nonempty = 0
    for q in range(len(queries)):
        want_i, want_d = og.query(queries[q, :3], k)
        assert cnt[q] == len(want_i), q
        assert np.array_equal(idx[q, :cnt[q]], want_i), q
        assert np.array_equal(dist[q, :cnt[q]].view(np.int32), want_d.view(np.int32)), q
        nonempty += cnt[q] > 0
    return nonempty


@pytest.mark.parametrize("n,radius,k,clustered", [(1, 0.1, 10, False), (1000, 0.05, 10, False),
                                                  (50_000, 0.01, 10, False), (50_000, 0.004, 64, True),
                                                  (200_000, 0.02, 16, True), (100_000, 0.05, 1, False)])
def test_build_and_query(gpu, n, radius, k, clustered):
    photons = _photons(n, seed=n, clustered=clustered)
    g, _ = _grid(photons, radius)
    og = oracle.Grid()
    og.build(photons[:, :3], radius)
    _check_build(g, og, n)
    rng = np.random.default_rng(1)
    queries = np.concatenate([photons[rng.integers(0, n, 2000), :3] + 0.5 * radius * rng.normal(size=(2000, 3)),
                              rng.uniform(-0.2, 1.2, (300, 3))]).astype(np.float32)
    nonempty = _check_queries(g, og, queries, k)
    if n > 100:
        assert nonempty > 500


@pytest.mark.parametrize("k", [65, 100, 1000, 4096])
def test_query_large_k(gpu, k):
    """k above 64: dense clustered photons, so that lists fill up and the k-th distance prunes
    (a larger radius for the largest k)."""
    n, radius = 200_000, 0.03 if k <= 1000 else 0.06
    photons = _photons(n, seed=77, clustered=True)
    g, _ = _grid(photons, radius)
    og = oracle.Grid()
    og.build(photons[:, :3], radius)
    rng = np.random.default_rng(2)
    queries = np.concatenate([photons[rng.integers(0, n, 300), :3] + 0.3 * radius * rng.normal(size=(300, 3)),
                              rng.uniform(-0.2, 1.2, (30, 3))]).astype(np.float32)
    assert _check_queries(g, og, queries, k) > 200
    _, _, cnt = _query(g, queries, k)
    assert (cnt == k).sum() > 50  # full lists (the k-th distance prunes)


def test_query_k_limit(gpu):
    photons = _photons(100, seed=1)
    g, _ = _grid(photons, 0.1)
    qb = N.DeviceBuffer.from_array(photons[:4, :3].copy())
    out = N.DeviceBuffer(4 * 4097 * 4)
    cnt = N.DeviceBuffer(16)
    r = N.lib.imba_hashgrid_query(g.h, qb.ptr, 12, 0, 4, 4097, out.ptr, None, cnt.ptr, None)
    assert r == 1  # IMBA_ERR_INVALID_ARG: k above 4096


def test_table_grows_only_and_rebuild(gpu):
    """The bucket table keeps its largest size across builds (rangesearch.h:33-34)."""
    g = S.HashGrid()
    og = oracle.Grid()
    for n, seed, table in ((30_000, 1, 60_000), (5_000, 2, 60_000), (60_000, 3, 120_000), (100, 4, 120_000)):
        photons = _photons(n, seed)
        _grid(photons, 0.02, g=g)
        og.build(photons[:, :3], 0.02)
        assert g.info().table_size == og.g.table_size == table
        _check_build(g, og, n)
        queries = np.random.default_rng(seed).random((500, 3), dtype=np.float32)
        _check_queries(g, og, queries, 10)
    assert g.info().table_size == 120_000


def test_large_bucket_ordering(gpu):
    """Thousands of photons at one point (one bucket larger than the LDS sort) stay ordered by
    index, and a query returns the k smallest indices among equal distances."""
    n = 9000
    photons = np.zeros((n, 4), np.float32)
    photons[:, :3] = 0.5
    photons[::7, :3] = np.random.default_rng(0).random((len(photons[::7]), 3), dtype=np.float32)
    g, _ = _grid(photons, 0.01)
    og = oracle.Grid()
    og.build(photons[:, :3], 0.01)
    _check_build(g, og, n)
    _check_queries(g, og, np.full((4, 3), 0.5, np.float32), 64)


def test_gather_records(gpu):
    photons = _photons(10_000, 5)
    g, pb = _grid(photons, 0.03)
    out = N.DeviceBuffer(pb.nbytes)
    N.check(N.lib.imba_hashgrid_gather(g.h, pb.ptr, 16, out.ptr, None), "gather")
    got = out.download(np.float32, photons.size).reshape(photons.shape)
    og = oracle.Grid()
    og.build(photons[:, :3], 0.03)
    assert np.array_equal(got, photons[og.order()])


def test_reset_forgets_the_table_history(gpu):
    """After imba_hashgrid_reset a grid builds and answers as a fresh one (table 2N), while
    keeping its allocations."""
    big, small = _photons(20000, 31), _photons(3000, 32, clustered=True)
    g, _ = _grid(big, 0.05)
    assert g.info().table_size == 40000
    g.reset()
    g, pb = _grid(small, 0.04, g=g)
    fresh, _ = _grid(small, 0.04)
    assert g.info().table_size == fresh.info().table_size == 6000
    q = _photons(2000, 33)
    (ia, da, ca), (ib, db, cb) = _query(g, q, 10), _query(fresh, q, 10)
    assert np.array_equal(ca, cb) and ca.max() > 0
    for j in range(len(q)):
        assert np.array_equal(ia[j, :ca[j]], ib[j, :cb[j]])
        assert np.array_equal(da[j, :ca[j]].view(np.int32), db[j, :cb[j]].view(np.int32))
