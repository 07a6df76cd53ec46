"""The allocation-free compaction (imba_compact_hits_noalloc) against imba_compact_hits and the
oracle restatement of RayQueue::compact_hits (render/scheduling/ray_queue.h:161-185).

Two calls in a row share the library's static block-offset array; the second is smaller, so
stale offsets from the first call would show up in its results.  Each call is compared, bit
for bit over every output buffer, with imba_compact_hits on the same inputs and its hit prefix
with the oracle.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from imbatracer_amd import native as N

pytestmark = pytest.mark.gpu


def _queue(n, hit_frac, seed, state_words=4):
    rng = np.random.default_rng(seed)
    rays = np.zeros(n, N.RAY_DTYPE)
    rays["org"] = rng.random((n, 4), dtype=np.float32)
    rays["dir"] = rng.random((n, 4), dtype=np.float32)
    hits = np.zeros(n, N.HIT_DTYPE)
    is_hit = rng.random(n) < hit_frac
    hits["tri_id"] = np.where(is_hit, rng.integers(0, 1 << 30, n), -1)
    hits["inst_id"] = rng.integers(-1, 100, n)
    hits["tmax"] = rng.random(n, dtype=np.float32)
    hits["u"] = rng.random(n, dtype=np.float32)
    states = rng.integers(-5, 1 << 30, (n, state_words)).astype(np.int32)
    states[:, 0] = np.arange(n)
    return rays, hits, states


def _run(fn, rays, hits, states, stream=None):
    n = len(rays)
    rb, hb, sb = (N.DeviceBuffer.from_array(x) for x in (rays, hits, states))
    ro, ho, so = N.DeviceBuffer(max(rb.nbytes, 1)), N.DeviceBuffer(max(hb.nbytes, 1)), N.DeviceBuffer(max(sb.nbytes, 1))
    k = C.c_int(-1)
    N.check(fn(rb.ptr, hb.ptr, sb.ptr, 4 * states.shape[1], n, ro.ptr, ho.ptr, so.ptr, C.byref(k), stream), fn.__name__)
    return (k.value, ro.download(np.uint8, rb.nbytes), ho.download(np.uint8, hb.nbytes),
            so.download(np.uint8, sb.nbytes))


def _check_call(n, frac, seed, state_words=4):
    rays, hits, states = _queue(n, frac, seed, state_words)
    print("noalloc compaction: n=%d hit fraction %.2f seed %d" % (n, frac, seed), flush=True)
    got = _run(N.lib.imba_compact_hits_noalloc, rays, hits, states)
    want = _run(N.lib.imba_compact_hits, rays, hits, states)
    assert got[0] == want[0]
    for a, b in zip(got[1:], want[1:]):
        assert np.array_equal(a, b)
    k_ref, r_ref, h_ref, s_ref = oracle.compact_hits(rays, hits, states)
    k = got[0]
    assert k == k_ref
    assert np.array_equal(got[2][:k * 16], np.ascontiguousarray(h_ref[:k]).view(np.uint8).ravel())
    assert np.array_equal(got[1][:k * 32], np.ascontiguousarray(r_ref[:k]).view(np.uint8).ravel())
    assert np.array_equal(got[3][:k * 4 * state_words], np.ascontiguousarray(s_ref[:k]).view(np.uint8).ravel())


def test_noalloc_two_calls_second_smaller(gpu):
    _check_call(3_000_000, 0.61, seed=1)
    _check_call(70_001, 0.23, seed=2, state_words=10)


def test_noalloc_on_a_stream_without_read_back(gpu):
    n = 500_000
    rays, hits, states = _queue(n, 0.5, seed=9)
    st = N.Stream()
    rb, hb, sb = (N.DeviceBuffer.from_array(x) for x in (rays, hits, states))
    ro, ho, so = N.DeviceBuffer(rb.nbytes), N.DeviceBuffer(hb.nbytes), N.DeviceBuffer(sb.nbytes)
    N.check(N.lib.imba_compact_hits_noalloc(rb.ptr, hb.ptr, sb.ptr, 16, n, ro.ptr, ho.ptr, so.ptr, None, st.handle),
            "noalloc")
    st.sync()
    k_ref, _, h_ref, s_ref = oracle.compact_hits(rays, hits, states)
    assert np.array_equal(ho.download(N.HIT_DTYPE, k_ref).view(np.int32), h_ref[:k_ref].view(np.int32))
    assert np.array_equal(so.download(np.int32, 4 * k_ref).reshape(-1, 4), s_ref[:k_ref])


def test_noalloc_rejects_counts_beyond_its_array(gpu):
    k = C.c_int(-1)
    limit = 134_217_728
    r = N.lib.imba_compact_hits_noalloc(None, None, None, 16, limit + 1, None, None, None, C.byref(k), None)
    assert r == 5  # IMBA_ERR_CAPACITY
    assert b"exceeds" in N.lib.imba_last_error()
