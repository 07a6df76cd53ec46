"""The multi-queue bounce loop (workloads.QueueLoop: per bounce one batched traversal and one
batched shading step + compaction of every queue, captured in a graph) against the same loop
run queue by queue with the single-queue calls, whose every traversal is checked against the
oracle on the same rays: every queue's size at every bounce, its last hits and its final
states are equal bit for bit, for graph replays and eager steps alike."""
import ctypes as C

import numpy as np
import pytest

import oracle
from imbatracer_amd import native as N
from imbatracer_amd import scene as S
from imbatracer_amd import workloads as W

pytestmark = pytest.mark.gpu

SIZES = [0, 1, 63, 64, 65, 300, 2048, 2049, 5000, 10000, 777, 4096]
SB = W.PT_STATE_BYTES


@pytest.fixture(scope="module")
def setup(gpu):
    scenes, shading, geo, arrays = [], [], [], []
    for k, (n, builder) in enumerate([(2000, N.BUILDER_SBVH), (5000, N.BUILDER_BINNED_SAH), (20000, N.BUILDER_SBVH)]):
        b, v, i = W.random_tri_scene(n, builder, seed=0x99 + k)
        scenes.append(S.Scene(b.desc()))
        arrays.append(b.arrays())
        normals, mats = b.tri_attrs()
        nm = int(mats.max()) + 1
        shading.append(W.PtShadingData(normals, mats, np.full((nm, 3), 0.5, np.float32), np.zeros((nm, 3), np.float32),
                                       np.zeros(nm, np.float32), np.zeros((0, 12), np.float32)))
        geo.append((v, i))
    queues = []
    for k, n in enumerate(SIZES):
        v, i = geo[k % 3]
        rays = W.c4_rays(v, i, max(n, 1), coherent=k % 2 == 0, seed=700 + k)[:n] if n else np.zeros(0, N.RAY_DTYPE)
        queues.append((k % 3, N.DeviceBuffer.from_array(rays) if n else N.DeviceBuffer(32), n))
    return scenes, shading, queues, arrays


def _snapshot(loop):
    c = loop.counts.download(np.int32, loop.n_counts)
    out = []
    for e in loop.q:
        n = e["n"]
        out.append((e["hits"].download(np.uint8, 16 * n), e["states"][loop.L % 2].download(np.uint8, SB * n)))
    return c, out


def _per_queue(loop, scenes, arrays):
    """The loop queue by queue with the single-queue calls; every traversal's hits are checked
    against the oracle on the rays it traversed.  Returns the number of rays checked."""
    st = loop.stream.handle
    ws = N.DeviceBuffer(int(N.lib.imba_compact_workspace_size(max(max(e["n"] for e in loop.q), 1))))
    N.check(N.lib.imba_memcpy_d2d(loop.counts.ptr, loop.counts0.ptr, 4 * loop.n_counts, st), "counts")
    checked = 0
    for k, e in enumerate(loop.q):
        n = e["n"]
        N.check(N.lib.imba_memset(e["hits"].ptr, 0, e["hits"].nbytes, st), "memset")
        N.check(N.lib.imba_memset(e["states"][loop.L % 2].ptr, 0, e["states"][loop.L % 2].nbytes, st), "memset")
        if n == 0:
            continue
        sh = loop.sh[e["scene"]]
        N.check(N.lib.imba_memcpy_d2d(e["rays"][0].ptr, e["rays0"].ptr, 32 * n, st), "rays")
        N.check(N.lib.imba_memcpy_d2d(e["states"][0].ptr, e["states0"].ptr, SB * n, st), "states")
        for b in range(loop.L):
            cur, nxt = b % 2, (b + 1) % 2
            cnt = loop._cnt(k, b)
            N.check(N.lib.imba_intersect_dev(scenes[e["scene"]].h, e["rays"][cur].ptr, e["hits"].ptr, n, cnt, None, st),
                    "intersect_dev")
            m = int(loop.counts.download(np.int32, loop.n_counts, loop.stream)[k * (loop.L + 1) + b])
            if m:
                rays = e["rays"][cur].download(N.RAY_DTYPE, m, loop.stream)
                hits = e["hits"].download(N.HIT_DTYPE, m, loop.stream)
                want, bad = oracle.trace(arrays[e["scene"]], rays, threads=8)
                assert bad == 0
                assert np.array_equal(hits.view(np.int32), want.view(np.int32)), (k, b)
                checked += m
            N.check(N.lib.imba_pt_shade(C.byref(sh), e["rays"][cur].ptr, e["hits"].ptr, e["states"][cur].ptr, n, cnt,
                                        None, e["rays"][cur].ptr, None, None, st), "pt_shade")
            N.check(N.lib.imba_compact_rays_dev(e["rays"][cur].ptr, e["states"][cur].ptr, SB, n, cnt,
                                                e["rays"][nxt].ptr, e["states"][nxt].ptr, loop._cnt(k, b + 1), ws.ptr,
                                                st), "compact_dev")
    loop.stream.sync()
    return checked


def test_loop_equals_per_queue(setup):
    scenes, shading, queues, arrays = setup
    loop = W.QueueLoop(scenes, shading, queues, max_path_len=6, rr_start=2)
    loop.build_graph()
    checked = _per_queue(loop, scenes, arrays)
    assert checked > sum(SIZES)
    want_c, want = _snapshot(loop)
    assert want_c.reshape(len(SIZES), 7)[:, 0].tolist() == SIZES
    assert want_c.reshape(len(SIZES), 7)[:, 1:].sum() > 0
    for graph in (True, False, True):
        for e in loop.q:  # clear what the loop must rewrite
            N.check(N.lib.imba_memset(e["hits"].ptr, 0, e["hits"].nbytes, None), "memset")
            N.check(N.lib.imba_memset(e["states"][loop.L % 2].ptr, 0, e["states"][loop.L % 2].nbytes, None), "memset")
        loop.step(graph)
        loop.stream.sync()
        got_c, got = _snapshot(loop)
        assert np.array_equal(got_c, want_c)
        for k, ((gh, gs), (wh, ws)) in enumerate(zip(got, want)):
            assert np.array_equal(gh, wh)
            # the final compacted states: only the survivors' prefix is defined
            m = int(got_c.reshape(len(SIZES), 7)[k, loop.L])
            assert np.array_equal(gs[:SB * m], ws[:SB * m])
