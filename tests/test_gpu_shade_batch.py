"""Batched shading + compaction (imba_shade_batch_*) against the single-queue calls.

Several queues over different scenes, of different sizes (empty, smaller than one chunk,
spanning many compaction tiles), some with their size on the device and below the capacity:
after one batched launch every queue's shaded rays and states, its shadow slots, its compacted
survivors and its new device count equal those of imba_pt_shade + imba_compact_rays_dev run on
that queue alone, bit for bit; after the shadow rays' any-hit traversal, imba_shade_batch_shadow
gives every queue the image of imba_pt_shadow (within 1e-6: a pixel's samples are added with
float atomics in no fixed order).  The launch is also replayed from a graph.
"""
import ctypes as C

import numpy as np
import pytest

from imbatracer_amd import native as N
from imbatracer_amd import scene as S
from imbatracer_amd import workloads as W

pytestmark = pytest.mark.gpu

STATE = W.PT_STATE_BYTES


class Queue:
    def __init__(self, scene, width, height, spp, count, seed):
        self.sc, self.n = scene, width * height * spp
        self.w, self.h = width, height
        self.count = count  # entries in use (<= capacity)
        self.gpu = S.Scene(scene.desc)
        pos, d, up, fov = scene.camera
        self.pos = np.asarray(pos, np.float32)
        self.r2w = np.zeros(16, np.float32)
        dd, uu = np.asarray(d, np.float32), np.asarray(up, np.float32)
        if self.n:
            N.check(N.lib.imba_gen_camera_matrix(N.ptr(self.pos), N.ptr(dd), N.ptr(uu), fov, width, height,
                                                 N.ptr(self.r2w)), "matrix")
        n = max(self.n, 1)
        self.rays0, self.states0 = N.DeviceBuffer(n * 32), N.DeviceBuffer(n * STATE)
        self.hits = N.DeviceBuffer(n * 16)
        if self.n:
            N.check(N.lib.imba_pt_camera(N.ptr(self.r2w), N.ptr(self.pos), width, height, spp, seed, 0, self.n,
                                         self.rays0.ptr, self.states0.ptr, None), "camera")
            N.check(N.lib.imba_intersect(self.gpu.h, self.rays0.ptr, self.hits.ptr, self.n, None, None), "trace")
        self.shading = W.PtShadingData(scene.tri_normals, scene.tri_materials, scene.albedo, scene.emission,
                                       scene.emitter_area, scene.lights)
        self.cnt = N.DeviceBuffer(8)
        self.cnt.upload(np.array([count, -1], np.int32))

    def fresh(self):
        """Copies of the traced queue: rays, states, outputs, image."""
        n = max(self.n, 1)
        r, s = N.DeviceBuffer(n * 32), N.DeviceBuffer(n * STATE)
        if self.n:
            N.check(N.lib.imba_memcpy_d2d(r.ptr, self.rays0.ptr, self.n * 32, None), "copy")
            N.check(N.lib.imba_memcpy_d2d(s.ptr, self.states0.ptr, self.n * STATE, None), "copy")
        ro, so = N.DeviceBuffer(n * 32), N.DeviceBuffer(n * STATE)
        img = N.DeviceBuffer(max(self.w * self.h * 12, 4))
        N.check(N.lib.imba_memset(img.ptr, 0, img.nbytes, None), "memset")
        nxt = N.DeviceBuffer(4)
        N.check(N.lib.imba_memset(nxt.ptr, 0xFF, 4, None), "memset")
        sr, ss, shh = N.DeviceBuffer(n * 32), N.DeviceBuffer(n * 16), N.DeviceBuffer(n * 16)
        return r, s, ro, so, img, nxt, sr, ss, shh

    def item(self, bufs, max_len, rr):
        r, s, ro, so, img, nxt, sr, ss, shh = bufs
        return {"rays": r, "hits": self.hits, "states": s, "capacity": self.n, "count_dev": self.cnt,
                "rays_out": ro, "states_out": so, "next_count_dev": nxt, "shadow_rays": sr, "shadow_states": ss,
                "shadow_hits": shh, "image": img, "shading": self.shading.struct(max_len, rr)}

    def single(self, bufs, max_len, rr):
        r, s, ro, so, img, nxt, sr, ss, shh = bufs
        if self.n == 0:
            return
        sh = self.shading.struct(max_len, rr)
        N.check(N.lib.imba_pt_shade(C.byref(sh), r.ptr, self.hits.ptr, s.ptr, self.n, self.cnt.ptr, img.ptr, r.ptr,
                                    sr.ptr, ss.ptr, None), "shade")
        self.occluded(bufs)
        N.check(N.lib.imba_pt_shadow(shh.ptr, ss.ptr, self.n, self.cnt.ptr, img.ptr, None), "shadow")
        ws = N.DeviceBuffer(int(N.lib.imba_compact_workspace_size(self.n)))
        N.check(N.lib.imba_compact_rays_dev(r.ptr, s.ptr, STATE, self.n, self.cnt.ptr, ro.ptr, so.ptr, nxt.ptr,
                                            ws.ptr, None), "compact")

    def occluded(self, bufs, stream=None):
        r, s, ro, so, img, nxt, sr, ss, shh = bufs
        if self.n:
            N.check(N.lib.imba_occluded_dev(self.gpu.h, sr.ptr, shh.ptr, self.n, self.cnt.ptr, None, N._s(stream)),
                    "occluded")

    def host(self, bufs):
        r, s, ro, so, img, nxt, sr, ss, shh = bufs
        n = max(self.n, 1)
        c = int(nxt.download(np.int32, 1)[0])
        m = min(max(int(self.cnt.download(np.int32, 1)[0]), 0), self.n)
        return (r.download(np.uint8, n * 32), s.download(np.uint8, n * STATE), c,
                ro.download(np.uint8, max(c, 0) * 32), so.download(np.uint8, max(c, 0) * STATE),
                sr.download(np.uint8, m * 32), ss.download(np.uint8, m * 16), shh.download(np.uint8, m * 16),
                img.download(np.float32, img.nbytes // 4))


def _same(g, w):
    """Everything bit for bit but the image (last element), which is compared within 1e-6."""
    assert g[2] == w[2]
    for x, y in zip(g[:-1], w[:-1]):
        if isinstance(x, np.ndarray):
            assert np.array_equal(x, y)
    assert np.allclose(g[-1], w[-1], rtol=1e-6, atol=0)


@pytest.fixture(scope="module")
def queues(gpu):
    a = W.SphereGridScene(side=3, level=2, builder=N.BUILDER_SBVH)
    b = W.SphereGridScene(side=5, level=1, builder=N.BUILDER_BINNED_SAH, seed=0x77)
    return [Queue(a, 64, 48, 2, 64 * 48 * 2, 11),      # many tiles
            Queue(b, 16, 8, 1, 100, 12),               # device count below the capacity
            Queue(a, 0, 0, 1, 0, 13),                  # empty
            Queue(b, 5, 7, 2, 70, 14),                 # less than one chunk
            Queue(a, 40, 30, 1, 0, 15),                # device count 0
            Queue(b, 96, 64, 1, 96 * 64, 16)]


@pytest.mark.parametrize("max_len,rr", [(8, 3), (2, 0)])
def test_batch_equals_single_queues(queues, max_len, rr):
    want, got = [], []
    for q in queues:
        bufs = q.fresh()
        q.single(bufs, max_len, rr)
        want.append(q.host(bufs))
    bufs = [q.fresh() for q in queues]
    batch = S.ShadeBatch([q.item(b, max_len, rr) for q, b in zip(queues, bufs)])
    batch.launch()
    for q, b in zip(queues, bufs):
        q.occluded(b)
    batch.shadow()
    N.check(N.lib.imba_stream_sync(None), "sync")
    lit = 0
    for q, b, w in zip(queues, bufs, want):
        g = q.host(b)
        if q.n == 0:
            assert g[2] == 0
            continue
        _same(g, w)
        lit += int((g[-1] > 0).sum())
    assert lit > 100  # the shadow rays and emitter hits did add light


def test_graph_replays(queues):
    st = N.Stream()
    bufs = [q.fresh() for q in queues]
    batch = S.ShadeBatch([q.item(b, 8, 3) for q, b in zip(queues, bufs)])

    def step():
        batch.launch(st)
        for q, b in zip(queues, bufs):
            q.occluded(b, st)
        batch.shadow(st)

    step()
    st.sync()
    first = [q.host(b) for q, b in zip(queues, bufs)]
    g = S.Graph.begin(st)
    try:
        step()
    finally:
        graph = S.Graph.end(st)
    for _ in range(3):
        # restore the inputs (rays and states were shaded in place), then replay
        for q, b in zip(queues, bufs):
            if q.n:
                N.check(N.lib.imba_memcpy_d2d(b[0].ptr, q.rays0.ptr, q.n * 32, None), "copy")
                N.check(N.lib.imba_memcpy_d2d(b[1].ptr, q.states0.ptr, q.n * STATE, None), "copy")
                N.check(N.lib.imba_memset(b[4].ptr, 0, b[4].nbytes, None), "memset")
        graph.launch(st)
        st.sync()
        again = [q.host(b) for q, b in zip(queues, bufs)]
        for x, y in zip(again, first):
            _same(x, y)
    del g
