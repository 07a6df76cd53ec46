"""Batched bounce + compaction (imba_shade_batch_*) against the single-queue calls.

Several queues over different scenes, of different sizes (empty, smaller than one chunk,
spanning many compaction tiles), some with their size on the device and below the capacity:
after one batched launch every queue's bounced rays and states, its compacted survivors, its
new device count and its image equal those of imba_pt_bounce_dev + imba_compact_rays_dev run
on that queue alone, bit for bit.  The launch is also replayed from a graph.
"""
import numpy as np
import pytest

from imbatracer_amd import native as N
from imbatracer_amd import scene as S
from imbatracer_amd import workloads as W

pytestmark = pytest.mark.gpu

STATE = W.STATE_BYTES


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
        self.tri_n = N.DeviceBuffer.from_array(scene.tri_normals)
        self.tri_m = N.DeviceBuffer.from_array(scene.tri_materials)
        self.albedo = N.DeviceBuffer.from_array(scene.albedo)
        self.emission = N.DeviceBuffer.from_array(scene.emission)
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
        return r, s, ro, so, img, nxt

    def item(self, bufs):
        r, s, ro, so, img, nxt = bufs
        return {"rays": r, "hits": self.hits, "states": s, "capacity": self.n, "count_dev": self.cnt,
                "rays_out": ro, "states_out": so, "next_count_dev": nxt, "tri_normals": self.tri_n,
                "tri_materials": self.tri_m, "tri_count": self.sc.tri_count, "albedo": self.albedo,
                "emission": self.emission, "material_count": len(self.sc.albedo), "offset": self.sc.offset,
                "image": img}

    def single(self, bufs, max_len, rr):
        r, s, ro, so, img, nxt = bufs
        if self.n == 0:
            return
        N.check(N.lib.imba_pt_bounce_dev(r.ptr, self.hits.ptr, s.ptr, self.n, self.cnt.ptr, self.tri_n.ptr,
                                         self.tri_m.ptr, self.sc.tri_count, self.albedo.ptr, self.emission.ptr,
                                         len(self.sc.albedo), max_len, rr, self.sc.offset, img.ptr, r.ptr, None),
                "bounce")
        ws = N.DeviceBuffer(int(N.lib.imba_compact_workspace_size(self.n)))
        N.check(N.lib.imba_compact_rays_dev(r.ptr, s.ptr, STATE, self.n, self.cnt.ptr, ro.ptr, so.ptr, nxt.ptr,
                                            ws.ptr, None), "compact")

    def host(self, bufs):
        r, s, ro, so, img, nxt = bufs
        n = max(self.n, 1)
        c = int(nxt.download(np.int32, 1)[0])
        return (r.download(np.uint8, n * 32), s.download(np.uint8, n * STATE), c,
                ro.download(np.uint8, max(c, 0) * 32), so.download(np.uint8, max(c, 0) * STATE),
                img.download(np.uint8, img.nbytes))


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
    batch = S.ShadeBatch([q.item(b) for q, b in zip(queues, bufs)])
    batch.launch(max_len, rr)
    N.check(N.lib.imba_stream_sync(None), "sync")
    for q, b, w in zip(queues, bufs, want):
        g = q.host(b)
        if q.n == 0:
            assert g[2] == 0
            continue
        assert g[2] == w[2]
        for x, y in zip(g, w):
            if isinstance(x, np.ndarray):
                assert np.array_equal(x, y)


def test_graph_replays(queues):
    st = N.Stream()
    bufs = [q.fresh() for q in queues]
    batch = S.ShadeBatch([q.item(b) for q, b in zip(queues, bufs)])
    batch.launch(8, 3, st)
    st.sync()
    first = [q.host(b) for q, b in zip(queues, bufs)]
    g = S.Graph.begin(st)
    try:
        batch.launch(8, 3, st)
    finally:
        graph = S.Graph.end(st)
    for _ in range(3):
        # restore the inputs (rays and states were bounced in place), then replay
        for q, b in zip(queues, bufs):
            if q.n:
                N.check(N.lib.imba_memcpy_d2d(b[0].ptr, q.rays0.ptr, q.n * 32, None), "copy")
                N.check(N.lib.imba_memcpy_d2d(b[1].ptr, q.states0.ptr, q.n * STATE, None), "copy")
                N.check(N.lib.imba_memset(b[4].ptr, 0, b[4].nbytes, None), "memset")
        graph.launch(st)
        st.sync()
        again = [q.host(b) for q, b in zip(queues, bufs)]
        for x, y in zip(again, first):
            assert x[2] == y[2]
            for u, v in zip(x, y):
                if isinstance(u, np.ndarray):
                    assert np.array_equal(u, v)
    del g
