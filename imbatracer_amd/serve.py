"""A request loop for path-traced frames on one GPU, with continuous cross-session batching.

Requests (scene, camera, resolution, spp, seed) arrive over time.  A request is admitted when
its buffers fit in a fixed device arena (admission by HBM footprint, first come first served)
and a counter slot is free; its queues then live in the arena until it completes.  Every step
of the loop:
  * admitted requests get their camera rays (the reference's per-pixel seeding);
  * the current queue of every active request, whatever its scene and bounce, is traversed:
    the short ones together in one batched launch (imba_trace_batch_update +
    imba_trace_batch_launch), so requests that arrived at different times share launches, and
    each queue long enough to fill the GPU alone in a launch of its own;
  * the requests' shading steps (emitter hits, one next-event shadow ray per surface hit,
    Lambertian bounce) and compactions follow in one batched launch (imba_shade_batch_*) with
    the queue sizes kept on the device, then the shadow rays' any-hit traversal (batched the
    same way) and their contributions (imba_shade_batch_shadow), and one read-back of all the
    requests' queue sizes ends the step.
Photon-mapping (PPM) and vertex connection and merging (VCM) requests advance through their
phases one per step; their light and camera queues join the step's batched traversal.
A request completes when its queue is empty or its paths reach the maximum length; its latency
is completion minus arrival.  The plain per-request path (`batched=False`) runs the same
kernels for one request at a time, in arrival order.

A request's image does not depend on the other requests it shares launches with: its rays,
seeds and per-ray results are its own (the image is a sum of float atomics, so equal up to the
order of additions).
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np

from . import native as N
from . import scene as S
from . import vcm as VC
from .workloads import PPM_STATE_BYTES, PT_STATE_BYTES, SHADOW_STATE_BYTES, PtShadingData

ALIGN = 256


class DeviceArena:
    """First-fit allocator of byte ranges inside one device allocation."""

    def __init__(self, nbytes: int, buf=None):
        self.buf = buf if buf is not None else N.DeviceBuffer(nbytes)
        self.size = nbytes
        self.free = [(0, nbytes)]  # sorted (offset, length)

    def alloc(self, nbytes: int):
        nbytes = (nbytes + ALIGN - 1) // ALIGN * ALIGN
        for k, (off, ln) in enumerate(self.free):
            if ln >= nbytes:
                if ln == nbytes:
                    self.free.pop(k)
                else:
                    self.free[k] = (off + nbytes, ln - nbytes)
                return off, nbytes
        return None

    def release(self, off: int, nbytes: int):
        self.free.append((off, nbytes))
        self.free.sort()
        merged = []
        for o, ln in self.free:
            if merged and merged[-1][0] + merged[-1][1] == o:
                merged[-1] = (merged[-1][0], merged[-1][1] + ln)
            else:
                merged.append((o, ln))
        self.free = merged

    def ptr(self, off: int):
        return C.c_void_p(self.buf.ptr.value + off)


class ServedScene:
    """A resident scene with the shading data of the path tracer (per-triangle normals and
    materials, per-material albedo, emission and emitter area, the triangle lights) and of the
    photon mapper."""

    def __init__(self, scene, cfg: N.TraceConfig | None = None):
        self.s = scene
        self.gpu = S.Scene(scene.desc)
        self.shading = PtShadingData(scene.tri_normals, scene.tri_materials, scene.albedo, scene.emission,
                                     scene.emitter_area, scene.lights)
        self.tri_n = N.DeviceBuffer.from_array(scene.tri_normals)
        self.tri_m = N.DeviceBuffer.from_array(scene.tri_materials)
        self.albedo = N.DeviceBuffer.from_array(scene.albedo)
        self.emission = N.DeviceBuffer.from_array(scene.emission)
        self.cfg = cfg
        # photon mapping: material kinds, the emitting triangle, the merge radius
        self.kinds = N.DeviceBuffer.from_array(np.ascontiguousarray(scene.kinds, np.int32))
        self.light_tri = np.ascontiguousarray(scene.light_tri, np.float32).reshape(-1)
        self.light_int = np.ascontiguousarray(scene.emission[len(scene.albedo) - 1], np.float32)
        lo, hi = np.asarray(scene.bounds[0], np.float64), np.asarray(scene.bounds[1], np.float64)
        self.radius = float(np.float32(0.005 * np.linalg.norm(hi - lo)))


@dataclass
class Request:
    rid: int
    session: int
    scene: int
    width: int
    height: int
    spp: int
    cam_pos: tuple
    cam_dir: tuple
    seed: int
    arrival: float  # seconds after the start of the run
    integrator: str = "pt"  # "pt", "ppm" (photon mapping: the PPM mode of VCM) or "vcm"
    # filled in by the server
    admitted: float = -1.0
    done: float = -1.0
    rays: int = 0
    counts: list = field(default_factory=list)

    @property
    def samples(self) -> int:
        return self.width * self.height * self.spp


PPM_DEPTH = 6   # light and camera path hits of a photon-mapping request
PPM_K = 10      # photons per merge query
VCM_DEPTH = 6   # light path hits and camera path vertices of a VCM request
VCM_K = 10      # light vertices per merge query
VCM_C = 1       # vertex connections per camera vertex
VCM_RR_START = 0  # Russian roulette from the first bounce, as the reference


def first_rays(req: Request) -> int:
    """Rays of a request's first traversal: its camera samples (PT), its light paths plus its
    camera samples (PPM, VCM: the light pass, then the camera pass)."""
    if req.integrator == "pt":
        return req.samples
    return req.width * req.height + req.samples


def footprint(req: Request) -> int:
    """Device bytes a request holds while active.  PT: two ray and state buffers, hits, the
    shadow slots (ray, state, hit), image and the compaction workspace.  PPM: two ray and state
    buffers and hits for max(light paths, samples) entries, the photon slots and the compacted
    photons, the merge queries and their results, image and workspace.  VCM: two ray and state
    buffers, hits and the shadow slots (and their compacted copies) for max(light paths,
    samples) entries, the light vertex slots, the cache and its merge records, the merge queries,
    image and workspace."""
    n = req.samples
    if req.integrator == "pt":
        ws = int(N.lib.imba_compact_workspace_size(n))
        return (2 * 32 * n + 2 * PT_STATE_BYTES * n + 16 * n + (32 + SHADOW_STATE_BYTES + 16) * n +
                12 * req.width * req.height + ws + 12 * ALIGN)
    p = req.width * req.height
    cap = max(p, n)
    if req.integrator == "vcm":
        slots = VCM_DEPTH * p
        ws = int(N.lib.imba_compact_workspace_size(max(slots, (1 + VCM_C) * cap)))
        return (2 * 32 * cap + 2 * VC.VCM_STATE_DTYPE.itemsize * cap + 16 * cap +
                (1 + VCM_C) * cap * (32 + 16 + 16 + 32 + 16) + (4 * 32 + 32) * slots + 64 * cap +
                12 * req.width * req.height + ws + 24 * ALIGN)
    slots = PPM_DEPTH * p
    ws = int(N.lib.imba_compact_workspace_size(max(slots, cap)))
    return (2 * 32 * cap + 2 * PPM_STATE_BYTES * cap + 16 * cap + 2 * 48 * slots + 64 * n + 8 * PPM_K * n + 4 * n +
            12 * req.width * req.height + ws + 16 * ALIGN)


class _Job:
    def __init__(self, req: Request, slot: int, block, scene: ServedScene, arena: DeviceArena):
        self.req, self.slot, self.block, self.sc = req, slot, block, scene
        n = req.samples
        self.n = n
        off = block[0]

        def take(nbytes):
            nonlocal off
            p = arena.ptr(off)
            off += (nbytes + ALIGN - 1) // ALIGN * ALIGN
            return p

        self.rays = [take(32 * n), take(32 * n)]
        self.states = [take(PT_STATE_BYTES * n), take(PT_STATE_BYTES * n)]
        self.hits = take(16 * n)
        self.shadow_rays, self.shadow_states, self.shadow_hits = take(32 * n), take(SHADOW_STATE_BYTES * n), take(16 * n)
        self.image = take(12 * req.width * req.height)
        self.workspace = take(int(N.lib.imba_compact_workspace_size(n)))
        self.bounce = 0
        self.cur = 0
        self.queued = n  # rays in the current queue (as of the last read-back)
        self.r2w = np.zeros(16, np.float32)
        self.pos = np.asarray(req.cam_pos, np.float32)
        d = np.asarray(req.cam_dir, np.float32)
        up = np.asarray((0.0, 1.0, 0.0), np.float32)
        N.check(N.lib.imba_gen_camera_matrix(N.ptr(self.pos), N.ptr(d), N.ptr(up), 60.0, req.width, req.height,
                                             N.ptr(self.r2w)), "camera matrix")


class _PPMJob:
    """A photon-mapping request (the PPM mode of the VCM integrator, vcm.cpp) as a sequence of
    phases advanced one per step: `light` (PPM_DEPTH bounces of width x height light paths,
    photons into slots), `photons` (slots compacted; the count is read back), `grid` (hash grid
    built, camera rays shot), `camera` (bounces until every path has written its merge query),
    `merge` (k-nearest queries and merging), then finished.  The light and camera queues join
    the step's batched traversal like PT queues.  Hash grids come from a pool of the server's
    (reset before use, so a request's table size is its own); their arrays live outside the
    arena."""

    def __init__(self, req: Request, slot: int, block, scene: ServedScene, arena: DeviceArena):
        self.req, self.slot, self.block, self.sc = req, slot, block, scene
        self.P = req.width * req.height
        self.n = req.samples
        self.cap = max(self.P, self.n)
        self.L = PPM_DEPTH
        off = block[0]

        def take(nbytes):
            nonlocal off
            p = arena.ptr(off)
            off += (nbytes + ALIGN - 1) // ALIGN * ALIGN
            return p

        cap, slots = self.cap, self.L * self.P
        self.rays = [take(32 * cap), take(32 * cap)]
        self.states = [take(PPM_STATE_BYTES * cap), take(PPM_STATE_BYTES * cap)]
        self.hits = take(16 * cap)
        self.slot_rays, self.slot_states = take(32 * slots), take(16 * slots)
        self.photon_rays, self.photon_states = take(32 * slots), take(16 * slots)
        self.queries = take(64 * self.n)
        self.out_idx, self.out_dist, self.out_count = take(4 * PPM_K * self.n), take(4 * PPM_K * self.n), take(4 * self.n)
        self.image = take(12 * req.width * req.height)
        self.workspace = take(int(N.lib.imba_compact_workspace_size(max(slots, cap))))
        self.grid = None  # taken from the server's pool when the photon count is known
        self.phase = "light"
        self.bounce = 0
        self.cur = 0
        self.queued = self.P
        self.photons = 0
        self.r2w = np.zeros(16, np.float32)
        self.pos = np.asarray(req.cam_pos, np.float32)
        d = np.asarray(req.cam_dir, np.float32)
        up = np.asarray((0.0, 1.0, 0.0), np.float32)
        N.check(N.lib.imba_gen_camera_matrix(N.ptr(self.pos), N.ptr(d), N.ptr(up), 60.0, req.width, req.height,
                                             N.ptr(self.r2w)), "camera matrix")

    # count slots of the request's row: light bounce b, camera bounce b, photon slots, photons
    def _light(self, srv, b):
        return srv.counts_ptr(self.slot, b)

    def _camera(self, srv, b):
        return srv.counts_ptr(self.slot, srv.L1 + b)

    def start(self, srv, st):
        c0 = np.zeros(srv.cs, np.int32)
        c0[0] = self.P
        c0[srv.L1] = self.n
        c0[2 * srv.L1] = self.L * self.P
        self.c0 = c0  # kept until the copy has run (the step ends with a stream synchronisation)
        N.check(N.lib.imba_memcpy_h2d(srv.counts_ptr(self.slot, 0), N.ptr(c0), c0.nbytes, st), "counts")
        N.check(N.lib.imba_memset(self.image, 0, 12 * self.req.width * self.req.height, st), "image")
        N.check(N.lib.imba_ppm_light(N.ptr(self.sc.light_tri), N.ptr(self.sc.light_int),
                                     (self.req.seed ^ 0x5A5A5A5A) & 0xFFFFFFFF, 0, self.P, self.rays[0],
                                     self.states[0], st), "ppm_light")

    def tracing(self) -> bool:
        return self.phase in ("light", "camera")

    def trace_item(self, srv):
        cnt = self._light(srv, self.bounce) if self.phase == "light" else self._camera(srv, self.bounce)
        cap = self.P if self.phase == "light" else self.n
        return (self.sc.gpu, self.rays[self.cur], self.hits, cap, cnt, False)

    def before_trace(self, srv, st):
        sc, req = self.sc, self.req
        if self.phase == "grid":
            self.grid = srv.take_grid()
            if self.photons > 0:
                N.check(N.lib.imba_hashgrid_build(self.grid.h, self.photon_rays, 32, 0, self.photons, sc.radius, st),
                        "grid build")
            N.check(N.lib.imba_ppm_camera(N.ptr(self.r2w), N.ptr(self.pos), req.width, req.height, req.spp, req.seed,
                                          0, self.n, self.rays[0], self.states[0], self.queries, st), "ppm_camera")
            self.phase, self.bounce, self.cur = "camera", 0, 0
        elif self.phase == "merge":
            if self.photons > 0:
                N.check(N.lib.imba_hashgrid_query(self.grid.h, self.queries, 64, 0, self.n, PPM_K, self.out_idx,
                                                  self.out_dist, self.out_count, st), "grid query")
                N.check(N.lib.imba_ppm_merge(self.queries, self.n, PPM_K, self.out_idx, self.out_dist, self.out_count,
                                             self.photon_rays, self.photon_states, sc.radius, self.P, self.image, st),
                        "merge")
            self.phase = "finished"

    def after_trace(self, srv, st):
        sc, b, cur = self.sc, self.bounce, self.cur
        if self.phase == "light":
            cnt = self._light(srv, b)
            N.check(N.lib.imba_ppm_light_bounce(self.rays[cur], self.hits, self.states[cur], self.P, cnt, sc.tri_n.ptr,
                                                sc.tri_m.ptr, sc.s.tri_count, sc.albedo.ptr, sc.kinds.ptr,
                                                len(sc.s.albedo), self.L, srv.rr_start,
                                                C.c_void_p(self.slot_rays.value + 32 * b * self.P),
                                                C.c_void_p(self.slot_states.value + 16 * b * self.P), self.rays[cur],
                                                st), "ppm_light_bounce")
            N.check(N.lib.imba_compact_rays_dev(self.rays[cur], self.states[cur], PPM_STATE_BYTES, self.P, cnt,
                                                self.rays[1 - cur], self.states[1 - cur], self._light(srv, b + 1),
                                                self.workspace, st), "compact")
            self.bounce, self.cur = b + 1, 1 - cur
            if self.bounce == self.L:
                N.check(N.lib.imba_compact_rays_dev(self.slot_rays, self.slot_states, 16, self.L * self.P,
                                                    srv.counts_ptr(self.slot, 2 * srv.L1), self.photon_rays,
                                                    self.photon_states, srv.counts_ptr(self.slot, 2 * srv.L1 + 1),
                                                    self.workspace, st), "compact photons")
                self.phase = "photons"
        elif self.phase == "camera":
            cnt = self._camera(srv, b)
            N.check(N.lib.imba_ppm_camera_bounce(self.rays[cur], self.hits, self.states[cur], self.n, cnt, sc.tri_n.ptr,
                                                 sc.tri_m.ptr, sc.s.tri_count, sc.albedo.ptr, sc.emission.ptr,
                                                 sc.kinds.ptr, len(sc.s.albedo), self.L, srv.rr_start, self.req.spp,
                                                 self.queries, self.image, self.rays[cur], st), "ppm_camera_bounce")
            N.check(N.lib.imba_compact_rays_dev(self.rays[cur], self.states[cur], PPM_STATE_BYTES, self.n, cnt,
                                                self.rays[1 - cur], self.states[1 - cur], self._camera(srv, b + 1),
                                                self.workspace, st), "compact")
            self.bounce, self.cur = b + 1, 1 - cur

    def update(self, srv, row) -> bool:
        """Advances the phase from the read-back counts; True when the request is complete."""
        if self.phase == "light":
            self.queued = int(row[self.bounce])
        elif self.phase == "photons":
            self.photons = int(row[2 * srv.L1 + 1])
            self.phase = "grid"
        elif self.phase == "camera":
            self.queued = int(row[srv.L1 + self.bounce])
            if self.queued == 0 or self.bounce >= self.L:
                self.phase = "merge"
        elif self.phase == "finished":
            light = [int(x) for x in row[:self.L]]
            cam = [int(x) for x in row[srv.L1:srv.L1 + self.L]]
            self.req.counts = light + cam
            self.req.rays = int(sum(light) + sum(cam))
            return True
        return False


class _VCMJob:
    """A VCM request (one iteration of VCMIntegrator<ALGO_VCM>, vcm.cpp, as
    imbatracer_amd.vcm.VcmRenderer renders it: width x height light paths, VCM_DEPTH hits and
    vertices, one connection and a k = VCM_K merge per camera vertex) as phases advanced one per
    step: `light` (VCM_DEPTH bounces: light vertices into slots, camera connections traced and
    added), `cache` (slots compacted; the cache size is read back), `grid` (hash grid and merge
    records built, camera rays shot), `camera` (bounces with next event, connections and the
    fused merge) until every path has ended.  The light and camera queues join the step's
    batched closest-hit traversal and the bounce's shadow rays its batched any-hit traversal."""

    def __init__(self, req: Request, slot: int, block, scene: ServedScene, arena: DeviceArena):
        self.req, self.slot, self.block, self.sc = req, slot, block, scene
        self.P = req.width * req.height
        self.n = req.samples
        self.cap = max(self.P, self.n)
        self.L = VCM_DEPTH
        off = block[0]

        def take(nbytes):
            nonlocal off
            p = arena.ptr(off)
            off += (nbytes + ALIGN - 1) // ALIGN * ALIGN
            return p

        cap, slots = self.cap, self.L * self.P
        sb = VC.VCM_STATE_DTYPE.itemsize
        self.rays = [take(32 * cap), take(32 * cap)]
        self.states = [take(sb * cap), take(sb * cap)]
        self.hits = take(16 * cap)
        ns = (1 + VCM_C) * cap
        self.shadow_rays, self.shadow_states, self.shadow_hits = take(32 * ns), take(16 * ns), take(16 * ns)
        # the occupied shadow slots, stable-compacted before their traversal (as VcmRenderer does)
        self.shadow_rays_c, self.shadow_states_c = take(32 * ns), take(16 * ns)
        self.slot_a, self.slot_b = take(32 * slots), take(32 * slots)
        self.vtx_a, self.vtx_b = take(32 * slots), take(32 * slots)
        self.records = take(32 * slots)
        self.queries = take(64 * cap)
        self.image = take(12 * req.width * req.height)
        self.workspace = take(int(N.lib.imba_compact_workspace_size(max(slots, ns))))
        self.grid = None  # taken from the server's pool when the cache size is known
        self.phase = "light"
        self.bounce = 0
        self.cur = 0
        self.queued = self.P
        self.vertices = 0
        r2w, w2r, fwd, ipd = VC.camera_frame(req.cam_pos, req.cam_dir, (0.0, 1.0, 0.0), 60.0, req.width, req.height)
        self.frame = VC.make_frame(scene.shading, scene.kinds, (np.asarray(req.cam_pos, np.float32), r2w, w2r, fwd, ipd),
                                   req.width, req.height, req.spp, self.P, scene.radius, VCM_DEPTH, VCM_DEPTH,
                                   VCM_RR_START, VCM_C, "vcm")

    # count slots of the request's row: light bounce b, camera bounce b, vertex slots, cache
    # size, shadow slots of the camera bounce
    def _light(self, srv, b):
        return srv.counts_ptr(self.slot, b)

    def _camera(self, srv, b):
        return srv.counts_ptr(self.slot, srv.L1 + b)

    def _cache(self, srv):
        return srv.counts_ptr(self.slot, 2 * srv.L1 + 1)

    def _shadow(self, srv):
        return srv.counts_ptr(self.slot, 2 * srv.L1 + 2)

    def _live_shadow(self, srv):
        return srv.counts_ptr(self.slot, 2 * srv.L1 + 3)

    def start(self, srv, st):
        c0 = np.zeros(srv.cs, np.int32)
        c0[0] = self.P
        c0[srv.L1] = self.n
        c0[2 * srv.L1] = self.L * self.P
        self.c0 = c0  # kept until the copy has run (the step ends with a stream synchronisation)
        N.check(N.lib.imba_memcpy_h2d(srv.counts_ptr(self.slot, 0), N.ptr(c0), c0.nbytes, st), "counts")
        N.check(N.lib.imba_memset(self.image, 0, 12 * self.req.width * self.req.height, st), "image")
        N.check(N.lib.imba_vcm_light(C.byref(self.frame), (self.req.seed ^ 0x5A5A5A5A) & 0xFFFFFFFF, 0, self.P,
                                     self.rays[0], self.states[0], st), "vcm_light")

    def tracing(self) -> bool:
        return self.phase in ("light", "camera")

    def trace_item(self, srv):
        cnt = self._light(srv, self.bounce) if self.phase == "light" else self._camera(srv, self.bounce)
        cap = self.P if self.phase == "light" else self.n
        return (self.sc.gpu, self.rays[self.cur], self.hits, cap, cnt, False)

    def before_trace(self, srv, st):
        if self.phase != "grid":
            return
        self.grid = srv.take_grid()
        if self.vertices > 0:
            N.check(N.lib.imba_hashgrid_build(self.grid.h, self.vtx_a, 32, 0, self.vertices, self.sc.radius, st),
                    "grid build")
            N.check(N.lib.imba_vcm_merge_records(self.grid.h, self.vtx_a, self.vtx_b, self.records, st),
                    "vcm_merge_records")
        N.check(N.lib.imba_vcm_camera(C.byref(self.frame), self.req.seed, 0, self.n, self.rays[0], self.states[0], st),
                "vcm_camera")
        self.phase, self.bounce, self.cur = "camera", 0, 0

    def bounce_step(self, srv, st):
        """The bounce kernel of the current phase (after the closest-hit traversal): light
        vertices and camera connections, or next event, connections and merge queries."""
        b, cur, fa = self.bounce, self.cur, C.byref(self.frame)
        if self.phase == "light":
            N.check(N.lib.imba_vcm_light_bounce(fa, self.rays[cur], self.hits, self.states[cur], self.P,
                                                self._light(srv, b), C.c_void_p(self.slot_a.value + 32 * b * self.P),
                                                C.c_void_p(self.slot_b.value + 32 * b * self.P), self.rays[cur],
                                                self.shadow_rays, self.shadow_states, st), "vcm_light_bounce")
        else:
            N.check(N.lib.imba_vcm_camera_bounce(fa, self.rays[cur], self.hits, self.states[cur], self.n,
                                                 self._camera(srv, b), self.vtx_a, self.vtx_b, self._cache(srv),
                                                 self.rays[cur], self.shadow_rays, self.shadow_states,
                                                 self._shadow(srv), self.queries, self.image, st), "vcm_camera_bounce")
        cap, cnt = (self.P, self._light(srv, b)) if self.phase == "light" else \
            ((1 + VCM_C) * self.n, self._shadow(srv))
        N.check(N.lib.imba_compact_rays_dev(self.shadow_rays, self.shadow_states, 16, cap, cnt, self.shadow_rays_c,
                                            self.shadow_states_c, self._live_shadow(srv), self.workspace, st),
                "compact shadows")

    def shadow_item(self, srv):
        """The bounce's shadow rays as a traversal item (any hit), with the estimate of their
        number the server uses to choose an own launch."""
        if self.phase == "light":
            return (self.sc.gpu, self.shadow_rays_c, self.shadow_hits, self.P, self._live_shadow(srv), True,
                    self.queued)
        ns = (1 + VCM_C) * self.n
        return (self.sc.gpu, self.shadow_rays_c, self.shadow_hits, ns, self._live_shadow(srv), True,
                (1 + VCM_C) * self.queued)

    def finish_step(self, srv, st):
        """After the shadow traversal: the unoccluded contributions, the merge (camera bounces)
        and the compaction; the last light bounce also compacts the vertex slots."""
        b, cur, fa = self.bounce, self.cur, C.byref(self.frame)
        sb = VC.VCM_STATE_DTYPE.itemsize
        if self.phase == "light":
            cnt = self._light(srv, b)
            N.check(N.lib.imba_pt_shadow(self.shadow_hits, self.shadow_states_c, self.P, self._live_shadow(srv),
                                         self.image, st), "pt_shadow")
            N.check(N.lib.imba_compact_rays_dev(self.rays[cur], self.states[cur], sb, self.P, cnt, self.rays[1 - cur],
                                                self.states[1 - cur], self._light(srv, b + 1), self.workspace, st),
                    "compact")
            self.bounce, self.cur = b + 1, 1 - cur
            if self.bounce == self.L:
                N.check(N.lib.imba_compact_rays_dev(self.slot_a, self.slot_b, 32, self.L * self.P,
                                                    srv.counts_ptr(self.slot, 2 * srv.L1), self.vtx_a, self.vtx_b,
                                                    self._cache(srv), self.workspace, st), "compact vertices")
                self.phase = "cache"
        elif self.phase == "camera":
            cnt = self._camera(srv, b)
            ns = (1 + VCM_C) * self.n
            N.check(N.lib.imba_pt_shadow(self.shadow_hits, self.shadow_states_c, ns, self._live_shadow(srv), self.image,
                                         st), "pt_shadow")
            if self.vertices > 0:
                N.check(N.lib.imba_vcm_merge(fa, self.grid.h, self.records, self.queries, self.n, cnt, VCM_K,
                                             self.image, st), "vcm_merge")
            N.check(N.lib.imba_compact_rays_dev(self.rays[cur], self.states[cur], sb, self.n, cnt, self.rays[1 - cur],
                                                self.states[1 - cur], self._camera(srv, b + 1), self.workspace, st),
                    "compact")
            self.bounce, self.cur = b + 1, 1 - cur

    def update(self, srv, row) -> bool:
        """Advances the phase from the read-back counts; True when the request is complete."""
        if self.phase == "light":
            self.queued = int(row[self.bounce])
        elif self.phase == "cache":
            self.vertices = int(row[2 * srv.L1 + 1])
            self.phase = "grid"
        elif self.phase == "camera":
            self.queued = int(row[srv.L1 + self.bounce])
            if self.queued == 0 or self.bounce >= self.L:
                light = [int(x) for x in row[:self.L]]
                cam = [int(x) for x in row[srv.L1:srv.L1 + self.L]]
                self.req.counts = light + cam
                self.req.rays = int(sum(light) + sum(cam))
                return True
        return False


class Server:
    # A queue at least this long fills the GPU by itself: it gets its own k_trace launch in a
    # batched step, and the shorter queues share one batched launch.
    OWN_LAUNCH_RAYS = 131072
    # The phased requests' own kernels (grid builds, VCM bounce, merge and compaction, photon
    # mapping bounces) run on this many side streams in a batched step, so that the small
    # launches of different requests overlap on the device.
    SIDE_STREAMS = 4

    def __init__(self, scenes: list, arena_bytes: int, max_active: int = 32, batched: bool = True,
                 max_path_len: int = 8, rr_start: int = 3, cfg: N.TraceConfig | None = None,
                 max_rays_in_flight: int | None = None):
        """max_rays_in_flight: admission also waits (first come, first served) while the
        requests in flight already hold this many first-traversal rays (first_rays), so that
        large requests are not spread over too many at once (each one's latency grows with the
        work it shares the GPU with) while small ones still batch up to max_active."""
        self.scenes = scenes
        self.arena = DeviceArena(arena_bytes)
        self.max_active = max_active if batched else 1
        self.max_rays = max_rays_in_flight if batched else None
        self.rays_in_flight = 0
        self.batched = batched
        self.max_path_len, self.rr_start = max_path_len, rr_start
        self.cfg = cfg
        self.stream = N.Stream()
        # per request slot: PT bounce counts [0, L]; PPM / VCM light counts [0, L], camera
        # counts [L1, L1 + L], photon (vertex) slots at 2 L1, photons (cache size) at 2 L1 + 1,
        # VCM camera shadow slots at 2 L1 + 2, VCM occupied shadow slots at 2 L1 + 3
        self.L1 = max(max_path_len, PPM_DEPTH, VCM_DEPTH) + 1
        self.cs = 2 * self.L1 + 4
        self.counts = N.DeviceBuffer(4 * self.cs * self.max_active)
        self.free_slots = list(range(self.max_active))
        self.pending: list[Request] = []
        self.active: list[_Job] = []
        self.done: list[Request] = []
        self.batch = S.TraceBatch([])
        self.shadow_batch = S.TraceBatch([])
        self.shade = S.ShadeBatch([])
        self.grids: list = []  # hash grids of finished photon-mapping requests, for reuse
        self.steps = 0
        self.t0 = None
        # device time of every step, from its first launch to the read-back of the queue sizes
        # (event pair recorded each step): gpu_ms / wall time is the share of the run the GPU had
        # work queued; the rest it waited for the host
        self._ev = (N.Event(), N.Event())
        self.gpu_ms = 0.0
        self.side = [N.Stream() for _ in range(self.SIDE_STREAMS)] if batched else []
        self._fork_ev = N.Event()
        self._join_ev = [N.Event() for _ in self.side]
        self._forked = 0

    def _fork(self, jobs) -> dict:
        """Streams for the per-request work of `jobs` (job id -> stream handle): on the side
        streams, round robin, behind the main stream's work so far (plain path: the main
        stream).  _join() makes the main stream wait for that work."""
        if not self.side or not jobs:
            return {id(j): self.stream.handle for j in jobs}
        self._fork_ev.record(self.stream)
        self._forked = min(len(jobs), len(self.side))
        for s in self.side[:self._forked]:
            s.wait(self._fork_ev)
        return {id(j): self.side[k % len(self.side)].handle for k, j in enumerate(jobs)}

    def _join(self):
        for k in range(self._forked):
            self._join_ev[k].record(self.side[k])
            self.stream.wait(self._join_ev[k])
        self._forked = 0

    def _shade_item(self, j) -> N.ShadeItem:
        return N.ShadeItem(j.rays[j.cur].value, j.hits.value, j.states[j.cur].value, j.n,
                           self.counts_ptr(j.slot, j.bounce).value, j.rays[1 - j.cur].value,
                           j.states[1 - j.cur].value, self.counts_ptr(j.slot, j.bounce + 1).value, j.shadow_rays.value,
                           j.shadow_states.value, j.shadow_hits.value, j.image.value, self._shading(j.sc))

    def _shading(self, sc) -> N.PtShading:
        return sc.shading.struct(self.max_path_len, self.rr_start)

    def take_grid(self):
        """A hash grid with no table-size history (reused: no allocation once warm)."""
        g = self.grids.pop() if self.grids else S.HashGrid()
        g.reset()
        return g

    def counts_ptr(self, slot: int, index: int):
        return self.counts.at(4 * (self.cs * slot + index))

    def now(self) -> float:
        return time.perf_counter() - self.t0

    def submit(self, req: Request):
        if footprint(req) > self.arena.size:
            raise ValueError("request %d needs %d bytes, more than the arena" % (req.rid, footprint(req)))
        self.pending.append(req)

    def _admit(self):
        st = self.stream.handle
        while self.pending and self.free_slots:
            req = self.pending[0]
            if self.max_rays and self.active and self.rays_in_flight + first_rays(req) > self.max_rays:
                break  # first come first served: wait for work in flight to finish
            block = self.arena.alloc(footprint(req))
            if block is None:
                break  # first come first served: wait for memory
            self.pending.pop(0)
            req.admitted = self.now()
            self.rays_in_flight += first_rays(req)
            if req.integrator in ("ppm", "vcm"):
                cls = _PPMJob if req.integrator == "ppm" else _VCMJob
                job = cls(req, self.free_slots.pop(0), block, self.scenes[req.scene], self.arena)
                job.start(self, st)
                self.active.append(job)
                continue
            job = _Job(req, self.free_slots.pop(0), block, self.scenes[req.scene], self.arena)
            n = job.n
            c0 = np.zeros(self.cs, np.int32)
            c0[0] = n
            job.c0 = c0  # kept until the copy has run (the step ends with a stream synchronisation)
            N.check(N.lib.imba_memcpy_h2d(self.counts_ptr(job.slot, 0), N.ptr(c0), c0.nbytes, st), "counts")
            N.check(N.lib.imba_memset(job.image, 0, 12 * req.width * req.height, st), "image")
            N.check(N.lib.imba_pt_camera(N.ptr(job.r2w), N.ptr(job.pos), req.width, req.height, req.spp, req.seed, 0,
                                         n, job.rays[0], job.states[0], st), "camera")
            self.active.append(job)

    def step(self):
        """One step: admit; the phased requests' non-traversal phases (grid build and camera
        rays, or queries and merging); one batched traversal over every active queue; the PT
        bounces and compactions in one batched launch and the VCM bounce kernels; one batched
        any-hit traversal over every shadow queue (PT next event, VCM connections and next
        event); the shadow contributions, VCM merges and compactions and the photon-mapping
        bounces; one read-back of all queue sizes."""
        self._ev[0].record(self.stream)
        self._admit()
        if not self.active:
            return
        st = self.stream.handle
        pt_jobs = [j for j in self.active if isinstance(j, _Job)]
        ppm_jobs = [j for j in self.active if isinstance(j, (_PPMJob, _VCMJob))]  # phased requests
        grid_jobs = [j for j in ppm_jobs if j.phase in ("grid", "merge")]
        js = self._fork(grid_jobs)
        for j in ppm_jobs:
            j.before_trace(self, js.get(id(j), st))
        self._join()
        items = [(j.sc.gpu, j.rays[j.cur], j.hits, j.n, self.counts_ptr(j.slot, j.bounce), False, j.queued)
                 for j in pt_jobs] + [j.trace_item(self) + (j.queued,) for j in ppm_jobs if j.tracing()]
        own = [it for it in items if not self.batched or it[6] >= self.OWN_LAUNCH_RAYS]
        shared = [it[:6] for it in items if self.batched and it[6] < self.OWN_LAUNCH_RAYS]
        for sc, rays, hits, cap, cnt, _, _ in own:
            N.check(N.lib.imba_intersect_dev(sc.h, rays, hits, cap, cnt, C.byref(self.cfg) if self.cfg else None, st),
                    "intersect_dev")
        if shared:
            self.batch.update(shared, self.stream)
            self.batch.launch(self.cfg, self.stream)
        vcm_steps = [j for j in ppm_jobs if isinstance(j, _VCMJob) and j.tracing()]
        sitems = []
        if self.batched and pt_jobs:
            # every PT request's shading step and compaction in one batched launch; their shadow
            # rays join the step's any-hit traversal below
            self.shade.update([self._shade_item(j) for j in pt_jobs], self.stream)
            self.shade.launch(self.stream)
            sitems += [(j.sc.gpu, j.shadow_rays, j.shadow_hits, j.n, self.counts_ptr(j.slot, j.bounce), True, j.queued)
                       for j in pt_jobs]
        js = self._fork(vcm_steps)
        for j in vcm_steps:
            j.bounce_step(self, js[id(j)])
            sitems.append(j.shadow_item(self))
        self._join()
        # every shadow queue of the step (PT next event, VCM camera connections / next event /
        # vertex connections): long ones in their own launch, the rest in one batched launch
        for sc, rays, hits, cap, cnt, _, q in sitems:
            if not self.batched or q >= self.OWN_LAUNCH_RAYS:
                N.check(N.lib.imba_occluded_dev(sc.h, rays, hits, cap, cnt, C.byref(self.cfg) if self.cfg else None,
                                                st), "occluded_dev")
        shared_s = [it[:6] for it in sitems if self.batched and it[6] < self.OWN_LAUNCH_RAYS]
        if shared_s:
            self.shadow_batch.update(shared_s, self.stream)
            self.shadow_batch.launch(self.cfg, self.stream)
        if self.batched and pt_jobs:
            self.shade.shadow(self.stream)
        elif pt_jobs:
            for j in pt_jobs:
                sc = j.sc
                cnt = self.counts_ptr(j.slot, j.bounce)
                sh = self._shading(sc)
                N.check(N.lib.imba_pt_shade(C.byref(sh), j.rays[j.cur], j.hits, j.states[j.cur], j.n, cnt, j.image,
                                            j.rays[j.cur], j.shadow_rays, j.shadow_states, st), "pt_shade")
                N.check(N.lib.imba_occluded_dev(sc.gpu.h, j.shadow_rays, j.shadow_hits, j.n, cnt,
                                                C.byref(self.cfg) if self.cfg else None, st), "occluded_dev")
                N.check(N.lib.imba_pt_shadow(j.shadow_hits, j.shadow_states, j.n, cnt, j.image, st), "pt_shadow")
                N.check(N.lib.imba_compact_rays_dev(j.rays[j.cur], j.states[j.cur], PT_STATE_BYTES, j.n, cnt,
                                                    j.rays[1 - j.cur], j.states[1 - j.cur],
                                                    self.counts_ptr(j.slot, j.bounce + 1), j.workspace, st),
                        "compact_dev")
        tail = vcm_steps + [j for j in ppm_jobs if isinstance(j, _PPMJob) and j.tracing()]
        js = self._fork(tail)
        for j in vcm_steps:
            j.finish_step(self, js[id(j)])
        for j in tail[len(vcm_steps):]:
            j.after_trace(self, js[id(j)])
        self._join()
        for j in pt_jobs:
            j.bounce += 1
            j.cur = 1 - j.cur
        allc = self.counts.download(np.int32, self.cs * self.max_active, self.stream)
        self._ev[1].record(self.stream)
        self.stream.sync()
        self.gpu_ms += self._ev[0].elapsed_ms(self._ev[1])
        allc = allc.reshape(self.max_active, self.cs)
        self.steps += 1
        still = []
        for j in self.active:
            c = allc[j.slot]
            if isinstance(j, (_PPMJob, _VCMJob)):
                finished = j.update(self, c)
            else:
                j.queued = int(c[j.bounce])
                finished = c[j.bounce] == 0 or j.bounce >= self.max_path_len
                if finished:
                    j.req.counts = [int(x) for x in c[:j.bounce]]
                    j.req.rays = int(sum(j.req.counts))
            if finished:
                j.req.done = self.now()
                self.rays_in_flight -= first_rays(j.req)
                self.done.append(j.req)
                self.arena.release(*j.block)
                self.free_slots.append(j.slot)
                if isinstance(j, (_PPMJob, _VCMJob)) and j.grid is not None:
                    self.grids.append(j.grid)
                    j.grid = None
            else:
                still.append(j)
        self.active = still

    def run(self, requests: list, keep_images: bool = False, progress=None):
        """Runs the requests with their arrival times (seconds after start); returns the
        completed requests (and {rid: image} with keep_images)."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        images = {}
        self.t0 = time.perf_counter()
        k = 0
        last_report = 0.0
        while k < len(reqs) or self.pending or self.active:
            now = self.now()
            while k < len(reqs) and reqs[k].arrival <= now:
                self.submit(reqs[k])
                k += 1
            if not self.pending and not self.active:
                if k < len(reqs):
                    time.sleep(max(0.0, min(0.002, reqs[k].arrival - now)))
                continue
            before = {j.req.rid: j for j in self.active}
            self.step()
            if keep_images:
                for rid, j in before.items():
                    if j.req.done >= 0 and rid not in images:
                        images[rid] = self._image_now(j)
            if progress and self.now() - last_report > 5.0:
                last_report = self.now()
                progress("%d/%d requests done, %d active, %d pending" % (len(self.done), len(reqs), len(self.active),
                                                                        len(self.pending)))
        return (self.done, images) if keep_images else self.done

    def _image_now(self, job) -> np.ndarray:
        out = np.empty(job.req.width * job.req.height * 3, np.float32)
        N.check(N.lib.imba_memcpy_d2h(N.ptr(out), job.image, out.nbytes, self.stream.handle), "image")
        self.stream.sync()
        return out


def poisson_requests(n: int, rate: float, n_sessions: int, n_scenes: int, scene_centres: list, seed: int = 7,
                     resolutions=((256, 256), (640, 480), (1280, 720), (1920, 1080)), spp=(1, 4),
                     integrators=("pt",), ceiling: float = float("inf")) -> list:
    """n requests with Poisson arrivals at `rate` per second from n_sessions sessions, each for a
    uniformly drawn scene, resolution, spp and integrator and a random camera looking at the
    scene centre (cx, cy, cz) from `radius` away, no higher than `ceiling`."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for rid in range(n):
        t += rng.exponential(1.0 / rate)
        sc = int(rng.integers(0, n_scenes))
        w, h = resolutions[int(rng.integers(0, len(resolutions)))]
        cx, cy, cz, radius = scene_centres[sc]
        ang = rng.uniform(0, 2 * np.pi)
        pos = (cx + radius * np.sin(ang), min(cy + rng.uniform(0.3, 0.9) * radius, ceiling), cz - radius * np.cos(ang))
        d = (cx - pos[0], cy * 0.3 - pos[1], cz - pos[2])
        req = Request(rid, int(rng.integers(0, n_sessions)), sc, w, h, int(rng.integers(spp[0], spp[1] + 1)),
                      pos, d, int(rng.integers(1, 2 ** 31 - 1)), t)
        if len(integrators) > 1:
            req.integrator = integrators[int(rng.integers(0, len(integrators)))]
        else:
            req.integrator = integrators[0]
        out.append(req)
    return out


def latency_stats(done: list) -> dict:
    lat = np.array([(r.done - r.arrival) * 1e3 for r in done])
    rays = sum(r.rays for r in done)
    span = max(r.done for r in done) - min(r.arrival for r in done)
    return {"requests": len(done), "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "rays": int(rays), "mrays_s": rays / span / 1e6,
            "requests_per_s": len(done) / span, "span_s": span}
