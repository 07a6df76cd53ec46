"""Seeded synthetic workloads of the measured configurations.

random_triangle_workload: 65,536 random triangles in the unit cube, binned-SAH BVH, 1024x1024
    pinhole primary rays from (0.5, 0.5, -1.5) looking +z (60 deg), and one occlusion ray per
    pixel towards the point light (0.5, 2, 0.5).
sphere_grid_scene: an open box of 12 triangles, a 32 x 32 grid of level-5 icospheres
    (20,480 triangles each, radius 0.4, spacing 1.0; 20,971,520 triangles in all), one
    triangular area light; every sphere is its own mesh with its own Lambertian albedo drawn
    uniformly from [0.2, 0.8]^3; SBVH per mesh; 1,026 identity instances.
PathTracer: a device-side frame loop over that scene: camera rays, then per bounce closest-hit
    traversal, the shading step of PathTracer::process_primary_rays (emitter hits with their MIS
    weight, one next-event shadow ray per non-emitting hit, Lambertian bounce), any-hit
    traversal of the shadow rays and their contributions, queue compaction; up to
    max_path_len - 1 bounces, Russian roulette from bounce `rr_start` on.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import native as N
from . import scene as S

# the 40-byte path-tracer state (PTState: pixel, sample, rng, throughput, bounces, last pdf,
# last specular) and the 16-byte shadow-ray state (pixel, contribution; pixel -1 = empty slot)
PT_STATE_BYTES = 40
PT_STATE_DTYPE = np.dtype([("pixel", "<i4"), ("sample", "<i4"), ("rng", "<u8"), ("tp", "<f4", 3), ("bounces", "<i4"),
                           ("last_pdf", "<f4"), ("last_specular", "<i4")])
SHADOW_STATE_BYTES = 16
SHADOW_STATE_DTYPE = np.dtype([("pixel", "<i4"), ("tp", "<f4", 3)])
# the 32-byte path state of the photon-mapping producers (path / pixel, sample, rng, throughput,
# bounces)
PPM_STATE_BYTES = 32
PPM_STATE_DTYPE = np.dtype([("pixel", "<i4"), ("sample", "<i4"), ("rng", "<u8"), ("tp", "<f4", 3), ("bounces", "<i4")])

C1_CAMERA = ((0.5, 0.5, -1.5), (0.0, 0.0, 1.0), (0.0, 1.0, 0.0), 60.0)
C1_LIGHT = np.array([0.5, 2.0, 0.5], np.float32)


def random_tri_scene(n: int, builder: int = N.BUILDER_BINNED_SAH, seed: int = 0x1234, threads: int = 1):
    """n seeded random triangles (one mesh, identity instance), built into a SceneBuild."""
    v, i = S.random_triangles(n, seed)
    b = S.SceneBuild()
    b.add_mesh(v, i)
    b.run(builder, threads)
    return b, v, i


def primary_rays(w: int, h: int) -> np.ndarray:
    """Pinhole rays of the random-triangle workload's camera, one per pixel."""
    pos, d, up, fov = C1_CAMERA
    return S.camera_rays(pos, d, up, fov, w, h)


def shadow_rays(primary: np.ndarray, hits: np.ndarray, light=C1_LIGHT, seed: int = 99) -> np.ndarray:
    """One occlusion ray per primary ray towards the point light (pt.cpp:33-41 conventions):
    origin at the primary hit (or a seeded point in the unit cube on a miss),
    tmin = 1e-3 * dist, tmax = dist - tmin."""
    rng = np.random.default_rng(seed)
    o = primary["org"][:, :3].astype(np.float32)
    d = primary["dir"][:, :3].astype(np.float32)
    t = hits["tmax"].astype(np.float32)[:, None]
    pts = (o + t * d).astype(np.float32)
    miss = hits["tri_id"] < 0
    pts[miss] = rng.random((int(miss.sum()), 3), dtype=np.float32)
    to = (light[None, :] - pts).astype(np.float32)
    dist = np.sqrt((to * to).sum(1)).astype(np.float32)
    dirs = (to / dist[:, None]).astype(np.float32)
    off = (np.float32(1e-3) * dist).astype(np.float32)
    return S.make_rays(pts, dirs, off, (dist - off).astype(np.float32))


C4_SCENE_TRIS = [10_000, 30_000, 100_000, 250_000, 500_000, 1_000_000, 2_000_000, 4_000_000]


def c4_queue_sizes(n_queues: int = 256, lo: int = 1_000, hi: int = 1_000_000, seed: int = 404) -> np.ndarray:
    """Queue sizes log-uniform in [lo, hi] (about 37M rays for 256 queues)."""
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(lo), np.log(hi), n_queues)).astype(np.int64)


def c4_rays(vertices: np.ndarray, indices: np.ndarray, n: int, coherent: bool, seed: int) -> np.ndarray:
    """One queue of the multi-queue workload over a random-triangle scene in the unit cube:
    coherent = pinhole rays from a seeded camera orbiting the cube; otherwise cosine-weighted
    hemisphere rays leaving random points of random triangles (tmin 1e-4)."""
    rng = np.random.default_rng(seed)
    if coherent:
        ang = rng.uniform(0.0, 2.0 * np.pi)
        pos = np.array([0.5 + 1.8 * np.sin(ang), rng.uniform(0.2, 0.8), 0.5 - 1.8 * np.cos(ang)], np.float32)
        look = (np.array([0.5, 0.5, 0.5], np.float32) - pos).astype(np.float32)
        w = max(1, int(np.ceil(np.sqrt(n))))
        h = int(np.ceil(n / w))
        return S.camera_rays(pos, look, (0.0, 1.0, 0.0), 50.0, w, h)[:n]
    tri = rng.integers(0, len(indices), n)
    p = vertices[:, :3][indices[tri, :3].astype(np.int64)].astype(np.float64)  # (n, 3, 3)
    r1, r2 = rng.random(n), rng.random(n)
    s = np.sqrt(r1)
    bary = np.stack([1.0 - s, s * (1.0 - r2), s * r2], 1)
    org = (bary[:, :, None] * p).sum(1)
    nrm = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    nrm /= np.maximum(np.linalg.norm(nrm, axis=1, keepdims=True), 1e-30)
    nrm *= np.where(rng.random(n) < 0.5, -1.0, 1.0)[:, None]
    # cosine-weighted direction around the normal
    u1, u2 = rng.random(n), rng.random(n)
    rr, phi = np.sqrt(u1), 2.0 * np.pi * u2
    local = np.stack([rr * np.cos(phi), rr * np.sin(phi), np.sqrt(np.maximum(0.0, 1.0 - u1))], 1)
    a = np.where(np.abs(nrm[:, :1]) > 0.9, np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
    t = np.cross(nrm, a)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    b = np.cross(nrm, t)
    d = local[:, :1] * t + local[:, 1:2] * b + local[:, 2:] * nrm
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return S.make_rays(org.astype(np.float32), d.astype(np.float32), np.float32(1e-4), np.float32(3.4028235e38))


def mwc_uniform(seed: int, n: int) -> np.ndarray:
    """n draws of the MWC64X generator's random_float() (render/random.h:19-46)."""
    out = np.empty(n, np.float32)
    s = seed & 0xFFFFFFFFFFFFFFFF
    for k in range(n):
        c = s >> 32
        x = s & 0xFFFFFFFF
        s = (x * 4294883355 + c) & 0xFFFFFFFFFFFFFFFF
        out[k] = np.float32(np.float32(x ^ c) / np.float32(4294967296.0))
    return out


def open_box(lo, hi):
    """12 triangles: floor (as 2 quads), back, left, right walls and ceiling; the front
    (camera) side is open."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    xm = 0.5 * (x0 + x1)
    quads = [
        [(x0, y0, z0), (xm, y0, z0), (xm, y0, z1), (x0, y0, z1)],  # floor, left half
        [(xm, y0, z0), (x1, y0, z0), (x1, y0, z1), (xm, y0, z1)],  # floor, right half
        [(x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1)],  # back
        [(x0, y0, z0), (x0, y0, z1), (x0, y1, z1), (x0, y1, z0)],  # left
        [(x1, y0, z0), (x1, y1, z0), (x1, y1, z1), (x1, y0, z1)],  # right
        [(x0, y1, z0), (x0, y1, z1), (x1, y1, z1), (x1, y1, z0)],  # ceiling
    ]
    v, idx = [], []
    for q in quads:
        b = len(v)
        v.extend(q)
        idx.extend([(b, b + 1, b + 2), (b, b + 2, b + 3)])
    v = np.array([(*p, 1.0) for p in v], np.float32)
    i = np.array([(*t, 0) for t in idx], np.uint32)
    return v, i


def triangle_area(p0, p1, p2) -> np.float32:
    """Tri::area (core/tri.h:20) in float32: length(cross(v1 - v0, v2 - v0)) / 2."""
    F = np.float32
    a = np.asarray(p1, F) - np.asarray(p0, F)
    b = np.asarray(p2, F) - np.asarray(p0, F)
    c = np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]], F)
    return F(np.sqrt((c[0] * c[0] + c[1] * c[1]) + c[2] * c[2]) / F(2))


class PtShadingData:
    """A scene's path-tracing shading data in device memory (N.PtShading): per-triangle
    geometric normals and materials, per-material albedo, emission and emitter area, and the
    triangle lights (12 floats each: p0, p1, p2, intensity)."""

    def __init__(self, tri_normals, tri_materials, albedo, emission, emitter_area, lights):
        F = np.float32
        self.tri_n = N.DeviceBuffer.from_array(np.ascontiguousarray(tri_normals, F))
        self.tri_m = N.DeviceBuffer.from_array(np.ascontiguousarray(tri_materials, np.int32))
        self.albedo = N.DeviceBuffer.from_array(np.ascontiguousarray(albedo, F))
        self.emission = N.DeviceBuffer.from_array(np.ascontiguousarray(emission, F))
        self.emitter_area = N.DeviceBuffer.from_array(np.ascontiguousarray(emitter_area, F))
        lights = np.ascontiguousarray(lights, F).reshape(-1, 12)
        self.lights = N.DeviceBuffer.from_array(lights if len(lights) else np.zeros((1, 12), F))
        self.tri_count, self.material_count, self.light_count = len(tri_materials), len(albedo), len(lights)

    def struct(self, max_path_len: int, rr_start: int) -> N.PtShading:
        return N.PtShading(self.tri_n.ptr.value, self.tri_m.ptr.value, self.tri_count, self.albedo.ptr.value,
                           self.emission.ptr.value, self.emitter_area.ptr.value, self.material_count,
                           self.lights.ptr.value, self.light_count, max_path_len, rr_start)


# Served requests look at a sphere grid from inside its box (walls at x = -1 and side + 1, the
# far wall at z = side + 1, the ceiling at y = 8 with the light just below it): from outside,
# a camera would see only the unlit exterior of the box.
SERVED_CEILING = 7.5


def served_view(side: int, spacing: float = 1.0):
    """(cx, cy, cz, radius) of the ring the served cameras of a side x side grid sit on."""
    ext = side * spacing
    return (ext * 0.5, 0.4, ext * 0.5, 0.5 * ext + 0.5)


class SphereGridScene:
    def __init__(self, side: int = 32, level: int = 5, radius: float = 0.4, spacing: float = 1.0,
                 builder: int = N.BUILDER_SBVH, threads: int = 8, seed: int = 0x5EED, specular_fraction: float = 0.0):
        self.side = side
        self.seed = seed
        ext = side * spacing
        self.bounds = ((-1.0, 0.0, -12.0), (ext + 1.0, 8.0, ext + 1.0))
        self.build = S.SceneBuild()
        n_mat = side * side + 2
        self.albedo = np.zeros((n_mat, 3), np.float32)
        self.emission = np.zeros((n_mat, 3), np.float32)
        # material 0: box, 1..side^2: spheres, last: light
        bv, bi = open_box(*self.bounds)
        self.build.add_mesh(bv, bi)
        self.albedo[0] = 0.6
        draws = mwc_uniform(seed, 3 * side * side)
        self.albedo[1:1 + side * side] = (0.2 + 0.6 * draws).reshape(-1, 3)
        # material kinds (photon mapping): 0 Lambertian, 1 mirror, 2 glass; a fraction of the
        # spheres, half mirrors and half glass, chosen by a second draw per sphere
        self.kinds = self._kinds(specular_fraction)
        for k in range(side * side):
            gx, gz = k % side, k // side
            centre = ((gx + 0.5) * spacing, radius, (gz + 0.5) * spacing)
            v, i = S.icosphere(level, radius, centre, 1 + k)
            self.build.add_mesh(v, i)
        light = n_mat - 1
        lv = np.array([[ext * 0.35, 7.9, ext * 0.35, 1], [ext * 0.65, 7.9, ext * 0.35, 1],
                       [ext * 0.5, 7.9, ext * 0.65, 1]], np.float32)
        # the light triangle is wound so that its geometric normal, cross(v1 - v0, v2 - v0),
        # faces the scene (down): emitters radiate on that side only (render/light.h:19-39)
        self.build.add_mesh(lv, np.array([[0, 1, 2, light]], np.uint32))
        self.emission[light] = 40.0
        self.light_tri = np.ascontiguousarray(lv[:, :3])
        self.emitter_area = np.zeros(n_mat, np.float32)
        self.emitter_area[light] = triangle_area(*self.light_tri)
        # TriangleLight(intensity, p0, p1, p2), as the reference creates one per emissive face
        self.lights = np.concatenate([self.light_tri.reshape(-1), self.emission[light]]).astype(np.float32)[None, :]
        self.build.run(builder, threads)
        self.desc = self.build.desc()
        self.tri_normals, self.tri_materials = self.build.tri_attrs()
        self.tri_count = len(self.tri_materials)
        self.camera = ((ext * 0.5, 6.0, -10.0), (0.0, -6.0, 26.0), (0.0, 1.0, 0.0), 60.0)
        self.offset = 1e-4 * float(np.linalg.norm(np.subtract(self.bounds[1], self.bounds[0])))


    def _kinds(self, specular_fraction: float) -> np.ndarray:
        side = self.side
        kinds = np.zeros(side * side + 2, np.int32)
        if specular_fraction > 0.0:
            u = mwc_uniform(self.seed ^ 0x5BEC, side * side)
            sph = kinds[1:1 + side * side]
            sph[u < 0.5 * specular_fraction] = 1
            sph[(u >= 0.5 * specular_fraction) & (u < specular_fraction)] = 2
        return kinds

    def with_specular(self, specular_fraction: float) -> "SphereGridScene":
        """The same geometry (same BVHs, not rebuilt) with the material kinds a scene built with
        this specular_fraction has."""
        import copy

        other = copy.copy(self)
        other.kinds = self._kinds(specular_fraction)
        return other


class PathTracer:
    """Device-resident PT frame loop over a SphereGridScene (or any scene with per-triangle
    normals and materials, per-material albedo / emission / emitter area and triangle lights).

    Per bounce: closest-hit traversal of the queue, the shading step (imba_pt_shade: emitter
    contributions, one next-event shadow ray per non-emitting hit into the shadow slots,
    Lambertian bounce in place), any-hit traversal of the shadow slots, their contributions
    (imba_pt_shadow), stable compaction of the continuing paths."""

    def __init__(self, scene: SphereGridScene, width: int, height: int, spp: int = 1, max_path_len: int = 8,
                 rr_start: int = 3, seed_base: int = 0x1234567, trace_cfg: N.TraceConfig | None = None,
                 shadow_cfg: N.TraceConfig | None = None, gpu_scene: S.Scene | None = None, camera=None):
        """gpu_scene: an uploaded scene to share (another frame loop over the same scene);
        camera: (position, direction, up, fov in degrees), else the scene's."""
        self.s = scene
        self.gpu_scene = gpu_scene if gpu_scene is not None else S.Scene(scene.desc)
        self.w, self.h, self.spp = width, height, spp
        self.n = width * height * spp
        self.max_path_len, self.rr_start, self.seed_base = max_path_len, rr_start, seed_base
        self.cfg = trace_cfg
        self.shadow_cfg = shadow_cfg
        self.bounce_cfg = None  # optional per-bounce launch configurations (see cfg_of)
        self.shadow_bounce_cfg = None
        pos, d, up, fov = camera if camera is not None else scene.camera
        self.cam_pos = np.asarray(pos, np.float32)
        self.r2w = np.zeros(16, np.float32)
        dd, uu = np.asarray(d, np.float32), np.asarray(up, np.float32)
        N.check(N.lib.imba_gen_camera_matrix(N.ptr(self.cam_pos), N.ptr(dd), N.ptr(uu), fov, width, height,
                                             N.ptr(self.r2w)), "camera matrix")
        self.rays = [N.DeviceBuffer(self.n * 32), N.DeviceBuffer(self.n * 32)]
        self.states = [N.DeviceBuffer(self.n * PT_STATE_BYTES), N.DeviceBuffer(self.n * PT_STATE_BYTES)]
        self.hits = N.DeviceBuffer(self.n * 16)
        self.shadow_rays = N.DeviceBuffer(self.n * 32)
        self.shadow_states = N.DeviceBuffer(self.n * SHADOW_STATE_BYTES)
        self.shadow_hits = N.DeviceBuffer(self.n * 16)
        self.counts = None
        self.shading = PtShadingData(scene.tri_normals, scene.tri_materials, scene.albedo, scene.emission,
                                     scene.emitter_area, scene.lights)
        self.sh = self.shading.struct(max_path_len, rr_start)
        self.image = N.DeviceBuffer(width * height * 12)
        self.stream = N.Stream()
        # frames add into the image: it starts at zero (device allocations are not cleared)
        N.check(N.lib.imba_memset(self.image.ptr, 0, width * height * 12, self.stream.handle), "memset")
        self.events = []
        self.dev_counts = None
        self.graph = None

    def frame(self, stats: bool = False, keep_queues: int = 0, timed: bool = False):
        """Runs one frame. Returns per-bounce ray counts, and with stats=True the per-bounce
        visit-count sums (nodes, triangles, instances) of the closest-hit traversal (those of the
        shadow traversal, over all its slots, go to self.shadow_visit); keep_queues > 0 also
        returns host copies of the first keep_queues ray queues (their states go to
        self.kept_states, their shadow slots and any-hit results to self.kept_shadow).
        self.shadow_counts holds the shadow rays (non-empty slots) of every bounce."""
        st = self.stream.handle
        n = self.n
        N.check(N.lib.imba_pt_camera(N.ptr(self.r2w), N.ptr(self.cam_pos), self.w, self.h, self.spp, self.seed_base, 0,
                                     n, self.rays[0].ptr, self.states[0].ptr, st), "camera")
        cur = 0
        counts, visit, queues, ev = [], [], [], []
        self.kept_states = []  # with keep_queues: the states of the kept queues
        self.kept_shadow = []  # with keep_queues: (shadow rays, shadow states, shadow hits) of the kept bounces
        self.kept_hits = []    # with keep_queues: the closest hits of the kept queues
        self.shadow_counts = []
        self.shadow_visit = []
        self.shadow_events = []
        if stats and self.counts is None:
            self.counts = N.DeviceBuffer(self.n * 12)
        for bounce in range(self.max_path_len):
            if n == 0:
                break
            counts.append(n)
            keep = len(queues) < keep_queues
            if keep:
                queues.append(self.rays[cur].download(N.RAY_DTYPE, n, st))
                self.kept_states.append(self.states[cur].download(PT_STATE_DTYPE, n, st))
            if stats:
                N.check(N.lib.imba_trace_stats(self.gpu_scene.h, self.rays[cur].ptr, self.hits.ptr, self.counts.ptr,
                                               n, 0, self._cfg_ref(bounce), st), "trace_stats")
                c = self.counts.download(np.uint32, 3 * n, st).reshape(-1, 3)
                visit.append(c.astype(np.int64).sum(0))
            else:
                if timed:
                    e0, e1 = N.Event(), N.Event()
                    e0.record(self.stream)
                N.check(N.lib.imba_intersect(self.gpu_scene.h, self.rays[cur].ptr, self.hits.ptr, n,
                                             self._cfg_ref(bounce), st), "intersect")
                if timed:
                    e1.record(self.stream)
                    ev.append((e0, e1, n))
            if keep:
                self.kept_hits.append(self.hits.download(N.HIT_DTYPE, n, st))
            N.check(N.lib.imba_pt_shade(C.byref(self.sh), self.rays[cur].ptr, self.hits.ptr, self.states[cur].ptr, n,
                                        None, self.image.ptr, self.rays[cur].ptr, self.shadow_rays.ptr,
                                        self.shadow_states.ptr, st), "pt_shade")
            sst = self.shadow_states.download(SHADOW_STATE_DTYPE, n, st)
            self.shadow_counts.append(int((sst["pixel"] >= 0).sum()))
            if stats:
                # visit counts of every slot of the launch (an empty slot tests the root once)
                N.check(N.lib.imba_trace_stats(self.gpu_scene.h, self.shadow_rays.ptr, self.shadow_hits.ptr,
                                               self.counts.ptr, n, 1, self._shadow_cfg_ref(bounce), st), "trace_stats")
                c = self.counts.download(np.uint32, 3 * n, st).reshape(-1, 3)
                self.shadow_visit.append(c.astype(np.int64).sum(0))
            else:
                if timed:
                    e0, e1 = N.Event(), N.Event()
                    e0.record(self.stream)
                N.check(N.lib.imba_occluded(self.gpu_scene.h, self.shadow_rays.ptr, self.shadow_hits.ptr, n,
                                            self._shadow_cfg_ref(bounce), st), "occluded")
                if timed:
                    e1.record(self.stream)
                    self.shadow_events.append((e0, e1, n))
            if keep:
                self.kept_shadow.append((self.shadow_rays.download(N.RAY_DTYPE, n, st), sst,
                                         self.shadow_hits.download(N.HIT_DTYPE, n, st)))
            N.check(N.lib.imba_pt_shadow(self.shadow_hits.ptr, self.shadow_states.ptr, n, None, self.image.ptr, st),
                    "pt_shadow")
            nn = C.c_int()
            N.check(N.lib.imba_compact_rays(self.rays[cur].ptr, self.states[cur].ptr, PT_STATE_BYTES, n,
                                            self.rays[1 - cur].ptr, self.states[1 - cur].ptr, C.byref(nn), st),
                    "compact_rays")
            n = nn.value
            cur = 1 - cur
        self.events = ev
        return counts, visit, queues

    def trace_times_ms(self):
        return [(e0.elapsed_ms(e1), n) for e0, e1, n in self.events]

    def shadow_times_ms(self):
        return [(e0.elapsed_ms(e1), n) for e0, e1, n in self.shadow_events]

    def cfg_of(self, bounce: int):
        """Launch configuration of a bounce's traversal: bounce_cfg[bounce] when set (the queues
        of a frame shrink from bounce to bounce), else cfg."""
        if self.bounce_cfg is not None and bounce < len(self.bounce_cfg) and self.bounce_cfg[bounce] is not None:
            return self.bounce_cfg[bounce]
        return self.cfg

    def shadow_cfg_of(self, bounce: int):
        """Launch configuration of a bounce's shadow-ray traversal (shadow_bounce_cfg, else
        shadow_cfg, else the closest-hit configuration)."""
        if (self.shadow_bounce_cfg is not None and bounce < len(self.shadow_bounce_cfg)
                and self.shadow_bounce_cfg[bounce] is not None):
            return self.shadow_bounce_cfg[bounce]
        return self.shadow_cfg if self.shadow_cfg is not None else self.cfg_of(bounce)

    def _cfg_ref(self, bounce: int):
        c = self.cfg_of(bounce)
        return C.byref(c) if c else None

    def _shadow_cfg_ref(self, bounce: int):
        c = self.shadow_cfg_of(bounce)
        return C.byref(c) if c else None

    # -- device-resident frame: counts stay on the GPU, the frame is one graph launch ----------
    def _frame_dev(self, batches):
        """Enqueues one frame whose queue sizes live in self.dev_counts (entry b = rays of bounce
        b): every bounce is launched, a bounce whose queue is empty does no work."""
        st = self.stream.handle
        n = self.n
        N.check(N.lib.imba_pt_camera(N.ptr(self.r2w), N.ptr(self.cam_pos), self.w, self.h, self.spp, self.seed_base, 0,
                                     n, self.rays[0].ptr, self.states[0].ptr, st), "camera")
        cur = 0
        for bounce in range(self.max_path_len):
            cnt = self.dev_counts.at(4 * bounce)
            if batches:
                batches[bounce].launch(self.cfg_of(bounce), self.stream)
            else:
                N.check(N.lib.imba_intersect_dev(self.gpu_scene.h, self.rays[cur].ptr, self.hits.ptr, n, cnt,
                                                 self._cfg_ref(bounce), st), "intersect_dev")
            N.check(N.lib.imba_pt_shade(C.byref(self.sh), self.rays[cur].ptr, self.hits.ptr, self.states[cur].ptr, n,
                                        cnt, self.image.ptr, self.rays[cur].ptr, self.shadow_rays.ptr,
                                        self.shadow_states.ptr, st), "pt_shade")
            N.check(N.lib.imba_occluded_dev(self.gpu_scene.h, self.shadow_rays.ptr, self.shadow_hits.ptr, n, cnt,
                                            self._shadow_cfg_ref(bounce), st), "occluded_dev")
            N.check(N.lib.imba_pt_shadow(self.shadow_hits.ptr, self.shadow_states.ptr, n, cnt, self.image.ptr, st),
                    "pt_shadow")
            N.check(N.lib.imba_compact_rays_dev(self.rays[cur].ptr, self.states[cur].ptr, PT_STATE_BYTES, n, cnt,
                                                self.rays[1 - cur].ptr, self.states[1 - cur].ptr,
                                                self.dev_counts.at(4 * (bounce + 1)), self.workspace.ptr, st),
                    "compact_rays_dev")
            cur = 1 - cur

    def build_graph(self, use_batch: bool = False):
        """Captures one device-resident frame into a graph (after one eager run of the same
        calls, which sizes the spill stacks and counters the captured launches use).
        use_batch traverses with the batched multi-queue kernel (one queue) instead of k_trace."""
        from .scene import Graph, TraceBatch

        n = self.n
        self.dev_counts = N.DeviceBuffer(4 * (self.max_path_len + 1))
        counts0 = np.zeros(self.max_path_len + 1, np.int32)
        counts0[0] = n
        self.dev_counts.upload(counts0)
        self.workspace = N.DeviceBuffer(int(N.lib.imba_compact_workspace_size(n)))
        batches = None
        if use_batch:
            batches = [TraceBatch([(self.gpu_scene, self.rays[b % 2], self.hits, n, self.dev_counts.at(4 * b), False)])
                       for b in range(self.max_path_len)]
        self._batches = batches
        self._frame_dev(batches)
        self.stream.sync()
        Graph.begin(self.stream)
        try:
            self._frame_dev(batches)
        finally:
            g = Graph.end(self.stream)
        self.graph = g
        return g

    def frame_graph(self):
        """One frame by graph launch; returns the per-bounce ray counts (non-empty bounces)."""
        self.launch_graph()
        return self.finish_graph()

    def launch_graph(self):
        """Enqueues one graph frame and the read-back of its queue sizes on this loop's stream
        (frames of several loops on their own streams then overlap on the device)."""
        self.graph.launch(self.stream)
        if getattr(self, "_counts_host", None) is None:
            self._counts_host = N.PinnedArray(self.max_path_len + 1, np.int32)  # the copy does not block
        N.check(N.lib.imba_memcpy_d2h(self._counts_host.ptr, self.dev_counts.ptr, self._counts_host.array.nbytes,
                                      self.stream.handle), "counts")

    def finish_graph(self):
        """Waits for the frame launch_graph enqueued; returns its per-bounce ray counts."""
        self.stream.sync()
        c = self._counts_host.array
        out = []
        for v in c[:self.max_path_len]:
            if v == 0:
                break
            out.append(int(v))
        return out


PHOTON_RAY_DTYPE = np.dtype([("pos", "<f4", 3), ("length", "<i4"), ("out_dir", "<f4", 3), ("pad", "<f4")])
PHOTON_STATE_DTYPE = np.dtype([("path", "<i4"), ("tp", "<f4", 3)])
QUERY_DTYPE = np.dtype([("pos", "<f4", 3), ("pixel", "<i4"), ("normal", "<f4", 3), ("pad0", "<f4"),
                        ("out_dir", "<f4", 3), ("pad1", "<f4"), ("weight", "<f4", 3), ("pad2", "<f4")])


class PhotonMapper:
    """Device-resident photon-mapping frame (the PPM mode of the VCM integrator,
    render/integrators/vcm.cpp with ALGO_PPM) over a SphereGridScene:

      light pass   light_paths paths from the triangle light, up to max_depth hits each; every
                   non-specular hit deposits a photon into slot (bounce, queue entry)
      photons      the occupied slots, stable-compacted (by bounce, then by path)
      grid         HashGrid over the photon positions with radius = radius_factor * the scene
                   diagonal (LightVertices::build, light_vertices.h:70-76)
      camera pass  one path per pixel and sample, continued through mirrors and glass; the first
                   non-specular hit becomes a merge query
      query        the k nearest photons within the radius of each query (HashGrid::query)
      merge        vertex_merging into the image

    merge="fused" replaces the last two steps by one launch that keeps each query's list in its
    wave and reads the photons' directions and states from records gathered into grid slot order
    when the grid is built (imba_ppm_merge_records / imba_ppm_merge_grid); query() still writes
    the index lists on request.  camera: (position, direction, up, fov in degrees), else the
    scene's.

    Every traversal, producer and compaction runs on the device with queue sizes kept on the
    device; the host reads one number per frame (the photon count the grid build needs)."""

    def __init__(self, scene: SphereGridScene, width: int, height: int, light_paths: int, max_depth: int = 6,
                 rr_start: int = 3, radius_factor: float = 0.005, k: int = 10, spp: int = 1,
                 seed_base: int = 0x1234567, light_seed: int = 0x7654321, trace_cfg: N.TraceConfig | None = None,
                 merge: str = "lists", camera=None):
        from .scene import HashGrid

        if merge not in ("lists", "fused"):
            raise ValueError("merge must be 'lists' or 'fused'")
        if merge == "fused" and not 1 <= k <= 64:
            raise ValueError("the fused merge keeps its list in one wave: 1 <= k <= 64")
        self.merge_mode = merge
        self.s = scene
        self.gpu_scene = S.Scene(scene.desc)
        self.w, self.h, self.spp = width, height, spp
        self.n_cam = width * height * spp
        self.P = light_paths
        self.L = max_depth
        self.rr_start, self.k = rr_start, k
        self.seed_base, self.light_seed = seed_base, light_seed
        self.cfg = trace_cfg
        lo, hi = np.asarray(scene.bounds[0], np.float64), np.asarray(scene.bounds[1], np.float64)
        self.radius = float(np.float32(radius_factor * np.linalg.norm(hi - lo)))
        pos, d, up, fov = camera if camera is not None else scene.camera
        self.cam_pos = np.asarray(pos, np.float32)
        self.r2w = np.zeros(16, np.float32)
        dd, uu = np.asarray(d, np.float32), np.asarray(up, np.float32)  # alive across the call
        N.check(N.lib.imba_gen_camera_matrix(N.ptr(self.cam_pos), N.ptr(dd), N.ptr(uu), fov, width, height,
                                             N.ptr(self.r2w)), "camera matrix")
        assert np.all(np.isfinite(self.r2w))
        cap = max(self.P, self.n_cam)
        self.cap = cap
        self.rays = [N.DeviceBuffer(cap * 32), N.DeviceBuffer(cap * 32)]
        self.states = [N.DeviceBuffer(cap * PPM_STATE_BYTES), N.DeviceBuffer(cap * PPM_STATE_BYTES)]
        self.hits = N.DeviceBuffer(cap * 16)
        slots = self.L * self.P
        self.slots = slots
        self.slot_rays = N.DeviceBuffer(slots * 32)
        self.slot_states = N.DeviceBuffer(slots * 16)
        self.photon_rays = N.DeviceBuffer(slots * 32)
        self.photon_states = N.DeviceBuffer(slots * 16)
        self.records = N.DeviceBuffer(slots * 32) if merge == "fused" else None
        self.queries = N.DeviceBuffer(self.n_cam * 64)
        self.out_idx = N.DeviceBuffer(self.n_cam * k * 4)
        self.out_dist = N.DeviceBuffer(self.n_cam * k * 4)
        self.out_count = N.DeviceBuffer(self.n_cam * 4)
        self.image = N.DeviceBuffer(width * height * 12)
        # device counts: light queue sizes [0, L], camera queue sizes [L+1, 2L+1], slot count,
        # photon count
        self.counts = N.DeviceBuffer(4 * (2 * self.L + 4))
        self.workspace = N.DeviceBuffer(int(N.lib.imba_compact_workspace_size(max(slots, cap))))
        self.tri_n = N.DeviceBuffer.from_array(scene.tri_normals)
        self.tri_m = N.DeviceBuffer.from_array(scene.tri_materials)
        self.albedo = N.DeviceBuffer.from_array(scene.albedo)
        self.emission = N.DeviceBuffer.from_array(scene.emission)
        self.kinds = N.DeviceBuffer.from_array(scene.kinds)
        light = len(scene.albedo) - 1
        self.light_tri = np.ascontiguousarray(scene.light_tri, np.float32).reshape(-1)
        self.light_int = np.ascontiguousarray(scene.emission[light], np.float32)
        self.grid = HashGrid()
        self.stream = N.Stream()
        self.clear_image()  # frames add into the image; device allocations are not cleared
        self.photon_count = 0
        self.events = {}
        # with keep_queues, every bounce's queue (rays, states) is read back before it is
        # traversed, into kept["light"] / kept["camera"] (tests)
        self.keep_queues = False
        self.kept = {"light": [], "camera": []}

    def _keep(self, which: str, cur: int, count_index: int):
        if not self.keep_queues:
            return
        c = self.counts.download(np.int32, 1, self.stream, offset=4 * count_index)
        self.stream.sync()
        n = int(c[0])
        rays = self.rays[cur].download(N.RAY_DTYPE, n, self.stream)
        states = self.states[cur].download(PPM_STATE_DTYPE, n, self.stream)
        self.stream.sync()
        self.kept[which].append((rays, states))

    def _cnt(self, j: int):
        return self.counts.at(4 * j)

    def _cfg(self):
        return C.byref(self.cfg) if self.cfg else None

    def _reset_counts(self):
        c = np.zeros(2 * self.L + 4, np.int32)
        c[0] = self.P
        c[self.L + 1] = self.n_cam
        c[2 * self.L + 2] = self.slots
        self._counts0 = c  # the host side of the copy stays alive until the stream has run it
        self.counts.upload(c, self.stream)

    def light_pass(self):
        st = self.stream.handle
        P, L = self.P, self.L
        N.check(N.lib.imba_ppm_light(N.ptr(self.light_tri), N.ptr(self.light_int), self.light_seed, 0, P,
                                     self.rays[0].ptr, self.states[0].ptr, st), "ppm_light")
        cur = 0
        self.kept["light"] = []
        for b in range(L):
            cnt = self._cnt(b)
            self._keep("light", cur, b)
            N.check(N.lib.imba_intersect_dev(self.gpu_scene.h, self.rays[cur].ptr, self.hits.ptr, P, cnt, self._cfg(),
                                             st), "intersect_dev")
            N.check(N.lib.imba_ppm_light_bounce(self.rays[cur].ptr, self.hits.ptr, self.states[cur].ptr, P, cnt,
                                                self.tri_n.ptr, self.tri_m.ptr, self.s.tri_count, self.albedo.ptr,
                                                self.kinds.ptr, len(self.s.albedo), L, self.rr_start,
                                                self.slot_rays.at(32 * b * P), self.slot_states.at(16 * b * P),
                                                self.rays[cur].ptr, st), "ppm_light_bounce")
            N.check(N.lib.imba_compact_rays_dev(self.rays[cur].ptr, self.states[cur].ptr, PPM_STATE_BYTES, P, cnt,
                                                self.rays[1 - cur].ptr, self.states[1 - cur].ptr, self._cnt(b + 1),
                                                self.workspace.ptr, st), "compact_rays_dev")
            cur = 1 - cur

    def compact_photons(self):
        N.check(N.lib.imba_compact_rays_dev(self.slot_rays.ptr, self.slot_states.ptr, 16, self.slots,
                                            self._cnt(2 * self.L + 2), self.photon_rays.ptr, self.photon_states.ptr,
                                            self._cnt(2 * self.L + 3), self.workspace.ptr, self.stream.handle),
                "compact photons")

    def read_photon_count(self) -> int:
        c = self.counts.download(np.int32, 1, self.stream, offset=4 * (2 * self.L + 3))
        self.stream.sync()
        self.photon_count = int(c[0])
        return self.photon_count

    def build_grid(self):
        self.grid.build(self.photon_rays, 32, 0, self.photon_count, self.radius, self.stream)
        if self.records is not None:
            N.check(N.lib.imba_ppm_merge_records(self.grid.h, self.photon_rays.ptr, self.photon_states.ptr,
                                                 self.records.ptr, self.stream.handle), "ppm_merge_records")

    def camera_pass(self):
        st = self.stream.handle
        n, L = self.n_cam, self.L
        N.check(N.lib.imba_ppm_camera(N.ptr(self.r2w), N.ptr(self.cam_pos), self.w, self.h, self.spp, self.seed_base,
                                      0, n, self.rays[0].ptr, self.states[0].ptr, self.queries.ptr, st), "ppm_camera")
        cur = 0
        self.kept["camera"] = []
        for b in range(L):
            cnt = self._cnt(L + 1 + b)
            self._keep("camera", cur, L + 1 + b)
            N.check(N.lib.imba_intersect_dev(self.gpu_scene.h, self.rays[cur].ptr, self.hits.ptr, n, cnt, self._cfg(),
                                             st), "intersect_dev")
            N.check(N.lib.imba_ppm_camera_bounce(self.rays[cur].ptr, self.hits.ptr, self.states[cur].ptr, n, cnt,
                                                 self.tri_n.ptr, self.tri_m.ptr, self.s.tri_count, self.albedo.ptr,
                                                 self.emission.ptr, self.kinds.ptr, len(self.s.albedo), L,
                                                 self.rr_start, self.spp, self.queries.ptr, self.image.ptr,
                                                 self.rays[cur].ptr, st), "ppm_camera_bounce")
            N.check(N.lib.imba_compact_rays_dev(self.rays[cur].ptr, self.states[cur].ptr, PPM_STATE_BYTES, n, cnt,
                                                self.rays[1 - cur].ptr, self.states[1 - cur].ptr,
                                                self._cnt(L + 2 + b), self.workspace.ptr, st), "compact_rays_dev")
            cur = 1 - cur

    def query(self):
        self.grid.query(self.queries, 64, 0, self.n_cam, self.k, self.out_idx, self.out_count, self.out_dist,
                        self.stream)

    def merge_grid(self):
        """The grid query and the merge in one launch (merge="fused")."""
        N.check(N.lib.imba_ppm_merge_grid(self.grid.h, self.records.ptr, self.queries.ptr, self.n_cam, self.k,
                                          self.radius, self.P, self.image.ptr, self.stream.handle), "ppm_merge_grid")

    def merge(self):
        N.check(N.lib.imba_ppm_merge(self.queries.ptr, self.n_cam, self.k, self.out_idx.ptr, self.out_dist.ptr,
                                     self.out_count.ptr, self.photon_rays.ptr, self.photon_states.ptr,
                                     self.radius, self.P, self.image.ptr, self.stream.handle), "ppm_merge")

    def clear_image(self):
        N.check(N.lib.imba_memset(self.image.ptr, 0, self.w * self.h * 12, self.stream.handle), "memset")

    def frame(self, timed: bool = False):
        """One frame; returns (light queue sizes, camera queue sizes, photon count).  With
        timed=True the phases are bracketed by events (see phase_ms)."""
        marks = []

        def mark(name):
            if timed:
                e = N.Event()
                e.record(self.stream)
                marks.append((name, e))

        self._reset_counts()
        mark("start")
        self.light_pass()
        mark("light")
        self.compact_photons()
        mark("photons")
        self.read_photon_count()
        mark("count")
        self.build_grid()
        mark("grid_build")
        self.camera_pass()
        mark("camera")
        if self.merge_mode == "fused":
            self.merge_grid()
        else:
            self.query()
            mark("query")
            self.merge()
        mark("merge")
        c = self.counts.download(np.int32, 2 * self.L + 4, self.stream)
        self.stream.sync()
        self.events = marks
        light = [int(v) for v in c[:self.L] if v > 0]
        cam = [int(v) for v in c[self.L + 1:2 * self.L + 1] if v > 0]
        return light, cam, self.photon_count

    def phase_ms(self) -> dict:
        out = {}
        for (_, e0), (name, e1) in zip(self.events, self.events[1:]):
            out[name] = e0.elapsed_ms(e1)
        return out

    def photons_host(self):
        n = self.photon_count
        return (self.photon_rays.download(PHOTON_RAY_DTYPE, n, self.stream),
                self.photon_states.download(PHOTON_STATE_DTYPE, n, self.stream))

    def queries_host(self):
        return self.queries.download(QUERY_DTYPE, self.n_cam, self.stream)

    def results_host(self):
        n, k = self.n_cam, self.k
        return (self.out_idx.download(np.int32, n * k, self.stream).reshape(n, k),
                self.out_dist.download(np.float32, n * k, self.stream).reshape(n, k),
                self.out_count.download(np.int32, n, self.stream))


class QueueLoop:
    """A bounce loop over many queues at once (the batched persistent traversal of many queues
    with a graph-captured bounce loop): per bounce one batched traversal of every queue
    (imba_trace_batch_launch) and one batched Lambertian bounce + compaction of every queue
    (imba_shade_batch_launch), queue sizes kept on the device.  A step restores every queue's
    first rays and states and runs max_path_len bounces; it is captured once into a graph.

    shading: per scene a PtShadingData (scenes without lights: no next-event rays); queues:
    list of (scene index, rays DeviceBuffer, ray count)."""

    def __init__(self, scenes: list, shading: list, queues: list, max_path_len: int = 8, rr_start: int = 3,
                 cfg: N.TraceConfig | None = None, seed: int = 0x5EED):
        from .scene import ShadeBatch, TraceBatch

        self.L, self.rr_start, self.cfg = max_path_len, rr_start, cfg
        self.stream = N.Stream()
        self.scenes = scenes
        self.shading = shading
        self.sh = [d.struct(max_path_len, rr_start) for d in shading]
        self.q = []
        L1 = max_path_len + 1
        self.counts = N.DeviceBuffer(4 * L1 * max(len(queues), 1))
        counts0 = np.zeros(L1 * len(queues), np.int32)
        rng = np.random.default_rng(seed)
        for k, (si, rays0, n) in enumerate(queues):
            st = np.zeros(max(n, 1), PT_STATE_DTYPE)
            st["pixel"] = np.arange(max(n, 1))
            st["rng"] = rng.integers(1, 2 ** 63, size=max(n, 1), dtype=np.int64).astype(np.uint64)
            st["tp"] = 1.0
            self.q.append({"scene": si, "n": n, "rays0": rays0, "states0": N.DeviceBuffer.from_array(st),
                           "rays": [N.DeviceBuffer(max(n, 1) * 32), N.DeviceBuffer(max(n, 1) * 32)],
                           "states": [N.DeviceBuffer(max(n, 1) * PT_STATE_BYTES), N.DeviceBuffer(max(n, 1) * PT_STATE_BYTES)],
                           "hits": N.DeviceBuffer(max(n, 1) * 16)})
            counts0[k * L1] = n
        self.counts0 = N.DeviceBuffer.from_array(counts0)
        self.n_counts = len(counts0)
        self.trace = []
        self.shade = []
        for b in range(max_path_len):
            cur, nxt = b % 2, (b + 1) % 2
            self.trace.append(TraceBatch([(scenes[e["scene"]], e["rays"][cur], e["hits"], e["n"], self._cnt(k, b),
                                           False) for k, e in enumerate(self.q)]))
            items = []
            for k, e in enumerate(self.q):
                items.append(N.ShadeItem(e["rays"][cur].ptr.value, e["hits"].ptr.value, e["states"][cur].ptr.value,
                                         e["n"], self._cnt(k, b).value, e["rays"][nxt].ptr.value,
                                         e["states"][nxt].ptr.value, self._cnt(k, b + 1).value, None, None, None,
                                         None, self.sh[e["scene"]]))
            self.shade.append(ShadeBatch(items))
        self.graph = None

    def _cnt(self, k: int, b: int):
        return self.counts.at(4 * (k * (self.L + 1) + b))

    def _enqueue(self):
        st = self.stream.handle
        N.check(N.lib.imba_memcpy_d2d(self.counts.ptr, self.counts0.ptr, 4 * self.n_counts, st), "counts")
        for e in self.q:
            if e["n"]:
                N.check(N.lib.imba_memcpy_d2d(e["rays"][0].ptr, e["rays0"].ptr, 32 * e["n"], st), "rays")
                N.check(N.lib.imba_memcpy_d2d(e["states"][0].ptr, e["states0"].ptr, PT_STATE_BYTES * e["n"], st), "states")
        for b in range(self.L):
            self.trace[b].launch(self.cfg, self.stream)
            self.shade[b].launch(self.stream)

    def build_graph(self):
        from .scene import Graph

        self._enqueue()  # sizes spill stacks and counters before capture
        self.stream.sync()
        Graph.begin(self.stream)
        try:
            self._enqueue()
        finally:
            self.graph = Graph.end(self.stream)

    def step(self, graph: bool = True):
        if graph:
            self.graph.launch(self.stream)
        else:
            self._enqueue()

    def rays_per_bounce(self) -> np.ndarray:
        """Rays traversed at each bounce by the last step (sum over the queues)."""
        c = self.counts.download(np.int32, self.n_counts, self.stream)
        self.stream.sync()
        return c.reshape(len(self.q), self.L + 1)[:, :self.L].sum(0)
