"""One device-resident iteration of vertex connection and merging (VCMIntegrator<algo>::render,
render/integrators/vcm.cpp:24-56) over a SphereGridScene:

  light pass   light_paths paths from the triangle lights; per bounce: closest-hit traversal,
               imba_vcm_light_bounce (partial MIS weights, a light vertex per Lambertian hit into
               slot (bounce, entry), its connection to the camera into the light shadow slots,
               the bounce), any-hit traversal of the camera connections and their contributions,
               compaction of the continuing paths
  cache        the occupied vertex slots, stable-compacted (by bounce, then by path): the light
               vertex cache that connections draw from and the grid indexes
  grid         HashGrid over the cache positions, radius = radius_factor * the scene diagonal;
               merge records gathered into grid slot order
  camera pass  per bounce: closest-hit traversal, imba_vcm_camera_bounce (emitter hits,
               next-event and connection shadow rays, merge queries, the bounce), any-hit
               traversal of the shadow rays and their contributions, the fused grid query +
               merge of the bounce's queries, compaction

The sub-algorithms of the reference's template (vcm.h:37-43, 144-148) run the passes their
techniques need (vcm.cpp:47-51, 98-99): "vcm" all of them; "bpt" (eta_vm = 0) the light pass,
the cache and the camera pass without grid or merging; "lt" the light pass alone (no vertices
stored, camera connections with weight 1); "pt" the camera pass alone (next event with weight
1, no connections or merging).  PPM runs as workloads.PhotonMapper.

Queue sizes stay on the device; the host reads one number per frame (the cache size the grid
build needs).  Every launch is a HIP kernel of libimba_hip.so."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import native as N
from . import scene as S

VCM_STATE_DTYPE = np.dtype([("pixel", "<i4"), ("sample", "<i4"), ("rng", "<u8"), ("tp", "<f4", (3,)),
                            ("path_length", "<i4"), ("dVC", "<f4"), ("dVCM", "<f4"), ("dVM", "<f4"), ("pad", "<i4")])
assert VCM_STATE_DTYPE.itemsize == 48
VTX_A_DTYPE = np.dtype([("pos", "<f4", (3,)), ("path_length", "<i4"), ("out", "<f4", (3,)), ("tri", "<i4")])
VTX_B_DTYPE = np.dtype([("key", "<i4"), ("tp", "<f4", (3,)), ("dVCM", "<f4"), ("dVM", "<f4"), ("dVC", "<f4"),
                        ("mat", "<i4")])
MERGE_REC_DTYPE = np.dtype([("out", "<f4", (3,)), ("dVCM", "<f4"), ("tp", "<f4", (3,)), ("dVM", "<f4")])
VCM_QUERY_DTYPE = np.dtype([("pos", "<f4", (3,)), ("pixel", "<i4"), ("n", "<f4", (3,)), ("dVCM", "<f4"),
                            ("out", "<f4", (3,)), ("dVM", "<f4"), ("tp", "<f4", (3,)), ("mat", "<i4")])
SHADOW_STATE_DTYPE = np.dtype([("pixel", "<i4"), ("tp", "<f4", (3,))])
for _d in (VTX_A_DTYPE, VTX_B_DTYPE, MERGE_REC_DTYPE):
    assert _d.itemsize == 32
assert VCM_QUERY_DTYPE.itemsize == 64


class VcmFrame(C.Structure):
    """ImbaVcmFrame."""
    _fields_ = [("tri_normals", C.c_void_p), ("tri_materials", C.c_void_p), ("tri_count", C.c_int),
                ("albedo", C.c_void_p), ("emission", C.c_void_p), ("emitter_area", C.c_void_p),
                ("kinds", C.c_void_p), ("material_count", C.c_int), ("lights", C.c_void_p), ("light_count", C.c_int),
                ("raster_to_world", C.c_float * 16), ("world_to_raster", C.c_float * 16),
                ("cam_pos", C.c_float * 3), ("cam_dir", C.c_float * 3), ("image_plane_dist", C.c_float),
                ("width", C.c_int), ("height", C.c_int), ("spp", C.c_int), ("light_paths", C.c_int),
                ("radius", C.c_float), ("max_path_len", C.c_int), ("light_max_depth", C.c_int),
                ("rr_start", C.c_int), ("num_connections", C.c_int), ("algo", C.c_int)]


ALGOS = {"vcm": 0, "bpt": 1, "lt": 3, "pt": 4}  # IMBA_VCM_ALGO_*


def camera_frame(pos, d, up, fov, width, height):
    """(raster_to_world16, world_to_raster16, forward3, image_plane_dist) of the pinhole camera."""
    p, dd, uu = (np.ascontiguousarray(x, np.float32) for x in (pos, d, up))
    r2w, w2r, fwd = np.zeros(16, np.float32), np.zeros(16, np.float32), np.zeros(3, np.float32)
    ipd = C.c_float()
    N.check(N.lib.imba_gen_camera_matrix(N.ptr(p), N.ptr(dd), N.ptr(uu), fov, width, height, N.ptr(r2w)), "camera")
    N.check(N.lib.imba_gen_camera_frame(N.ptr(p), N.ptr(dd), N.ptr(uu), fov, width, height, N.ptr(w2r), N.ptr(fwd),
                                        C.byref(ipd)), "camera frame")
    return r2w, w2r, fwd, np.float32(ipd.value)


def make_frame(sh, kinds, cam, width: int, height: int, spp: int, light_paths: int, radius: float,
               max_path_len: int, light_max_depth: int, rr_start: int, num_connections: int, algo: str) -> VcmFrame:
    """ImbaVcmFrame over a scene's shading data (workloads.PtShadingData), its material kinds (a
    device buffer) and a camera (pos, raster_to_world16, world_to_raster16, forward3,
    image_plane_dist from camera_frame)."""
    pos, r2w, w2r, fwd, ipd = cam
    f = VcmFrame()
    f.tri_normals, f.tri_materials, f.tri_count = sh.tri_n.ptr.value, sh.tri_m.ptr.value, sh.tri_count
    f.albedo, f.emission, f.emitter_area = sh.albedo.ptr.value, sh.emission.ptr.value, sh.emitter_area.ptr.value
    f.kinds, f.material_count = kinds.ptr.value, sh.material_count
    f.lights, f.light_count = sh.lights.ptr.value, sh.light_count
    f.raster_to_world[:] = [float(x) for x in r2w]
    f.world_to_raster[:] = [float(x) for x in w2r]
    f.cam_pos[:] = [float(x) for x in pos]
    f.cam_dir[:] = [float(x) for x in fwd]
    f.image_plane_dist = float(ipd)
    f.width, f.height, f.spp, f.light_paths = width, height, spp, light_paths
    f.radius = radius
    f.max_path_len, f.light_max_depth, f.rr_start, f.num_connections = (max_path_len, light_max_depth, rr_start,
                                                                         num_connections)
    f.algo = ALGOS[algo]
    return f


class VcmRenderer:
    def __init__(self, scene, width: int, height: int, light_paths: int, spp: int = 1, max_path_len: int = 6,
                 light_max_depth: int = 6, rr_start: int = 0, radius_factor: float = 0.005, k: int = 10,
                 num_connections: int = 1, seed_base: int = 0x1234567, light_seed: int = 0x7654321,
                 trace_cfg: N.TraceConfig | None = None, shadow_cfg: N.TraceConfig | None = None,
                 gpu_scene: S.Scene | None = None, algo: str = "vcm", camera=None):
        """gpu_scene: an uploaded scene to share (another renderer over the same scene); algo:
        one of ALGOS; camera: (position, direction, up, fov in degrees), else the scene's."""
        from .scene import HashGrid
        from .workloads import PtShadingData

        if not 1 <= k <= 64:
            raise ValueError("the fused merge keeps its list in one wave: 1 <= k <= 64")
        if algo not in ALGOS:
            raise ValueError("algo must be one of %s" % sorted(ALGOS))
        self.algo = algo
        self.s = scene
        self.gpu_scene = gpu_scene if gpu_scene is not None else S.Scene(scene.desc)
        self.w, self.h, self.spp = width, height, spp
        self.n_cam = width * height * spp
        self.P = light_paths
        self.L = light_max_depth
        self.Lc = max_path_len
        self.k = k
        self.C = num_connections
        self.seed_base, self.light_seed = seed_base, light_seed
        self.cfg, self.shadow_cfg = trace_cfg, shadow_cfg
        lo, hi = np.asarray(scene.bounds[0], np.float64), np.asarray(scene.bounds[1], np.float64)
        self.radius = float(np.float32(radius_factor * np.linalg.norm(hi - lo)))
        pos, d, up, fov = camera if camera is not None else scene.camera
        self.r2w, self.w2r, self.fwd, self.ipd = camera_frame(pos, d, up, fov, width, height)
        self.cam_pos = np.asarray(pos, np.float32)
        self.sh = PtShadingData(scene.tri_normals, scene.tri_materials, scene.albedo, scene.emission,
                                scene.emitter_area, scene.lights)
        self.kinds = N.DeviceBuffer.from_array(np.ascontiguousarray(scene.kinds, np.int32))
        self.frame_args = make_frame(self.sh, self.kinds, (self.cam_pos, self.r2w, self.w2r, self.fwd, self.ipd),
                                     width, height, spp, light_paths, self.radius, max_path_len, light_max_depth,
                                     rr_start, num_connections, algo)
        cap = max(self.P, self.n_cam)
        self.cap = cap
        self.rays = [N.DeviceBuffer(cap * 32), N.DeviceBuffer(cap * 32)]
        self.states = [N.DeviceBuffer(cap * 48), N.DeviceBuffer(cap * 48)]
        self.hits = N.DeviceBuffer(cap * 16)
        ns = (1 + self.C) * cap
        self.shadow_rays = N.DeviceBuffer(ns * 32)
        self.shadow_states = N.DeviceBuffer(ns * 16)
        self.shadow_hits = N.DeviceBuffer(ns * 16)
        # the occupied shadow slots, stable-compacted before their traversal (a wave then holds
        # no empty slots beside live rays)
        self.compact_shadows = True
        self.shadow_rays_c = N.DeviceBuffer(ns * 32)
        self.shadow_states_c = N.DeviceBuffer(ns * 16)
        self.slots = self.L * self.P
        self.slot_a = N.DeviceBuffer(self.slots * 32)
        self.slot_b = N.DeviceBuffer(self.slots * 32)
        self.vtx_a = N.DeviceBuffer(self.slots * 32)
        self.vtx_b = N.DeviceBuffer(self.slots * 32)
        self.records = N.DeviceBuffer(self.slots * 32)
        self.queries = N.DeviceBuffer(cap * 64)
        self.image = N.DeviceBuffer(width * height * 12)
        # device counts: light queues [0, L], camera queues [L+1, L+1+Lc], slot count, vertex
        # count, shadow count
        self.n_counts = self.L + self.Lc + 6
        self.counts = N.DeviceBuffer(4 * self.n_counts)
        self.workspace = N.DeviceBuffer(int(N.lib.imba_compact_workspace_size(max(self.slots, ns))))
        self.grid = HashGrid()
        self.stream = N.Stream()
        self.vertex_count = 0
        self.events = []
        # keep_queues: every bounce's inputs and outputs read back (tests)
        self.keep_queues = False
        self.kept = {"light": [], "camera": []}
        # time_kernels: events around every traversal, shadow traversal and merge launch
        # (kernel_ms() sums them per kind)
        self.time_kernels = False
        self.kernel_events = []
        # host side of the device counts (pinned: copies on the stream do not block the host)
        self._pinned_counts0 = N.PinnedArray(self.n_counts, np.int32)
        self._pinned_counts = N.PinnedArray(self.n_counts, np.int32)
        self._ev_cache = N.Event()
        self._marks = []
        self._timed_frame = False
        self.clear_image()

    def _timed(self, kind: str, fn):
        if not self.time_kernels:
            fn()
            return
        e0, e1 = N.Event(), N.Event()
        e0.record(self.stream)
        fn()
        e1.record(self.stream)
        self.kernel_events.append((kind, e0, e1))

    def kernel_ms(self) -> dict:
        """Milliseconds per kind of launch in the last frame with time_kernels set."""
        out = {}
        for kind, e0, e1 in self.kernel_events:
            out[kind] = out.get(kind, 0.0) + e0.elapsed_ms(e1)
        return out

    # device count slots
    def _cnt(self, j: int):
        return self.counts.at(4 * j)

    def _light_cnt(self, b):
        return self._cnt(b)

    def _cam_cnt(self, b):
        return self._cnt(self.L + 1 + b)

    def _slot_cnt(self):
        return self._cnt(self.L + self.Lc + 2)

    def _vtx_cnt(self):
        return self._cnt(self.L + self.Lc + 3)

    def _shadow_cnt(self):
        return self._cnt(self.L + self.Lc + 4)

    def _live_shadow_cnt(self):
        return self._cnt(self.L + self.Lc + 5)

    def _slot_hits(self, shadow_states):
        """The any-hit answers of the last shadow traversal in slot order (tests): with
        compact_shadows, the occupied slots' answers scattered back to their slots (empty
        slots read tri_id -1)."""
        n = len(shadow_states)
        if not self.compact_shadows:
            return self.shadow_hits.download(N.HIT_DTYPE, n, self.stream)
        live = np.nonzero(shadow_states["pixel"] >= 0)[0]
        out = np.zeros(n, N.HIT_DTYPE)
        out["tri_id"] = -1
        out["inst_id"] = -1
        out[live] = self.shadow_hits.download(N.HIT_DTYPE, len(live), self.stream)
        return out

    def _trace_shadows(self, capacity: int, count_dev):
        """Any-hit traversal of the shadow slots [0, *count_dev) and the unoccluded ones'
        contributions.  With compact_shadows the occupied slots are first stable-compacted
        (shadow_rays_c / shadow_states_c, their count in the live-shadow count slot) and
        shadow_hits holds the answers in that order; otherwise in slot order."""
        st = self.stream.handle
        rays, states, cnt = self.shadow_rays, self.shadow_states, count_dev
        if self.compact_shadows:
            def compact_and_trace():
                N.check(N.lib.imba_compact_rays_dev(self.shadow_rays.ptr, self.shadow_states.ptr, 16, capacity,
                                                    count_dev, self.shadow_rays_c.ptr, self.shadow_states_c.ptr,
                                                    self._live_shadow_cnt(), self.workspace.ptr, st), "compact shadows")
                N.check(N.lib.imba_occluded_dev(self.gpu_scene.h, self.shadow_rays_c.ptr, self.shadow_hits.ptr,
                                                capacity, self._live_shadow_cnt(), self._cfg(self.shadow_cfg), st),
                        "occluded_dev")
            self._timed("shadow", compact_and_trace)
            rays, states, cnt = self.shadow_rays_c, self.shadow_states_c, self._live_shadow_cnt()
        else:
            self._timed("shadow", lambda: N.check(N.lib.imba_occluded_dev(
                self.gpu_scene.h, rays.ptr, self.shadow_hits.ptr, capacity, cnt, self._cfg(self.shadow_cfg), st),
                "occluded_dev"))
        N.check(N.lib.imba_pt_shadow(self.shadow_hits.ptr, states.ptr, capacity, cnt, self.image.ptr, st),
                "pt_shadow")

    def _cfg(self, cfg):
        return C.byref(cfg) if cfg else None

    def _count_host(self, j: int) -> int:
        c = self.counts.download(np.int32, 1, self.stream, offset=4 * j)
        self.stream.sync()
        return int(c[0])

    def clear_image(self):
        N.check(N.lib.imba_memset(self.image.ptr, 0, self.w * self.h * 12, self.stream.handle), "memset")

    def _reset_counts(self):
        c = self._pinned_counts0.array
        c[:] = 0
        c[0] = self.P if self.algo != "pt" else 0
        c[self.L + 1] = self.n_cam if self.algo != "lt" else 0
        c[self.L + self.Lc + 2] = self.slots
        N.check(N.lib.imba_memcpy_h2d(self.counts.ptr, self._pinned_counts0.ptr, 4 * self.n_counts,
                                      self.stream.handle), "counts")

    def light_pass(self):
        st = self.stream.handle
        P, fa = self.P, C.byref(self.frame_args)
        N.check(N.lib.imba_vcm_light(fa, self.light_seed, 0, P, self.rays[0].ptr, self.states[0].ptr, st), "vcm_light")
        cur = 0
        self.kept["light"] = []
        for b in range(self.L):
            cnt = self._light_cnt(b)
            keep = None
            if self.keep_queues:
                n = self._count_host(b)
                keep = {"rays": self.rays[cur].download(N.RAY_DTYPE, n, self.stream),
                        "states": self.states[cur].download(VCM_STATE_DTYPE, n, self.stream)}
            self._timed("trace", lambda: N.check(N.lib.imba_intersect_dev(
                self.gpu_scene.h, self.rays[cur].ptr, self.hits.ptr, P, cnt, self._cfg(self.cfg), st), "intersect_dev"))
            if keep is not None:
                keep["hits"] = self.hits.download(N.HIT_DTYPE, len(keep["rays"]), self.stream)
            N.check(N.lib.imba_vcm_light_bounce(fa, self.rays[cur].ptr, self.hits.ptr, self.states[cur].ptr, P, cnt,
                                                self.slot_a.at(32 * b * P), self.slot_b.at(32 * b * P),
                                                self.rays[cur].ptr, self.shadow_rays.ptr, self.shadow_states.ptr, st),
                    "vcm_light_bounce")
            self._trace_shadows(P, cnt)
            if keep is not None:
                n = len(keep["rays"])
                keep["out_rays"] = self.rays[cur].download(N.RAY_DTYPE, n, self.stream)
                keep["out_states"] = self.states[cur].download(VCM_STATE_DTYPE, n, self.stream)
                keep["vtx_a"] = self.slot_a.download(VTX_A_DTYPE, n, self.stream, offset=32 * b * P)
                keep["vtx_b"] = self.slot_b.download(VTX_B_DTYPE, n, self.stream, offset=32 * b * P)
                keep["shadow_rays"] = self.shadow_rays.download(N.RAY_DTYPE, n, self.stream)
                keep["shadow_states"] = self.shadow_states.download(SHADOW_STATE_DTYPE, n, self.stream)
                keep["shadow_hits"] = self._slot_hits(keep["shadow_states"])
                self.kept["light"].append(keep)
            N.check(N.lib.imba_compact_rays_dev(self.rays[cur].ptr, self.states[cur].ptr, 48, P, cnt,
                                                self.rays[1 - cur].ptr, self.states[1 - cur].ptr,
                                                self._light_cnt(b + 1), self.workspace.ptr, st), "compact_rays_dev")
            cur = 1 - cur

    def build_cache(self):
        """The light vertex cache: the occupied slots compacted, the cache size read back (the
        grid build sizes its table from it), the grid and the merge records."""
        self.compact_cache()
        self.read_cache_size()
        self.build_grid()

    def compact_cache(self):
        st = self.stream.handle
        N.check(N.lib.imba_compact_rays_dev(self.slot_a.ptr, self.slot_b.ptr, 32, self.slots, self._slot_cnt(),
                                            self.vtx_a.ptr, self.vtx_b.ptr, self._vtx_cnt(), self.workspace.ptr, st),
                "compact vertices")
        # the cache size into pinned memory (the copy does not block the host), and an event after it
        N.check(N.lib.imba_memcpy_d2h(self._pinned_counts.ptr, self.counts.ptr, 4 * self.n_counts, st), "counts")
        self._ev_cache.record(self.stream)

    def read_cache_size(self):
        self._ev_cache.sync()
        self.vertex_count = int(self._pinned_counts.array[self.L + self.Lc + 3])

    def build_grid(self):
        st = self.stream.handle
        self.grid.build(self.vtx_a, 32, 0, self.vertex_count, self.radius, self.stream)
        N.check(N.lib.imba_vcm_merge_records(self.grid.h, self.vtx_a.ptr, self.vtx_b.ptr, self.records.ptr, st),
                "vcm_merge_records")

    def camera_pass(self):
        st = self.stream.handle
        n, fa = self.n_cam, C.byref(self.frame_args)
        ns = (1 + self.C) * n
        N.check(N.lib.imba_vcm_camera(fa, self.seed_base, 0, n, self.rays[0].ptr, self.states[0].ptr, st),
                "vcm_camera")
        cur = 0
        self.kept["camera"] = []
        for b in range(self.Lc):
            cnt = self._cam_cnt(b)
            keep = None
            if self.keep_queues:
                m = self._count_host(self.L + 1 + b)
                keep = {"rays": self.rays[cur].download(N.RAY_DTYPE, m, self.stream),
                        "states": self.states[cur].download(VCM_STATE_DTYPE, m, self.stream)}
            self._timed("trace", lambda: N.check(N.lib.imba_intersect_dev(
                self.gpu_scene.h, self.rays[cur].ptr, self.hits.ptr, n, cnt, self._cfg(self.cfg), st), "intersect_dev"))
            if keep is not None:
                keep["hits"] = self.hits.download(N.HIT_DTYPE, len(keep["rays"]), self.stream)
                keep["image_before"] = self.image.download(np.float32, self.w * self.h * 3, self.stream)
            N.check(N.lib.imba_vcm_camera_bounce(fa, self.rays[cur].ptr, self.hits.ptr, self.states[cur].ptr, n, cnt,
                                                 self.vtx_a.ptr, self.vtx_b.ptr, self._vtx_cnt(), self.rays[cur].ptr,
                                                 self.shadow_rays.ptr, self.shadow_states.ptr, self._shadow_cnt(),
                                                 self.queries.ptr, self.image.ptr, st), "vcm_camera_bounce")
            if keep is not None:
                m = len(keep["rays"])
                keep["image_emit"] = self.image.download(np.float32, self.w * self.h * 3, self.stream)
            self._trace_shadows(ns, self._shadow_cnt())
            if keep is not None:
                keep["image_shadow"] = self.image.download(np.float32, self.w * self.h * 3, self.stream)
            if self.algo == "vcm":
                self._timed("merge", lambda: N.check(N.lib.imba_vcm_merge(
                    fa, self.grid.h, self.records.ptr, self.queries.ptr, n, cnt, self.k, self.image.ptr, st),
                    "vcm_merge"))
            if keep is not None:
                keep["image_merge"] = self.image.download(np.float32, self.w * self.h * 3, self.stream)
                keep["out_rays"] = self.rays[cur].download(N.RAY_DTYPE, m, self.stream)
                keep["out_states"] = self.states[cur].download(VCM_STATE_DTYPE, m, self.stream)
                keep["shadow_rays"] = self.shadow_rays.download(N.RAY_DTYPE, (1 + self.C) * m, self.stream)
                keep["shadow_states"] = self.shadow_states.download(SHADOW_STATE_DTYPE, (1 + self.C) * m, self.stream)
                keep["shadow_hits"] = self._slot_hits(keep["shadow_states"])
                keep["queries"] = self.queries.download(VCM_QUERY_DTYPE, m, self.stream)
                self.kept["camera"].append(keep)
            N.check(N.lib.imba_compact_rays_dev(self.rays[cur].ptr, self.states[cur].ptr, 48, n, cnt,
                                                self.rays[1 - cur].ptr, self.states[1 - cur].ptr,
                                                self._cam_cnt(b + 1), self.workspace.ptr, st), "compact_rays_dev")
            cur = 1 - cur

    def _mark(self, name):
        if self._timed_frame:
            e = N.Event()
            e.record(self.stream)
            self._marks.append((name, e))

    def launch_light(self, timed: bool = False):
        """First half of an iteration: the light pass and the light vertex cache, enqueued; the
        cache size is copied to the host behind an event."""
        self._timed_frame = timed
        self._marks = []
        self.kernel_events = []
        self._reset_counts()
        self._mark("start")
        if self.algo != "pt":
            self.light_pass()
        self._mark("light")
        if self.algo in ("vcm", "bpt"):
            self.compact_cache()

    def launch_camera(self):
        """Second half: waits for the cache size only, then enqueues the grid, the camera pass
        and the read-back of every queue size."""
        if self.algo in ("vcm", "bpt"):
            self.read_cache_size()
        if self.algo == "vcm":
            self.build_grid()
        self._mark("cache_grid")
        if self.algo != "lt":
            self.camera_pass()
        self._mark("camera")
        N.check(N.lib.imba_memcpy_d2h(self._pinned_counts.ptr, self.counts.ptr, 4 * self.n_counts,
                                      self.stream.handle), "counts")

    def finish(self):
        """Waits for the iteration; returns (light queue sizes, camera queue sizes, light vertex
        count)."""
        self.stream.sync()
        self.events = self._marks
        c = self._pinned_counts.array
        light = [int(v) for v in c[:self.L] if v > 0]
        cam = [int(v) for v in c[self.L + 1:self.L + 1 + self.Lc] if v > 0]
        return light, cam, self.vertex_count

    def frame(self, timed: bool = False):
        """One iteration; returns (light queue sizes, camera queue sizes, light vertex count).
        With timed=True the phases are bracketed by events (phase_ms).  Iterations of several
        renderers on their own streams overlap when their launch_light / launch_camera /
        finish calls interleave."""
        self.launch_light(timed)
        self.launch_camera()
        return self.finish()

    def phase_ms(self) -> dict:
        out = {}
        for (_, e0), (name, e1) in zip(self.events, self.events[1:]):
            out[name] = e0.elapsed_ms(e1)
        return out

    def cache_host(self):
        n = self.vertex_count
        return (self.vtx_a.download(VTX_A_DTYPE, n, self.stream), self.vtx_b.download(VTX_B_DTYPE, n, self.stream))

    def records_host(self):
        return self.records.download(MERGE_REC_DTYPE, self.vertex_count, self.stream)

    def image_host(self):
        return self.image.download(np.float32, self.w * self.h * 3, self.stream).reshape(self.h, self.w, 3)
