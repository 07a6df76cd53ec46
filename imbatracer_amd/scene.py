"""Python side of the library: scene assembly, resident scenes, traversal, compaction and the
photon hash grid, all through the C ABI (imbatracer_amd.native).

Arrays handed to the device functions are DeviceBuffer objects (device memory); host arrays
are numpy arrays.  Every call raises ImbaError on failure.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import native as N
from .native import (BUILDER_BINNED_SAH, BUILDER_SBVH, HIT_DTYPE, INSTANCE_DTYPE, MASK_DTYPE, NODE_DTYPE, RAY_DTYPE,
                     DeviceBuffer, check)

IDENTITY = np.eye(4, dtype=np.float32)


# ---------------------------------------------------------------------------------------------
# Synthetic inputs
def random_triangles(n: int, seed: int = 0x1234):
    """Random triangles in the unit cube (vertices (3n,4) float32, indices (n,4) uint32)."""
    v = np.empty((3 * n, 4), np.float32)
    i = np.empty((n, 4), np.uint32)
    check(N.lib.imba_gen_random_triangles(seed, n, N.ptr(v), N.ptr(i)), "imba_gen_random_triangles")
    return v, i


def icosphere(level: int, radius: float, centre, material: int = 0):
    nv, nt = C.c_int(), C.c_int()
    check(N.lib.imba_gen_icosphere_counts(level, C.byref(nv), C.byref(nt)), "icosphere counts")
    v = np.empty((nv.value, 4), np.float32)
    i = np.empty((nt.value, 4), np.uint32)
    check(N.lib.imba_gen_icosphere(level, radius, centre[0], centre[1], centre[2], material, N.ptr(v), N.ptr(i)),
          "imba_gen_icosphere")
    return v, i


def camera_rays(pos, direction, up, fov_deg: float, width: int, height: int) -> np.ndarray:
    rays = np.empty(width * height, RAY_DTYPE)
    p = np.asarray(pos, np.float32)
    d = np.asarray(direction, np.float32)
    u = np.asarray(up, np.float32)
    check(N.lib.imba_gen_camera_rays(N.ptr(p), N.ptr(d), N.ptr(u), fov_deg, width, height, N.ptr(rays)),
          "imba_gen_camera_rays")
    return rays


def make_rays(org: np.ndarray, dirs: np.ndarray, tmin, tmax) -> np.ndarray:
    n = len(org)
    r = np.empty(n, RAY_DTYPE)
    r["org"][:, :3] = org
    r["org"][:, 3] = tmin
    r["dir"][:, :3] = dirs
    r["dir"][:, 3] = tmax
    return r


# ---------------------------------------------------------------------------------------------
class SceneBuild:
    """Meshes + instances -> BVH2 traversal arrays (host)."""

    def __init__(self):
        self.h = C.c_void_p()
        check(N.lib.imba_build_create(C.byref(self.h)), "imba_build_create")
        self._keep = []

    def add_mesh(self, vertices: np.ndarray, indices: np.ndarray, texcoords: np.ndarray | None = None) -> int:
        v = np.ascontiguousarray(vertices, np.float32)
        if v.shape[1] == 3:
            v = np.ascontiguousarray(np.concatenate([v, np.ones((len(v), 1), np.float32)], 1))
        i = np.ascontiguousarray(indices, np.uint32)
        if i.ndim == 2 and i.shape[1] == 3:
            i = np.ascontiguousarray(np.concatenate([i, np.zeros((len(i), 1), np.uint32)], 1))
        tc = None if texcoords is None else np.ascontiguousarray(texcoords, np.float32)
        d = N.MeshDesc(v.ctypes.data, len(v), i.ctypes.data, len(i), None if tc is None else tc.ctypes.data)
        mid = C.c_int()
        check(N.lib.imba_build_add_mesh(self.h, C.byref(d), C.byref(mid)), "imba_build_add_mesh")
        return mid.value

    def add_instance(self, mesh_id: int, mat4=IDENTITY) -> int:
        m = np.ascontiguousarray(mat4, np.float32).reshape(16)
        iid = C.c_int()
        check(N.lib.imba_build_add_instance(self.h, mesh_id, N.ptr(m), C.byref(iid)), "imba_build_add_instance")
        return iid.value

    def set_mask(self, material: int, texels: np.ndarray):
        t = np.ascontiguousarray(texels, np.uint8)
        check(N.lib.imba_build_set_mask(self.h, material, t.shape[1], t.shape[0], N.ptr(t)), "imba_build_set_mask")

    def run(self, builder: int = BUILDER_SBVH, threads: int = 1):
        check(N.lib.imba_build_run(self.h, builder, threads), "imba_build_run")
        return self

    def desc(self) -> N.SceneDesc:
        d = N.SceneDesc()
        check(N.lib.imba_build_get_desc(self.h, C.byref(d)), "imba_build_get_desc")
        return d

    def cpu_layout(self, builder: int = BUILDER_SBVH, threads: int = 1):
        """Also writes the CPU layout (BVH4 nodes, 4-triangle packets) of the built scene."""
        check(N.lib.imba_build_cpu_layout(self.h, builder, threads), "imba_build_cpu_layout")
        return self

    def cpu_desc(self) -> N.CpuSceneDesc:
        d = N.CpuSceneDesc()
        check(N.lib.imba_build_get_cpu_desc(self.h, C.byref(d)), "imba_build_get_cpu_desc")
        return d

    def cpu_arrays(self) -> dict:
        """Copies of the CPU-layout host arrays (same keys as arrays(); nodes are NODE4_DTYPE)."""
        return self._copy(self.cpu_desc(), N.NODE4_DTYPE)

    def arrays(self) -> dict:
        """Copies of the host arrays."""
        return self._copy(self.desc(), NODE_DTYPE)

    @staticmethod
    def _copy(d, node_dtype) -> dict:
        def arr(p, n, dt):
            if n == 0 or not p:
                return np.zeros(0, dt)
            buf = (C.c_char * (n * np.dtype(dt).itemsize)).from_address(p)
            return np.frombuffer(buf, dtype=dt).copy()

        return {
            "root": d.root,
            "nodes": arr(d.nodes, d.node_count, node_dtype),
            "instances": arr(d.instances, d.instance_count, INSTANCE_DTYPE),
            "tris": arr(d.tris, d.tri_vec4_count, np.dtype(("<f4", 4))).reshape(-1, 4),
            "indices": arr(d.indices, d.index_count, np.int32),
            "texcoords": arr(d.texcoords, d.texcoord_count, np.dtype(("<f4", 2))).reshape(-1, 2),
            "masks": arr(d.masks, d.mask_count, MASK_DTYPE),
            "mask_buffer": arr(d.mask_buffer, d.mask_buffer_size, np.uint8),
        }

    def set_cache(self, mesh_id: int, path):
        """Cache file of a mesh's acceleration structures (loaded if present, else written)."""
        check(N.lib.imba_build_set_cache(self.h, mesh_id, None if path is None else str(path).encode()),
              "imba_build_set_cache")

    def cache_stats(self):
        """(meshes loaded from cache files, meshes built) by the last run / cpu_layout."""
        lo, bu = C.c_int(), C.c_int()
        check(N.lib.imba_build_cache_stats(self.h, C.byref(lo), C.byref(bu)), "imba_build_cache_stats")
        return lo.value, bu.value

    def tri_attrs(self):
        n = C.c_void_p()
        m = C.c_void_p()
        cnt = C.c_int()
        check(N.lib.imba_build_get_tri_attrs(self.h, C.byref(n), C.byref(m), C.byref(cnt)), "tri attrs")
        normals = np.frombuffer((C.c_float * (3 * cnt.value)).from_address(n.value), np.float32).reshape(-1, 3).copy()
        mats = np.frombuffer((C.c_int * cnt.value).from_address(m.value), np.int32).copy()
        return normals, mats

    def __del__(self):
        try:
            N.lib.imba_build_release(self.h)
        except Exception:
            pass


def desc_from_arrays(a: dict) -> tuple[N.SceneDesc, list]:
    """SceneDesc over host numpy arrays (the list keeps them alive)."""
    keep = []

    def p(x):
        x = np.ascontiguousarray(x)
        keep.append(x)
        return x.ctypes.data if x.size else None

    d = N.SceneDesc(a["root"], p(a["nodes"]), len(a["nodes"]), p(a["instances"]), len(a["instances"]),
                    p(a["tris"]), len(a["tris"]), p(a["indices"]), len(a["indices"]),
                    p(a["texcoords"]), len(a["texcoords"]), p(a["masks"]), len(a["masks"]),
                    p(a["mask_buffer"]), len(a["mask_buffer"]))
    return d, keep


def bvh_store(path, block_type: int, nodes: np.ndarray, tris: np.ndarray, node_offset: int = 0, tris_offset: int = 0,
              tri_id_offset: int = 0):
    """Writes one tree (node records, Vec4 records) as a cache-file block."""
    n = np.ascontiguousarray(nodes)
    t = np.ascontiguousarray(tris, np.float32).reshape(-1, 4)
    check(N.lib.imba_bvh_store(str(path).encode(), block_type, n.ctypes.data if n.size else None, len(n),
                               t.ctypes.data if t.size else None, len(t), node_offset, tris_offset, tri_id_offset),
          "imba_bvh_store")


def bvh_load(path, block_type: int, node_offset: int = 0, tris_offset: int = 0, tri_id_offset: int = 0):
    """Reads a cache-file block: (nodes, Vec4 records (n, 4) float32)."""
    nc, vc = C.c_int(), C.c_int()
    p = str(path).encode()
    code = N.lib.imba_bvh_load(p, block_type, node_offset, tris_offset, tri_id_offset, None, 0, None, 0,
                               C.byref(nc), C.byref(vc))
    if code not in (5, N.IMBA_OK):  # 5: capacity (counts filled in)
        check(code, "imba_bvh_load")
    dt = NODE_DTYPE if block_type == N.CACHE_BVH2 else N.NODE4_DTYPE
    nodes = np.empty(nc.value, dt)
    tris = np.empty((vc.value, 4), np.float32)
    check(N.lib.imba_bvh_load(p, block_type, node_offset, tris_offset, tri_id_offset, nodes.ctypes.data, len(nodes),
                              tris.ctypes.data, len(tris), C.byref(nc), C.byref(vc)), "imba_bvh_load")
    return nodes, tris


def wide_tree(arrays: dict) -> dict:
    """The 4-wide tree the upload derives from host arrays: nodes (NODE4_DTYPE), instances,
    root, max_stack."""
    d, keep = desc_from_arrays(arrays)
    h = C.c_void_p()
    check(N.lib.imba_wide_create(C.byref(d), C.byref(h)), "imba_wide_create")
    try:
        pn, pi = C.c_void_p(), C.c_void_p()
        nn, ni, root, ms = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(N.lib.imba_wide_get(h, C.byref(pn), C.byref(nn), C.byref(pi), C.byref(ni), C.byref(root), C.byref(ms)),
              "imba_wide_get")
        nodes = np.frombuffer((C.c_char * (nn.value * 128)).from_address(pn.value), N.NODE4_DTYPE).copy()
        inst = np.frombuffer((C.c_char * (ni.value * 64)).from_address(pi.value), INSTANCE_DTYPE).copy()
        pq, nq = C.c_void_p(), C.c_int()
        check(N.lib.imba_wide_get_quantized(h, C.byref(pq), C.byref(nq)), "imba_wide_get_quantized")
        qnodes = np.frombuffer((C.c_char * (nq.value * 64)).from_address(pq.value), N.QNODE4_DTYPE).copy() \
            if nq.value else np.zeros(0, N.QNODE4_DTYPE)
        return {"nodes": nodes, "qnodes": qnodes, "instances": inst, "root": root.value, "max_stack": ms.value}
    finally:
        N.lib.imba_wide_release(h)


def emit_mesh(vertices, indices, builder: int, width: int, cost: int) -> np.ndarray:
    v = np.ascontiguousarray(vertices, np.float32)
    i = np.ascontiguousarray(indices, np.uint32)
    d = N.MeshDesc(v.ctypes.data, len(v), i.ctypes.data, len(i), None)
    need = C.c_size_t()
    N.lib.imba_bvh_emit_mesh(C.byref(d), builder, width, cost, None, 0, C.byref(need))
    out = np.empty(need.value, np.int32)
    check(N.lib.imba_bvh_emit_mesh(C.byref(d), builder, width, cost, N.ptr(out), out.size, C.byref(need)), "emit")
    return out


def emit_boxes(boxes, centres, width: int, cost: int) -> np.ndarray:
    b = np.ascontiguousarray(boxes, np.float32)
    c = np.ascontiguousarray(centres, np.float32)
    need = C.c_size_t()
    N.lib.imba_bvh_emit_boxes(N.ptr(b), N.ptr(c), len(b), width, cost, None, 0, C.byref(need))
    out = np.empty(need.value, np.int32)
    check(N.lib.imba_bvh_emit_boxes(N.ptr(b), N.ptr(c), len(b), width, cost, N.ptr(out), out.size, C.byref(need)),
          "emit boxes")
    return out


# ---------------------------------------------------------------------------------------------
class Scene:
    """A device-resident, validated scene."""

    def __init__(self, desc: N.SceneDesc):
        self.h = C.c_void_p()
        check(N.lib.imba_scene_upload(C.byref(desc), C.byref(self.h)), "imba_scene_upload")

    def device_desc(self) -> N.SceneDesc:
        d = N.SceneDesc()
        check(N.lib.imba_scene_device_desc(self.h, C.byref(d)), "device desc")
        return d

    def footprint(self) -> int:
        s = C.c_size_t()
        check(N.lib.imba_scene_footprint(self.h, C.byref(s)), "footprint")
        return s.value

    def error_count(self) -> int:
        c = C.c_int()
        check(N.lib.imba_trace_error_count(self.h, C.byref(c)), "error count")
        return c.value

    def intersect(self, rays: DeviceBuffer, hits: DeviceBuffer, count: int, cfg: N.TraceConfig | None = None,
                  stream=None):
        check(N.lib.imba_intersect(self.h, rays.ptr, hits.ptr, count, C.byref(cfg) if cfg else None, N._s(stream)),
              "imba_intersect")

    def occluded(self, rays: DeviceBuffer, hits: DeviceBuffer, count: int, cfg: N.TraceConfig | None = None,
                 stream=None):
        check(N.lib.imba_occluded(self.h, rays.ptr, hits.ptr, count, C.byref(cfg) if cfg else None, N._s(stream)),
              "imba_occluded")

    def intersect_stats(self, rays: DeviceBuffer, hits: DeviceBuffer, counts: DeviceBuffer, count: int, any_hit=False,
                        stream=None):
        f = N.lib.imba_occluded_stats if any_hit else N.lib.imba_intersect_stats
        check(f(self.h, rays.ptr, hits.ptr, counts.ptr, count, N._s(stream)), "stats")

    def trace_stats(self, rays: DeviceBuffer, hits: DeviceBuffer, counts: DeviceBuffer, count: int, any_hit=False,
                    cfg: N.TraceConfig | None = None, stream=None):
        """Visit counts of the tree `cfg` selects (bvh_width 2 or 4)."""
        check(N.lib.imba_trace_stats(self.h, rays.ptr, hits.ptr, counts.ptr, count, int(any_hit),
                                     C.byref(cfg) if cfg else None, N._s(stream)), "imba_trace_stats")

    def stack_depths(self, rays: np.ndarray, any_hit: bool = False, cfg: N.TraceConfig | None = None) -> np.ndarray:
        """Rays of the host array `rays` by the deepest traversal stack entry they used (64 bins,
        the last one for 63 and deeper) on the tree `cfg` selects."""
        rb = DeviceBuffer.from_array(rays)
        hb = DeviceBuffer(max(len(rays), 1) * HIT_DTYPE.itemsize)
        hist = DeviceBuffer.from_array(np.zeros(64, np.uint32))
        check(N.lib.imba_trace_stack_depths(self.h, rb.ptr, hb.ptr, len(rays), int(any_hit),
                                            C.byref(cfg) if cfg else None, hist.ptr, None), "imba_trace_stack_depths")
        return hist.download(np.uint32, 64)

    def trace_host(self, rays: np.ndarray, any_hit: bool = False, cfg: N.TraceConfig | None = None) -> np.ndarray:
        """Convenience: upload rays, trace, download hits."""
        rb = DeviceBuffer.from_array(rays)
        hb = DeviceBuffer(len(rays) * HIT_DTYPE.itemsize)
        (self.occluded if any_hit else self.intersect)(rb, hb, len(rays), cfg)
        return hb.download(HIT_DTYPE, len(rays))

    def autotune(self, rays: DeviceBuffer, hits: DeviceBuffer, count: int, any_hit: bool = False, reps: int = 3,
                 stream=None):
        """Fastest launch shape on these rays: (TraceConfig, mean ms per launch)."""
        cfg, ms = N.TraceConfig(), C.c_float()
        check(N.lib.imba_autotune(self.h, rays.ptr, hits.ptr, count, int(any_hit), reps, C.byref(cfg), C.byref(ms),
                                  N._s(stream)), "imba_autotune")
        return cfg, ms.value

    def __del__(self):
        try:
            N.lib.imba_scene_release(self.h)
        except Exception:
            pass


class TraceBatch:
    """Many queues (scene, rays, hits, count[, device count], any_hit) traversed together."""

    def __init__(self, items):
        arr = self._items(items)
        self.h = C.c_void_p()
        check(N.lib.imba_trace_batch_create(arr, len(items), C.byref(self.h)), "imba_trace_batch_create")

    def _items(self, items):
        arr = (N.TraceItem * max(1, len(items)))()
        self._keep = []
        for k, it in enumerate(items):
            scene, rays, hits, count = it[:4]
            count_dev = it[4] if len(it) > 4 else None
            any_hit = it[5] if len(it) > 5 else False
            cdev = count_dev.ptr if isinstance(count_dev, DeviceBuffer) else count_dev
            rp = rays.ptr if isinstance(rays, DeviceBuffer) else rays
            hp = hits.ptr if isinstance(hits, DeviceBuffer) else hits
            arr[k] = N.TraceItem(scene.h, rp, hp, count, cdev, int(any_hit))
            self._keep.append(scene)
        return arr

    def update(self, items, stream=None):
        """Replaces the queues; the new table is copied on `stream` (launch on the same stream)."""
        arr = self._items(items)
        check(N.lib.imba_trace_batch_update(self.h, arr, len(items), N._s(stream)), "imba_trace_batch_update")

    def launch(self, cfg: N.TraceConfig | None = None, stream=None):
        check(N.lib.imba_trace_batch_launch(self.h, C.byref(cfg) if cfg else None, N._s(stream)),
              "imba_trace_batch_launch")

    def __del__(self):
        try:
            N.lib.imba_trace_batch_release(self.h)
        except Exception:
            pass


def _p(x):
    """Device pointer of a DeviceBuffer, a c_void_p or an int (None stays None)."""
    if x is None:
        return None
    if isinstance(x, DeviceBuffer):
        return x.ptr.value
    if isinstance(x, C.c_void_p):
        return x.value
    return x


class ShadeBatch:
    """Many queues shaded (path-tracing step) and compacted together: items are N.ShadeItem or
    dicts with its field names (DeviceBuffer / pointer values; "shading" an N.PtShading)."""

    def __init__(self, items):
        arr = self._items(items)
        self.h = C.c_void_p()
        check(N.lib.imba_shade_batch_create(arr, len(items), C.byref(self.h)), "imba_shade_batch_create")

    def _items(self, items):
        arr = (N.ShadeItem * max(1, len(items)))()
        for k, it in enumerate(items):
            if isinstance(it, N.ShadeItem):
                arr[k] = it
                continue
            e = N.ShadeItem()
            for name, _ in N.ShadeItem._fields_:
                v = it.get(name)
                if name == "capacity":
                    setattr(e, name, int(v or 0))
                elif name == "shading":
                    e.shading = v
                else:
                    setattr(e, name, _p(v))
            arr[k] = e
        return arr

    def update(self, items, stream=None):
        arr = self._items(items)
        check(N.lib.imba_shade_batch_update(self.h, arr, len(items), N._s(stream)), "imba_shade_batch_update")

    def launch(self, stream=None):
        check(N.lib.imba_shade_batch_launch(self.h, N._s(stream)), "imba_shade_batch_launch")

    def shadow(self, stream=None):
        check(N.lib.imba_shade_batch_shadow(self.h, N._s(stream)), "imba_shade_batch_shadow")

    def __del__(self):
        try:
            N.lib.imba_shade_batch_release(self.h)
        except Exception:
            pass


class Graph:
    """A captured sequence of library calls on one stream: `with Graph.capture(stream) as g:`."""

    def __init__(self, h):
        self.h = h

    @staticmethod
    def begin(stream):
        check(N.lib.imba_graph_begin(N._s(stream)), "imba_graph_begin")

    @staticmethod
    def end(stream) -> "Graph":
        h = C.c_void_p()
        check(N.lib.imba_graph_end(N._s(stream), C.byref(h)), "imba_graph_end")
        return Graph(h)

    def launch(self, stream=None):
        check(N.lib.imba_graph_launch(self.h, N._s(stream)), "imba_graph_launch")

    def __del__(self):
        try:
            N.lib.imba_graph_release(self.h)
        except Exception:
            pass


# ---------------------------------------------------------------------------------------------
def compact_hits(rays: DeviceBuffer, hits: DeviceBuffer, states: DeviceBuffer, state_stride: int, count: int,
                 rays_out: DeviceBuffer, hits_out: DeviceBuffer, states_out: DeviceBuffer, stream=None) -> int:
    n = C.c_int()
    check(N.lib.imba_compact_hits(rays.ptr, hits.ptr, states.ptr, state_stride, count, rays_out.ptr, hits_out.ptr,
                                  states_out.ptr, C.byref(n), N._s(stream)), "imba_compact_hits")
    return n.value


def compact_rays(rays: DeviceBuffer, states: DeviceBuffer, state_stride: int, count: int, rays_out: DeviceBuffer,
                 states_out: DeviceBuffer, stream=None) -> int:
    n = C.c_int()
    check(N.lib.imba_compact_rays(rays.ptr, states.ptr, state_stride, count, rays_out.ptr, states_out.ptr, C.byref(n),
                                  N._s(stream)), "imba_compact_rays")
    return n.value


class HashGrid:
    def __init__(self):
        self.h = C.c_void_p()
        check(N.lib.imba_hashgrid_create(C.byref(self.h)), "imba_hashgrid_create")

    def build(self, photons: DeviceBuffer, stride: int, pos_offset: int, count: int, radius: float, stream=None):
        check(N.lib.imba_hashgrid_build(self.h, photons.ptr, stride, pos_offset, count, radius, N._s(stream)),
              "imba_hashgrid_build")

    def query(self, queries: DeviceBuffer, stride: int, offset: int, nq: int, k: int, out_idx: DeviceBuffer,
              out_count: DeviceBuffer, out_dist: DeviceBuffer | None = None, stream=None):
        check(N.lib.imba_hashgrid_query(self.h, queries.ptr, stride, offset, nq, k, out_idx.ptr,
                                        out_dist.ptr if out_dist is not None else None, out_count.ptr, N._s(stream)),
              "imba_hashgrid_query")

    def reset(self):
        """Forget the table-size history; allocations are kept."""
        check(N.lib.imba_hashgrid_reset(self.h), "imba_hashgrid_reset")

    def info(self) -> N.HashGridInfo:
        i = N.HashGridInfo()
        check(N.lib.imba_hashgrid_info(self.h, C.byref(i)), "imba_hashgrid_info")
        return i

    def __del__(self):
        try:
            N.lib.imba_hashgrid_release(self.h)
        except Exception:
            pass
