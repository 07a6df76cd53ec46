"""CPU oracle — TEST INFRASTRUCTURE ONLY.

ctypes bindings of oracle/liboracle.so (plain-C restatement of the reference's hot path) and,
where it was built, of oracle/_ref/libref_builders.so (the reference's own BVH builders).
Only tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py use this package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref_builders.so")
REF_SRC = "/root/reference/src"

RAY_DTYPE = np.dtype([("org", "<f4", (4,)), ("dir", "<f4", (4,))])
HIT_DTYPE = np.dtype([("tri_id", "<i4"), ("inst_id", "<i4"), ("tmax", "<f4"), ("u", "<f4")])


def build(with_ref: bool = True):
    """Builds liboracle.so, and the reference-builder harness when the reference is present."""
    subprocess.run(["make", "-s", "-B", "-C", HERE, "liboracle.so"], check=True)
    if with_ref and os.path.isdir(REF_SRC):
        subprocess.run(["make", "-s", "-B", "-C", HERE, "ref", "REF=" + REF_SRC], check=True)


class SceneDesc(C.Structure):
    _fields_ = [("root", C.c_int),
                ("nodes", C.c_void_p), ("node_count", C.c_int),
                ("instances", C.c_void_p), ("instance_count", C.c_int),
                ("tris", C.c_void_p), ("tri_vec4_count", C.c_int),
                ("indices", C.c_void_p), ("index_count", C.c_int),
                ("texcoords", C.c_void_p), ("texcoord_count", C.c_int),
                ("masks", C.c_void_p), ("mask_count", C.c_int),
                ("mask_buffer", C.c_void_p), ("mask_buffer_size", C.c_int)]


class OracleGrid(C.Structure):
    _fields_ = [("photon_count", C.c_int), ("table_size", C.c_int),
                ("bbox_min", C.c_float * 3), ("bbox_max", C.c_float * 3),
                ("radius", C.c_float), ("radius_sqr", C.c_float), ("cell_size", C.c_float),
                ("inv_cell_size", C.c_float),
                ("cell_ends", C.POINTER(C.c_int)), ("order", C.POINTER(C.c_int)), ("positions", C.POINTER(C.c_float))]


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise ImportError("oracle/liboracle.so is not built (run make -C oracle)")
        L = C.CDLL(LIB)
        vp, i32 = C.c_void_p, C.c_int
        L.oracle_trace.argtypes = [C.POINTER(SceneDesc), vp, i32, i32, vp, vp]
        L.oracle_trace.restype = i32
        L.oracle_trace_mt.argtypes = [C.POINTER(SceneDesc), vp, i32, i32, vp, i32]
        L.oracle_trace_mt.restype = i32
        L.cpu4_trace.argtypes = [C.POINTER(SceneDesc), vp, i32, i32, vp]
        L.cpu4_trace.restype = i32
        L.cpu4_trace_mt.argtypes = [C.POINTER(SceneDesc), vp, i32, i32, vp, i32]
        L.cpu4_trace_mt.restype = i32
        L.oracle_brute_closest.argtypes = [vp, i32, vp, i32, vp, vp]
        L.oracle_brute_closest.restype = None
        L.oracle_compact_hits.argtypes = [vp, vp, vp, i32, i32]
        L.oracle_compact_hits.restype = i32
        L.oracle_compact_rays.argtypes = [vp, vp, i32, i32]
        L.oracle_compact_rays.restype = i32
        L.oracle_sort_by_material.argtypes = [vp, i32, vp, vp, i32, i32, vp, vp]
        L.oracle_sort_by_material.restype = None
        L.oracle_grid_build.argtypes = [C.POINTER(OracleGrid), vp, i32, i32, C.c_float]
        L.oracle_grid_build.restype = None
        L.oracle_grid_query.argtypes = [C.POINTER(OracleGrid), vp, i32, vp, vp]
        L.oracle_grid_query.restype = i32
        L.oracle_grid_query_many.argtypes = [C.POINTER(OracleGrid), vp, i32, i32, i32, vp, vp, vp]
        L.oracle_grid_query_many.restype = None
        L.oracle_grid_free.argtypes = [C.POINTER(OracleGrid)]
        L.oracle_grid_free.restype = None
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_LIB)


def ref_lib():
    global _ref
    if _ref is None:
        R = C.CDLL(REF_LIB)
        R.ref_emit_mesh.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                    C.c_size_t, C.POINTER(C.c_size_t)]
        R.ref_emit_mesh.restype = C.c_int
        R.ref_emit_boxes.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_size_t,
                                     C.POINTER(C.c_size_t)]
        R.ref_emit_boxes.restype = C.c_int
        R.ref_rng.argtypes = [C.c_uint64, C.c_int, C.c_void_p, C.c_void_p]
        R.ref_rng.restype = None
        R.ref_random_triangles.argtypes = [C.c_uint64, C.c_int, C.c_void_p]
        R.ref_random_triangles.restype = None
        _ref = R
    return _ref


def _p(a):
    return C.c_void_p(a.ctypes.data)


def scene_desc(arrays: dict):
    keep = []

    def p(x):
        x = np.ascontiguousarray(x)
        keep.append(x)
        return x.ctypes.data if x.size else None

    d = SceneDesc(arrays["root"], p(arrays["nodes"]), len(arrays["nodes"]), p(arrays["instances"]),
                  len(arrays["instances"]), p(arrays["tris"]), len(arrays["tris"]), p(arrays["indices"]),
                  len(arrays["indices"]), p(arrays["texcoords"]), len(arrays["texcoords"]), p(arrays["masks"]),
                  len(arrays["masks"]), p(arrays["mask_buffer"]), len(arrays["mask_buffer"]))
    return d, keep


def trace(arrays: dict, rays: np.ndarray, any_hit: bool = False, counts: bool = False, threads: int = 1):
    """Returns hits (and per-ray visit counts (n,3) when counts=True)."""
    d, keep = scene_desc(arrays)
    rays = np.ascontiguousarray(rays)
    hits = np.empty(len(rays), HIT_DTYPE)
    if counts:
        cnt = np.empty((len(rays), 3), np.uint32)
        bad = lib().oracle_trace(C.byref(d), _p(rays), len(rays), int(any_hit), _p(hits), _p(cnt))
        return hits, cnt, bad
    if threads > 1:
        bad = lib().oracle_trace_mt(C.byref(d), _p(rays), len(rays), int(any_hit), _p(hits), threads)
    else:
        bad = lib().oracle_trace(C.byref(d), _p(rays), len(rays), int(any_hit), _p(hits), None)
    return hits, bad


def trace_cpu4(arrays: dict, rays: np.ndarray, any_hit: bool = False, threads: int = 1):
    """The CPU-layout traversal (BVH4 + 4-triangle packets; arrays from SceneBuild.cpu_arrays()).
    Returns (hits, number of rays that met a malformed structure)."""
    d, keep = scene_desc(arrays)
    rays = np.ascontiguousarray(rays)
    hits = np.empty(len(rays), HIT_DTYPE)
    if threads > 1:
        bad = lib().cpu4_trace_mt(C.byref(d), _p(rays), len(rays), int(any_hit), _p(hits), threads)
    else:
        bad = lib().cpu4_trace(C.byref(d), _p(rays), len(rays), int(any_hit), _p(hits))
    return hits, bad


def sort_by_material(indices, rays, hits, num_mats):
    """(rays with dir.w = material key, sorted indices, bucket starts)."""
    ind = np.ascontiguousarray(indices, np.int32)
    rays = np.ascontiguousarray(rays).copy()
    hits = np.ascontiguousarray(hits)
    out = np.empty(len(rays), np.int32)
    starts = np.empty(num_mats + 1, np.int32)
    lib().oracle_sort_by_material(_p(ind), ind.size // 4, _p(rays), _p(hits), len(rays), num_mats, _p(out),
                                  _p(starts))
    return rays, out, starts


def brute_closest(tri_verts: np.ndarray, rays: np.ndarray):
    tv = np.ascontiguousarray(tri_verts, np.float32).reshape(-1, 9)
    rays = np.ascontiguousarray(rays)
    ids = np.empty(len(rays), np.int32)
    ts = np.empty(len(rays), np.float64)
    lib().oracle_brute_closest(_p(tv), len(tv), _p(rays), len(rays), _p(ids), _p(ts))
    return ids, ts


def _record_bytes(states):
    return states.dtype.itemsize * (1 if states.ndim == 1 else int(np.prod(states.shape[1:])))


def compact_hits(rays, hits, states):
    rays, hits, states = rays.copy(), hits.copy(), np.ascontiguousarray(states).copy()
    stride = _record_bytes(states)
    n = lib().oracle_compact_hits(_p(rays), _p(hits), _p(states), stride, len(rays))
    return n, rays, hits, states


def compact_rays(rays, states):
    rays, states = rays.copy(), np.ascontiguousarray(states).copy()
    stride = _record_bytes(states)
    n = lib().oracle_compact_rays(_p(rays), _p(states), stride, len(rays))
    return n, rays, states


class Grid:
    """Serial HashGrid restatement; keeps its bucket table across builds like the reference."""

    def __init__(self):
        self.g = OracleGrid()

    def build(self, positions: np.ndarray, radius: float):
        pos = np.ascontiguousarray(positions, np.float32)
        stride = pos.shape[1] if pos.ndim == 2 else 3
        lib().oracle_grid_build(C.byref(self.g), _p(pos), stride, len(pos), radius)

    def query(self, q, k: int):
        q = np.ascontiguousarray(q, np.float32)
        idx = np.empty(max(k, 1), np.int32)
        dist = np.empty(max(k, 1), np.float32)
        n = lib().oracle_grid_query(C.byref(self.g), _p(q), k, _p(idx), _p(dist))
        return idx[:n].copy(), dist[:n].copy()

    def query_many(self, q, k: int):
        """Queries for every row of q (n x stride floats, position first): (idx, dist, count)."""
        q = np.ascontiguousarray(q, np.float32)
        n = len(q)
        idx = np.empty((n, max(k, 1)), np.int32)
        dist = np.empty((n, max(k, 1)), np.float32)
        cnt = np.empty(n, np.int32)
        lib().oracle_grid_query_many(C.byref(self.g), _p(q), q.shape[1] if q.ndim == 2 else 3, n, k, _p(idx), _p(dist),
                                     _p(cnt))
        return idx, dist, cnt

    def cell_ends(self):
        return np.ctypeslib.as_array(self.g.cell_ends, (self.g.table_size,)).copy()

    def order(self):
        return np.ctypeslib.as_array(self.g.order, (self.g.photon_count,)).copy()

    def __del__(self):
        try:
            lib().oracle_grid_free(C.byref(self.g))
        except Exception:
            pass


def ref_emit_mesh(vertices, indices, builder: int, width: int, cost: int) -> np.ndarray:
    v = np.ascontiguousarray(vertices, np.float32)
    i = np.ascontiguousarray(indices, np.uint32)
    need = C.c_size_t()
    R = ref_lib()
    R.ref_emit_mesh(_p(v), len(v), _p(i), len(i), builder, width, cost, None, 0, C.byref(need))
    out = np.empty(need.value, np.int32)
    R.ref_emit_mesh(_p(v), len(v), _p(i), len(i), builder, width, cost, _p(out), out.size, C.byref(need))
    return out


def ref_rng(seed: int, n: int):
    raw = np.empty(n, np.uint32)
    uni = np.empty(n, np.float32)
    ref_lib().ref_rng(seed, n, _p(raw), _p(uni))
    return raw, uni


def ref_random_triangles(seed: int, n: int) -> np.ndarray:
    v = np.empty((3 * n, 4), np.float32)
    ref_lib().ref_random_triangles(seed, n, _p(v))
    return v


def ref_emit_boxes(boxes, centres, width: int, cost: int) -> np.ndarray:
    b = np.ascontiguousarray(boxes, np.float32)
    c = np.ascontiguousarray(centres, np.float32)
    need = C.c_size_t()
    R = ref_lib()
    R.ref_emit_boxes(_p(b), _p(c), len(b), width, cost, None, 0, C.byref(need))
    out = np.empty(need.value, np.int32)
    R.ref_emit_boxes(_p(b), _p(c), len(b), width, cost, _p(out), out.size, C.byref(need))
    return out
