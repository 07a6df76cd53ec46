"""ctypes bindings of libimba_hip.so (the C ABI declared in include/imba_hip.h).

The library is loaded from this package directory.  There is no fallback: if the library is
missing, importing this module raises ImportError telling how to build it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libimba_hip.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        "libimba_hip.so is not built (expected at %s); run `python -c 'import __graft_entry__ as g; g.build()'` "
        "or `python -m imbatracer_amd.build`" % LIB_PATH)

lib = C.CDLL(LIB_PATH)

# ---------------------------------------------------------------------------------------------
# numpy views of the ABI records
RAY_DTYPE = np.dtype([("org", "<f4", (4,)), ("dir", "<f4", (4,))])
HIT_DTYPE = np.dtype([("tri_id", "<i4"), ("inst_id", "<i4"), ("tmax", "<f4"), ("u", "<f4")])
BBOX_FIELDS = [("lo_x", "<f4"), ("hi_x", "<f4"), ("lo_y", "<f4"), ("hi_y", "<f4"), ("lo_z", "<f4"), ("hi_z", "<f4")]
NODE_DTYPE = np.dtype([("left_bb", np.dtype(BBOX_FIELDS)), ("right_bb", np.dtype(BBOX_FIELDS)),
                       ("left", "<i4"), ("right", "<i4"), ("pad", "<i4", (2,))])
INSTANCE_DTYPE = np.dtype([("transf", "<f4", (3, 4)), ("id", "<i4"), ("next", "<i4"), ("pad", "<i4", (2,))])
MASK_DTYPE = np.dtype([("width", "<i4"), ("height", "<i4"), ("offset", "<i4")])
NODE4_DTYPE = np.dtype([("min_x", "<f4", (4,)), ("max_x", "<f4", (4,)), ("min_y", "<f4", (4,)), ("max_y", "<f4", (4,)),
                        ("min_z", "<f4", (4,)), ("max_z", "<f4", (4,)), ("children", "<i4", (4,)), ("pad", "<i4", (4,))])
assert RAY_DTYPE.itemsize == 32 and HIT_DTYPE.itemsize == 16
assert NODE_DTYPE.itemsize == 64 and INSTANCE_DTYPE.itemsize == 64 and MASK_DTYPE.itemsize == 12
assert NODE4_DTYPE.itemsize == 128

DUMMY_CHILD = 0x76543210
LEAF_END_BITS = 0x80000000
BUILDER_SBVH = 0
BUILDER_BINNED_SAH = 1
COST_GPU_MESH, COST_GPU_TOP, COST_CPU = 0, 1, 2
CACHE_BVH2, CACHE_BVH4 = 1, 2

IMBA_OK = 0
ERRORS = {1: "invalid argument", 2: "HIP error", 3: "out of memory", 4: "bad BVH", 5: "capacity",
          6: "no device", 7: "kernel error"}


class ImbaError(RuntimeError):
    def __init__(self, code, where):
        self.code = code
        msg = lib.imba_last_error().decode(errors="replace")
        super().__init__("%s failed (%s): %s" % (where, ERRORS.get(code, code), msg))


def check(code, where=""):
    if code != IMBA_OK:
        raise ImbaError(code, where)


# ---------------------------------------------------------------------------------------------
class SceneDesc(C.Structure):
    _fields_ = [("root", C.c_int),
                ("nodes", C.c_void_p), ("node_count", C.c_int),
                ("instances", C.c_void_p), ("instance_count", C.c_int),
                ("tris", C.c_void_p), ("tri_vec4_count", C.c_int),
                ("indices", C.c_void_p), ("index_count", C.c_int),
                ("texcoords", C.c_void_p), ("texcoord_count", C.c_int),
                ("masks", C.c_void_p), ("mask_count", C.c_int),
                ("mask_buffer", C.c_void_p), ("mask_buffer_size", C.c_int)]


class CpuSceneDesc(C.Structure):
    """ImbaCpuSceneDesc: the fields of SceneDesc, with 4-wide nodes."""
    _fields_ = SceneDesc._fields_


class MeshDesc(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("vertex_count", C.c_int),
                ("indices", C.c_void_p), ("tri_count", C.c_int),
                ("texcoords", C.c_void_p)]


class TraceConfig(C.Structure):
    _fields_ = [("block_size", C.c_int), ("stack_lds", C.c_int), ("persistent", C.c_int), ("blocks_per_cu", C.c_int),
                ("bvh_width", C.c_int), ("quantized", C.c_int), ("treelet", C.c_int)]


class TraceItem(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("rays", C.c_void_p), ("hits", C.c_void_p), ("count", C.c_int),
                ("count_dev", C.c_void_p), ("any_hit", C.c_int)]


class PtShading(C.Structure):
    """ImbaPtShading: a scene's shading data for the path tracer (device pointers)."""
    _fields_ = [("tri_normals", C.c_void_p), ("tri_materials", C.c_void_p), ("tri_count", C.c_int),
                ("albedo", C.c_void_p), ("emission", C.c_void_p), ("emitter_area", C.c_void_p),
                ("material_count", C.c_int), ("lights", C.c_void_p), ("light_count", C.c_int),
                ("max_path_len", C.c_int), ("rr_start", C.c_int)]


class ShadeItem(C.Structure):
    """ImbaShadeItem: one queue of a batched shading step + compaction."""
    _fields_ = [("rays", C.c_void_p), ("hits", C.c_void_p), ("states", C.c_void_p), ("capacity", C.c_int),
                ("count_dev", C.c_void_p), ("rays_out", C.c_void_p), ("states_out", C.c_void_p),
                ("next_count_dev", C.c_void_p), ("shadow_rays", C.c_void_p), ("shadow_states", C.c_void_p),
                ("shadow_hits", C.c_void_p), ("image", C.c_void_p), ("shading", PtShading)]


class HashGridInfo(C.Structure):
    _fields_ = [("photon_count", C.c_int), ("table_size", C.c_int),
                ("bbox_min", C.c_float * 3), ("bbox_max", C.c_float * 3),
                ("radius", C.c_float), ("radius_sqr", C.c_float), ("cell_size", C.c_float), ("inv_cell_size", C.c_float),
                ("cell_ends", C.c_void_p), ("order", C.c_void_p), ("positions", C.c_void_p)]


vp, ip, i32, f32, sz = C.c_void_p, C.POINTER(C.c_int), C.c_int, C.c_float, C.c_size_t


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_sig("imba_last_error", C.c_char_p)
_sig("imba_version", C.c_char_p)
_sig("imba_device_count", i32, ip)
_sig("imba_set_device", i32, i32)
_sig("imba_get_device", i32, ip)
_sig("imba_device_info", i32, ip, ip, ip, C.c_char_p, i32)
_sig("imba_malloc", i32, C.POINTER(vp), sz)
_sig("imba_free", i32, vp)
_sig("imba_host_alloc", i32, C.POINTER(vp), sz)
_sig("imba_host_free", i32, vp)
_sig("imba_memcpy_h2d", i32, vp, vp, sz, vp)
_sig("imba_memcpy_d2h", i32, vp, vp, sz, vp)
_sig("imba_memcpy_d2d", i32, vp, vp, sz, vp)
_sig("imba_memset", i32, vp, i32, sz, vp)
_sig("imba_stream_create", i32, C.POINTER(vp))
_sig("imba_stream_destroy", i32, vp)
_sig("imba_stream_sync", i32, vp)
_sig("imba_event_create", i32, C.POINTER(vp))
_sig("imba_event_destroy", i32, vp)
_sig("imba_event_record", i32, vp, vp)
_sig("imba_event_elapsed_ms", i32, C.POINTER(f32), vp, vp)
_sig("imba_event_sync", i32, vp)
_sig("imba_stream_wait_event", i32, vp, vp)
_sig("imba_build_create", i32, C.POINTER(vp))
_sig("imba_build_release", i32, vp)
_sig("imba_build_add_mesh", i32, vp, C.POINTER(MeshDesc), ip)
_sig("imba_build_add_instance", i32, vp, i32, vp, ip)
_sig("imba_build_set_mask", i32, vp, i32, i32, i32, vp)
_sig("imba_build_run", i32, vp, i32, i32)
_sig("imba_build_get_desc", i32, vp, C.POINTER(SceneDesc))
_sig("imba_build_get_tri_attrs", i32, vp, C.POINTER(vp), C.POINTER(vp), ip)
_sig("imba_build_cpu_layout", i32, vp, i32, i32)
_sig("imba_build_get_cpu_desc", i32, vp, C.POINTER(CpuSceneDesc))
_sig("imba_bvh_store", i32, C.c_char_p, i32, vp, i32, vp, i32, i32, i32, i32)
_sig("imba_bvh_load", i32, C.c_char_p, i32, i32, i32, i32, vp, i32, vp, i32, ip, ip)
_sig("imba_bvh_has_block", i32, C.c_char_p, i32)
_sig("imba_build_set_cache", i32, vp, i32, C.c_char_p)
_sig("imba_build_cache_stats", i32, vp, ip, ip)
_sig("imba_bvh_emit_mesh", i32, C.POINTER(MeshDesc), i32, i32, i32, vp, sz, C.POINTER(sz))
_sig("imba_bvh_emit_boxes", i32, vp, vp, i32, i32, i32, vp, sz, C.POINTER(sz))
_sig("imba_scene_validate", i32, C.POINTER(SceneDesc), ip, ip)
_sig("imba_scene_upload", i32, C.POINTER(SceneDesc), C.POINTER(vp))
_sig("imba_trace_desc", i32, C.POINTER(SceneDesc), vp, vp, i32, i32, vp)
_sig("imba_ffi_error_count", i32, ip)
_sig("imba_scene_release", i32, vp)
_sig("imba_scene_device_desc", i32, vp, C.POINTER(SceneDesc))
_sig("imba_scene_footprint", i32, vp, C.POINTER(sz))
_sig("imba_intersect", i32, vp, vp, vp, i32, C.POINTER(TraceConfig), vp)
_sig("imba_occluded", i32, vp, vp, vp, i32, C.POINTER(TraceConfig), vp)
_sig("imba_intersect_dev", i32, vp, vp, vp, i32, vp, C.POINTER(TraceConfig), vp)
_sig("imba_occluded_dev", i32, vp, vp, vp, i32, vp, C.POINTER(TraceConfig), vp)
_sig("imba_intersect_stats", i32, vp, vp, vp, vp, i32, vp)
_sig("imba_occluded_stats", i32, vp, vp, vp, vp, i32, vp)
_sig("imba_trace_error_count", i32, vp, ip)
_sig("imba_trace_stats", i32, vp, vp, vp, vp, i32, i32, C.POINTER(TraceConfig), vp)
_sig("imba_trace_stack_depths", i32, vp, vp, vp, i32, i32, C.POINTER(TraceConfig), vp, vp)
_sig("imba_wide_create", i32, C.POINTER(SceneDesc), C.POINTER(vp))
_sig("imba_wide_get", i32, vp, C.POINTER(vp), ip, C.POINTER(vp), ip, ip, ip)
_sig("imba_wide_get_quantized", i32, vp, C.POINTER(vp), ip)
_sig("imba_wide_release", i32, vp)
QNODE4_DTYPE = np.dtype([("origin", "<f4", (3,)), ("exp", "i1", (3,)), ("count", "u1"), ("children", "<i4", (4,)),
                         ("lo_x", "u1", (4,)), ("hi_x", "u1", (4,)), ("lo_y", "u1", (4,)), ("hi_y", "u1", (4,)),
                         ("lo_z", "u1", (4,)), ("hi_z", "u1", (4,)), ("pad", "<u4", (2,))])
assert QNODE4_DTYPE.itemsize == 64
_sig("imba_trace_batch_create", i32, C.POINTER(TraceItem), i32, C.POINTER(vp))
_sig("imba_trace_batch_launch", i32, vp, C.POINTER(TraceConfig), vp)
_sig("imba_trace_batch_update", i32, vp, C.POINTER(TraceItem), i32, vp)
_sig("imba_trace_batch_release", i32, vp)
_sig("imba_autotune", i32, vp, vp, vp, i32, i32, i32, C.POINTER(TraceConfig), C.POINTER(f32), vp)
_sig("imba_read_bandwidth", i32, vp, C.c_size_t, i32, vp, C.c_size_t, vp, C.POINTER(f32))
_sig("imba_graph_begin", i32, vp)
_sig("imba_graph_end", i32, vp, C.POINTER(vp))
_sig("imba_graph_launch", i32, vp, vp)
_sig("imba_graph_release", i32, vp)
_sig("imba_compact_hits", i32, vp, vp, vp, i32, i32, vp, vp, vp, ip, vp)
_sig("imba_compact_rays", i32, vp, vp, i32, i32, vp, vp, ip, vp)
_sig("imba_compact_hits_noalloc", i32, vp, vp, vp, i32, i32, vp, vp, vp, ip, vp)
_sig("imba_compact_workspace_size", sz, i32)
_sig("imba_compact_rays_dev", i32, vp, vp, i32, i32, vp, vp, vp, vp, vp, vp)
_sig("imba_sort_by_material", i32, vp, i32, vp, vp, i32, i32, vp, vp, vp)
_sig("imba_hashgrid_create", i32, C.POINTER(vp))
_sig("imba_hashgrid_release", i32, vp)
_sig("imba_hashgrid_build", i32, vp, vp, i32, i32, i32, f32, vp)
_sig("imba_hashgrid_query", i32, vp, vp, i32, i32, i32, i32, vp, vp, vp, vp)
_sig("imba_hashgrid_info", i32, vp, C.POINTER(HashGridInfo))
_sig("imba_hashgrid_reset", i32, vp)
_sig("imba_hashgrid_gather", i32, vp, vp, i32, vp, vp)
_sig("imba_gen_random_triangles", i32, C.c_uint64, i32, vp, vp)
_sig("imba_gen_icosphere_counts", i32, i32, ip, ip)
_sig("imba_gen_icosphere", i32, i32, f32, f32, f32, f32, C.c_uint32, vp, vp)
_sig("imba_gen_camera_rays", i32, vp, vp, vp, f32, i32, i32, vp)
_sig("imba_gen_camera_matrix", i32, vp, vp, vp, f32, i32, i32, vp)
_sig("imba_gen_camera_frame", i32, vp, vp, vp, f32, i32, i32, vp, vp, C.POINTER(f32))
_sig("imba_vcm_light", i32, vp, C.c_uint32, i32, i32, vp, vp, vp)
_sig("imba_vcm_camera", i32, vp, C.c_uint32, i32, i32, vp, vp, vp)
_sig("imba_vcm_light_bounce", i32, vp, vp, vp, vp, i32, vp, vp, vp, vp, vp, vp, vp)
_sig("imba_vcm_camera_bounce", i32, vp, vp, vp, vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp)
_sig("imba_vcm_merge_records", i32, vp, vp, vp, vp, vp)
_sig("imba_vcm_merge", i32, vp, vp, vp, vp, i32, vp, i32, vp, vp)
_sig("imba_pt_camera", i32, vp, vp, i32, i32, i32, C.c_uint32, i32, i32, vp, vp, vp)
_sig("imba_pt_shade", i32, C.POINTER(PtShading), vp, vp, vp, i32, vp, vp, vp, vp, vp, vp)
_sig("imba_pt_shadow", i32, vp, vp, i32, vp, vp, vp)
_sig("imba_shade_batch_create", i32, vp, i32, C.POINTER(vp))
_sig("imba_shade_batch_update", i32, vp, vp, i32, vp)
_sig("imba_shade_batch_launch", i32, vp, vp)
_sig("imba_shade_batch_shadow", i32, vp, vp)
_sig("imba_shade_batch_release", i32, vp)
_sig("imba_ppm_light", i32, vp, vp, C.c_uint32, i32, i32, vp, vp, vp)
_sig("imba_ppm_light_bounce", i32, vp, vp, vp, i32, vp, vp, vp, i32, vp, vp, i32, i32, i32, vp, vp, vp, vp)
_sig("imba_ppm_camera", i32, vp, vp, i32, i32, i32, C.c_uint32, i32, i32, vp, vp, vp, vp)
_sig("imba_ppm_camera_bounce", i32, vp, vp, vp, i32, vp, vp, vp, i32, vp, vp, vp, i32, i32, i32, i32, vp, vp, vp, vp)
_sig("imba_ppm_merge", i32, vp, i32, i32, vp, vp, vp, vp, vp, f32, i32, vp, vp)
_sig("imba_ppm_merge_records", i32, vp, vp, vp, vp, vp)
_sig("imba_ppm_merge_grid", i32, vp, vp, vp, i32, i32, f32, i32, vp, vp)
_sig("intersect_gpu_masked_instanced", None, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, i32)
_sig("occluded_gpu_masked_instanced", None, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, i32)


def ptr(a: np.ndarray):
    """Address of a C-contiguous numpy array."""
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)


def device_count() -> int:
    n = C.c_int(0)
    check(lib.imba_device_count(C.byref(n)), "imba_device_count")
    return n.value


# ---------------------------------------------------------------------------------------------
# Device memory
class DeviceBuffer:
    """A hipMalloc'd buffer with numpy upload/download."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        check(lib.imba_malloc(C.byref(self.ptr), max(self.nbytes, 1)), "imba_malloc")

    @classmethod
    def from_array(cls, a: np.ndarray, stream=None):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a, stream)
        return b

    def upload(self, a: np.ndarray, stream=None, offset: int = 0):
        a = np.ascontiguousarray(a)
        assert offset + a.nbytes <= self.nbytes
        check(lib.imba_memcpy_h2d(C.c_void_p(self.ptr.value + offset), ptr(a), a.nbytes, _s(stream)), "h2d")

    def download(self, dtype, count: int, stream=None, offset: int = 0) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        assert offset + out.nbytes <= self.nbytes
        check(lib.imba_memcpy_d2h(ptr(out), C.c_void_p(self.ptr.value + offset), out.nbytes, _s(stream)), "d2h")
        return out

    def at(self, offset: int):
        return C.c_void_p(self.ptr.value + offset)

    def free(self):
        if self.ptr and self.ptr.value:
            check(lib.imba_free(self.ptr), "imba_free")
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PinnedArray:
    """A numpy array over pinned host memory (hipHostMalloc): asynchronous copies into it do not
    block the host."""

    def __init__(self, count: int, dtype):
        self.dtype = np.dtype(dtype)
        self.ptr = C.c_void_p()
        nbytes = max(1, count * self.dtype.itemsize)
        check(lib.imba_host_alloc(C.byref(self.ptr), nbytes), "imba_host_alloc")
        buf = (C.c_char * nbytes).from_address(self.ptr.value)
        self.array = np.frombuffer(buf, dtype=self.dtype, count=count)

    def __del__(self):
        try:
            if self.ptr and self.ptr.value:
                lib.imba_host_free(self.ptr)
                self.ptr = C.c_void_p()
        except Exception:
            pass


class Stream:
    def __init__(self):
        self.handle = C.c_void_p()
        check(lib.imba_stream_create(C.byref(self.handle)), "imba_stream_create")

    def sync(self):
        check(lib.imba_stream_sync(self.handle), "imba_stream_sync")

    def wait(self, event: "Event"):
        """Work enqueued on this stream from now on waits for the work recorded before event."""
        check(lib.imba_stream_wait_event(self.handle, event.handle), "imba_stream_wait_event")

    def __del__(self):
        try:
            if self.handle:
                lib.imba_stream_destroy(self.handle)
        except Exception:
            pass


class Event:
    def __init__(self):
        self.handle = C.c_void_p()
        check(lib.imba_event_create(C.byref(self.handle)), "imba_event_create")

    def record(self, stream=None):
        check(lib.imba_event_record(self.handle, stream.handle if isinstance(stream, Stream) else stream), "record")

    def sync(self):
        check(lib.imba_event_sync(self.handle), "imba_event_sync")

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        check(lib.imba_event_elapsed_ms(C.byref(ms), self.handle, end.handle), "elapsed")
        return ms.value

    def __del__(self):
        try:
            lib.imba_event_destroy(self.handle)
        except Exception:
            pass


def _s(stream):
    return stream.handle if isinstance(stream, Stream) else stream
