"""Python binding of librtx.so (include/rtx.h, include/rtx_scene.h).

Plumbing for tests and benchmarks: numpy arrays in host memory, device
buffers allocated through the library itself (rtx_malloc), ctypes calls.
The library must be built (``python treelet-prefetching-for-rt_amd/rtx/_build.py``
or ``__graft_entry__.build()``); importing this module without it raises.

Names follow the reference path they replace:
  Context.trace(..., mode=PLAIN)    VulkanRayTracing::traceRay              (vulkan_ray_tracing.cc:2309)
  Context.trace(..., mode=TREELET)  VulkanRayTracing::traceRayWithTreelets  (vulkan_ray_tracing.cc:1522)
  Context.build_treelets(budget)    VulkanRayTracing::createTreelets        (vulkan_ray_tracing.cc:823)
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librtx.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
_lib = ctypes.CDLL(LIB_PATH)

# ---- constants (include/rtx.h) ----------------------------------------------
OK = 0
STATUS_NAMES = {
    0: "RTX_OK", 1: "RTX_ERR_INVALID_ARGUMENT", 2: "RTX_ERR_BAD_NODE_TYPE", 3: "RTX_ERR_OFFSET_OUT_OF_ARENA",
    4: "RTX_ERR_NODE_LARGER_THAN_BUDGET", 5: "RTX_ERR_TRACE_OVERFLOW", 6: "RTX_ERR_STACK_OVERFLOW",
    7: "RTX_ERR_HIP", 8: "RTX_ERR_NO_DEVICE", 9: "RTX_ERR_NO_SCENE", 10: "RTX_ERR_NO_PARTITION",
    11: "RTX_ERR_UNSUPPORTED", 12: "RTX_ERR_OUT_OF_MEMORY", 13: "RTX_ERR_TIMEOUT", 14: "RTX_ERR_STALE",
    15: "RTX_ERR_CYCLE",
}
MODE_PLAIN, MODE_TREELET = 0, 1
PREFETCH_NONE, PREFETCH_ALWAYS, PREFETCH_POPULARITY, PREFETCH_PARTIAL, PREFETCH_LEARNED = 0, 1, 2, 3, 4
PREFETCH_LATTER = 5
RAY_FLAG_OPAQUE, RAY_FLAG_TERMINATE_ON_FIRST_HIT, RAY_FLAG_SKIP_CLOSEST_HIT_SHADER = 0x1, 0x4, 0x8
TX_BVH_STRUCTURE, TX_BVH_INTERNAL_NODE, TX_BVH_INSTANCE_LEAF = 0, 1, 2
TX_BVH_PRIMITIVE_LEAF_DESCRIPTOR, TX_BVH_QUAD_LEAF, TX_BVH_QUAD_LEAF_HIT, TX_BVH_PROCEDURAL_LEAF = 3, 4, 5, 6
NO_TREELET = 0xFFFFFFFF
WORKLOAD_SPHERE_AND_QUAD, WORKLOAD_INSTANCED_SPHERES = 0, 1
BUILD_DEFAULT, BUILD_SAH_OPTIMAL = 0, 1

RAY_DTYPE = np.dtype([("org", "<f4", 3), ("tmin", "<f4"), ("dir", "<f4", 3), ("tmax", "<f4"),
                      ("flags", "<u4"), ("cull_mask", "<u4"), ("sbt_record_offset", "<u4"),
                      ("sbt_record_stride", "<u4"), ("miss_index", "<u4"), ("reserved", "<u4", 3)])
HIT_DTYPE = np.dtype([("hit", "<u4"), ("primitive_index", "<u4"), ("geometry_index", "<u4"),
                      ("instance_index", "<u4"), ("t", "<f4"), ("p", "<f4", 3), ("bary", "<f4", 3),
                      ("n_records", "<u4"), ("n_nodes", "<u4"), ("fetch_bytes", "<u4"), ("n_any_hits", "<u4"),
                      ("max_level", "<u4"), ("instance_leaf", "<u8"), ("accessed_bytes", "<u4"),
                      ("reserved", "<u4")])
RECORD_DTYPE = np.dtype([("offset", "<u8"), ("treelet", "<u4"), ("size", "<u2"), ("type", "u1"),
                         ("level", "u1")])
assert RAY_DTYPE.itemsize == 64 and HIT_DTYPE.itemsize == 80 and RECORD_DTYPE.itemsize == 16
NO_NODE = 0xFFFFFFFFFFFFFFFF


class RtxError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        super().__init__(f"{what}: {STATUS_NAMES.get(status, status)}")


class _TraceOptions(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_uint32), ("prefetch", ctypes.c_uint32), ("threshold", ctypes.c_float),
                ("emit_records", ctypes.c_uint32), ("waves_per_cu", ctypes.c_uint32),
                ("lds_stack", ctypes.c_uint32), ("stage_bytes", ctypes.c_uint32), ("refill", ctypes.c_uint32),
                ("slot_bytes", ctypes.c_uint32), ("min_score", ctypes.c_float), ("pred_fraction", ctypes.c_float),
                ("lanes_per_ray", ctypes.c_uint32), ("claim_start", ctypes.c_uint32),
                ("record_layout", ctypes.c_uint32), ("record_slab", ctypes.c_uint32),
                ("claim_order", ctypes.c_void_p)]


class _TraceStats(ctypes.Structure):
    _fields_ = [("access_type", ctypes.c_uint64 * 7), ("num_hits", ctypes.c_uint64), ("num_any_hits", ctypes.c_uint64),
                ("n_rays", ctypes.c_uint64), ("tot_nodes", ctypes.c_uint64), ("max_nodes_per_ray", ctypes.c_uint32),
                ("max_tree_depth", ctypes.c_uint32), ("fetch_bytes", ctypes.c_uint64),
                ("accessed_bytes", ctypes.c_uint64), ("from_records", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class _CoherenceConfig(ctypes.Structure):
    _fields_ = [("hash", ctypes.c_char), ("world", ctypes.c_uint8), ("reserved", ctypes.c_uint16),
                ("francois_bits", ctypes.c_uint32), ("grid_bits", ctypes.c_uint32), ("sphere_bits", ctypes.c_uint32),
                ("two_point_ratio", ctypes.c_float), ("packet", ctypes.c_uint32),
                ("world_min", ctypes.c_float * 3), ("world_max", ctypes.c_float * 3)]


COHERENCE_WORLD_REFERENCE, COHERENCE_WORLD_TLAS, COHERENCE_WORLD_GIVEN = 0, 1, 2


class _LayoutSummary(ctypes.Structure):
    _fields_ = [("stride", ctypes.c_uint32), ("block_bytes", ctypes.c_uint32), ("layout_bytes", ctypes.c_uint64),
                ("mapped_nodes", ctypes.c_uint64), ("metadata_entries", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


class _PrefetchStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("node_reads", "lds_hits", "staged_treelets", "staged_bytes",
                                               "wasted_bytes", "predictions", "prediction_hits", "fallbacks",
                                               "timely_bytes", "too_early_bytes", "too_late_bytes", "late_bytes")]


class _TreeletSummary(ctypes.Structure):
    _fields_ = [("n_treelets", ctypes.c_uint32), ("budget_bytes", ctypes.c_uint32),
                ("n_list_entries", ctypes.c_uint64), ("total_bytes", ctypes.c_uint64),
                ("max_treelet_bytes", ctypes.c_uint32), ("max_treelet_nodes", ctypes.c_uint32)]


class _BatchJob(ctypes.Structure):
    _fields_ = [("scene", ctypes.c_uint32), ("n_rays", ctypes.c_uint32), ("rays_dev", ctypes.c_void_p),
                ("hits_dev", ctypes.c_void_p)]


class _WorkloadDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("n_instances", ctypes.c_uint32), ("sphere_slices", ctypes.c_uint32),
                ("sphere_stacks", ctypes.c_uint32), ("shared_blas", ctypes.c_uint32), ("width", ctypes.c_uint32),
                ("height", ctypes.c_uint32), ("seed", ctypes.c_uint64), ("box_size", ctypes.c_float),
                ("scale_min", ctypes.c_float), ("scale_max", ctypes.c_float), ("displacement", ctypes.c_float),
                ("ray_flags", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


_vp = ctypes.c_void_p
_u32, _u64, _i32, _f32 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_float
_P = ctypes.POINTER


def _sig(name, restype, *args):
    f = getattr(_lib, name)
    f.restype = restype
    f.argtypes = list(args)
    return f


_sig("rtx_status_string", ctypes.c_char_p, _i32)
_sig("rtx_version", ctypes.c_char_p)
_sig("rtx_device_count", _i32, _P(_i32))
_sig("rtx_create", _i32, _i32, _P(_vp))
_sig("rtx_destroy", _i32, _vp)
_sig("rtx_synchronize", _i32, _vp)
_sig("rtx_malloc", _i32, _vp, ctypes.c_size_t, _P(_vp))
_sig("rtx_free", _i32, _vp, _vp)
_sig("rtx_memcpy_h2d", _i32, _vp, _vp, _vp, ctypes.c_size_t)
_sig("rtx_memcpy_d2h", _i32, _vp, _vp, _vp, ctypes.c_size_t)
_sig("rtx_memcpy_d2d", _i32, _vp, _vp, _vp, ctypes.c_size_t)
_sig("rtx_memset", _i32, _vp, _vp, _i32, ctypes.c_size_t)
_sig("rtx_timer_start", _i32, _vp)
_sig("rtx_timer_stop", _i32, _vp, _P(_f32))
_sig("rtx_device_info", _i32, _vp, _P(_i32), _P(_i32), ctypes.c_char_p, ctypes.c_size_t)
_sig("rtx_set_scene", _i32, _vp, _vp, ctypes.c_size_t, _u64)
_sig("rtx_set_scene_device", _i32, _vp, _vp, ctypes.c_size_t, _u64)
_sig("rtx_validate_scene", _i32, _vp, _P(_u64))
_sig("rtx_build_treelets", _i32, _vp, _u32, _P(_TreeletSummary))
_sig("rtx_get_treelets", _i32, _vp, _vp, _vp, _vp, _vp)
_sig("rtx_set_treelets", _i32, _vp, _u32, _u32, _vp, _vp, _vp, _vp)
_sig("rtx_trace", _i32, _vp, _vp, _u32, _P(_TraceOptions), _vp, _vp, _u64, _vp)
_sig("rtx_trace_static", _i32, _vp, _vp, _u32, _P(_TraceOptions), _vp)
_sig("rtx_trace_async", _i32, _vp, _vp, _u32, _P(_TraceOptions), _vp)
_sig("rtx_get_stream", _i32, _vp, _P(_vp))
_sig("rtx_build_treelet_layout", _i32, _vp, _u32, _P(_LayoutSummary))
_sig("rtx_coherence_order", _i32, _vp, _vp, _u32, _P(_CoherenceConfig), _vp, _vp)
_sig("rtx_trace_stats_of", _i32, _vp, _vp, _u32, _vp, _u64, _P(_TraceStats))
_sig("rtx_get_treelet_layout", _i32, _vp, _vp, _vp)
_sig("rtx_adapter_create", _i32, _i32, _vp, _u64, _u64, _u32, _P(_vp))
_sig("rtx_adapter_destroy", _i32, _vp)
_sig("rtx_adapter_trace_ray", _i32, _vp, _i32, _u32, _u32, _u32, _u32, _u32, _P(_f32), _f32, _P(_f32), _f32, _vp, _vp,
     _u64, _P(_u64))
_sig("rtx_adapter_trace_lanes", _i32, _vp, _i32, _vp, _u32, _vp, _vp, _u64, _vp)
_sig("rtx_get_prefetch_stats", _i32, _vp, _P(_PrefetchStats))
_sig("rtx_last_kernel_ms", _i32, _vp, _P(_f32))
_sig("rtx_add_scene", _i32, _vp, _vp, ctypes.c_size_t, _u64, _P(_u32))
_sig("rtx_build_scene_treelets", _i32, _vp, _u32, _u32, _P(_TreeletSummary))
_sig("rtx_trace_batch", _i32, _vp, _P(_BatchJob), _u32, _P(_TraceOptions))
_sig("rtx_batch_graph_create", _i32, _vp, _P(_BatchJob), _u32, _P(_TraceOptions), _P(_vp))
_sig("rtx_batch_graph_launch", _i32, _vp, _u32, _P(_f32))
_sig("rtx_batch_graph_destroy", _i32, _vp)
_sig("rtx_autotune_batch", _i32, _vp, _P(_BatchJob), _u32, _P(_TraceOptions), _P(_f32))
_sig("rtx_autotune_batch_treelets", _i32, _vp, _P(_BatchJob), _u32, _vp, _u32, _P(_TraceOptions), _P(_u32), _P(_f32))
_sig("rtx_set_prefetch_table", _i32, _vp, _vp, _u64)
_sig("rtx_mlp_create", _i32, _vp, _u32, _u32, _u64, _P(_vp))
_sig("rtx_mlp_destroy", _i32, _vp)
_sig("rtx_mlp_param_count", _i32, _vp, _P(_u64))
_sig("rtx_mlp_get_params", _i32, _vp, _vp, _vp)
_sig("rtx_mlp_set_params", _i32, _vp, _vp)
_sig("rtx_mlp_loss_grad", _i32, _vp, _vp, _u32, _P(_f32))
_sig("rtx_mlp_train", _i32, _vp, _vp, _u32, _u32, _u32, _f32, _P(_f32))
_sig("rtx_mlp_predict", _i32, _vp, _vp, _u32, _vp, _vp)
_sig("rtx_mlp_gemm", _i32, _vp, _i32, _i32, _i32, _i32, _i32, _vp, _vp, _vp)
_sig("rtx_format_treelet_trace", _i32, _vp, _vp, _u32, _vp, _u32, _u64, _u32, _vp, ctypes.c_size_t,
     _P(ctypes.c_size_t))
_sig("rtx_as_dump_write", _i32, _vp, _u64, _u64, ctypes.c_char_p)
_sig("rtx_as_dump_info", _i32, ctypes.c_char_p, _P(_u64), _P(_u64))
_sig("rtx_as_dump_read", _i32, ctypes.c_char_p, _vp, _u64)
_sig("rtx_builder_create", _i32, _P(_vp))
_sig("rtx_builder_destroy", _i32, _vp)
_sig("rtx_builder_set_quality", _i32, _vp, _u32)
_sig("rtx_builder_add_blas", _i32, _vp, _vp, _u32, _u32, _P(_u32))
_sig("rtx_builder_add_instance", _i32, _vp, _u32, _vp, _u32, _u32)
_sig("rtx_builder_finish", _i32, _vp, _u32, _P(_u64), _P(_u64))
_sig("rtx_builder_finish_device", _i32, _vp, _vp, _P(_u64), _P(_u64))
_sig("rtx_workload_scene_device", _i32, _P(_WorkloadDesc), _vp, _P(_vp))
_sig("rtx_workload_builder", _i32, _P(_WorkloadDesc), _P(_vp))
_sig("rtx_builder_copy_arena", _i32, _vp, _vp, ctypes.c_size_t)
_sig("rtx_builder_arena_info", _i32, _vp, _P(_u64), _P(_u64))
_sig("rtx_builder_counts", _i32, _vp, _P(_u64), _P(_u64), _P(_u64), _P(_u64))
_sig("rtx_builder_blas_offset", _i32, _vp, _u32, _P(_u64))
_sig("rtx_mesh_displaced_sphere", _i32, _u32, _u32, _f32, _f32, _u64, _vp)
_sig("rtx_workload_scene", _i32, _P(_WorkloadDesc), _u32, _P(_vp))
_sig("rtx_workload_primary_rays", _i32, _P(_WorkloadDesc), _vp)
_sig("rtx_workload_secondary_rays", _i32, _vp, _vp, _vp, _u32, _u32, _u64, _vp, _vp, _vp, _P(_u32))

#: every symbol declared in include/rtx.h and include/rtx_scene.h
EXPORTED_SYMBOLS = [
    "rtx_status_string", "rtx_version", "rtx_device_count", "rtx_create", "rtx_destroy", "rtx_synchronize",
    "rtx_malloc", "rtx_free", "rtx_memcpy_h2d", "rtx_memcpy_d2h", "rtx_memcpy_d2d", "rtx_memset", "rtx_timer_start",
    "rtx_timer_stop", "rtx_device_info", "rtx_set_scene", "rtx_set_scene_device", "rtx_validate_scene",
    "rtx_build_treelets", "rtx_get_treelets", "rtx_set_treelets", "rtx_build_treelet_layout", "rtx_get_treelet_layout", "rtx_trace", "rtx_trace_static", "rtx_trace_async", "rtx_get_stream", "rtx_coherence_order", "rtx_trace_stats_of", "rtx_get_prefetch_stats",
    "rtx_last_kernel_ms", "rtx_add_scene", "rtx_build_scene_treelets", "rtx_trace_batch", "rtx_batch_graph_create",
    "rtx_batch_graph_launch", "rtx_batch_graph_destroy", "rtx_autotune_batch", "rtx_autotune_batch_treelets",
    "rtx_set_prefetch_table", "rtx_mlp_create", "rtx_mlp_destroy", "rtx_mlp_param_count", "rtx_mlp_get_params",
    "rtx_mlp_set_params", "rtx_mlp_loss_grad", "rtx_mlp_train", "rtx_mlp_predict", "rtx_mlp_gemm",
    "rtx_builder_create", "rtx_builder_destroy", "rtx_builder_set_quality", "rtx_builder_add_blas", "rtx_builder_add_instance",
    "rtx_builder_finish", "rtx_builder_finish_device", "rtx_workload_scene_device", "rtx_workload_builder", "rtx_builder_copy_arena", "rtx_builder_arena_info", "rtx_builder_counts", "rtx_builder_blas_offset",
    "rtx_mesh_displaced_sphere", "rtx_workload_scene", "rtx_workload_primary_rays", "rtx_workload_secondary_rays",
    "rtx_format_treelet_trace", "rtx_as_dump_write", "rtx_as_dump_info", "rtx_as_dump_read",
    "rtx_adapter_create", "rtx_adapter_destroy", "rtx_adapter_trace_ray", "rtx_adapter_trace_lanes",
]


def _check(st, what):
    if st != OK:
        raise RtxError(st, what)


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def version():
    return _lib.rtx_version().decode()


def device_count():
    n = ctypes.c_int(0)
    _check(_lib.rtx_device_count(ctypes.byref(n)), "rtx_device_count")
    return n.value


# ---- scenes ------------------------------------------------------------------
class Scene:
    """A finished acceleration-structure arena plus the builder that made it."""

    def __init__(self, handle):
        self._h = ctypes.c_void_p(handle)
        nbytes, tlas = ctypes.c_uint64(0), ctypes.c_uint64(0)
        tri, blas, inst, internal = (ctypes.c_uint64(0) for _ in range(4))
        _check(_lib.rtx_builder_counts(self._h, ctypes.byref(tri), ctypes.byref(blas), ctypes.byref(inst),
                                       ctypes.byref(internal)), "rtx_builder_counts")
        self.n_triangles, self.n_blas, self.n_instances, self.n_internal = tri.value, blas.value, inst.value, internal.value
        self._arena = None
        self.tlas_offset = 0

    @property
    def arena(self):
        return self._arena

    def _capture_finished(self):
        nb, tl = ctypes.c_uint64(0), ctypes.c_uint64(0)
        _check(_lib.rtx_builder_arena_info(self._h, ctypes.byref(nb), ctypes.byref(tl)), "rtx_builder_arena_info")
        buf = np.empty(nb.value, dtype=np.uint8)
        _check(_lib.rtx_builder_copy_arena(self._h, _ptr(buf), nb.value), "rtx_builder_copy_arena")
        self._arena = buf
        self.tlas_offset = tl.value

    def blas_offset(self, blas_id):
        o = ctypes.c_uint64(0)
        _check(_lib.rtx_builder_blas_offset(self._h, blas_id, ctypes.byref(o)), "rtx_builder_blas_offset")
        return o.value

    def secondary_rays(self, rays, hits, kind=0, seed=1, light=(0.0, 0.0, 0.0)):
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        hits = np.ascontiguousarray(hits, dtype=HIT_DTYPE)
        out = np.zeros(len(rays), dtype=RAY_DTYPE)
        src = np.zeros(len(rays), dtype=np.uint32)
        n_out = ctypes.c_uint32(0)
        lp = np.asarray(light, dtype=np.float32)
        _check(_lib.rtx_workload_secondary_rays(self._h, _ptr(rays), _ptr(hits), len(rays), kind, seed, _ptr(lp),
                                                _ptr(out), _ptr(src), ctypes.byref(n_out)),
               "rtx_workload_secondary_rays")
        return out[:n_out.value].copy(), src[:n_out.value].copy()

    def close(self):
        if self._h:
            _lib.rtx_builder_destroy(self._h)
            self._h = ctypes.c_void_p(None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SceneBuilder:
    """rtx_builder: add triangle BLASes and instances, then finish()."""

    def __init__(self):
        h = ctypes.c_void_p()
        _check(_lib.rtx_builder_create(ctypes.byref(h)), "rtx_builder_create")
        self._h = h

    def add_blas(self, tris, geometry_index=0):
        t = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
        bid = ctypes.c_uint32(0)
        _check(_lib.rtx_builder_add_blas(self._h, _ptr(t), len(t), geometry_index, ctypes.byref(bid)),
               "rtx_builder_add_blas")
        return bid.value

    def add_instance(self, blas_id, object_to_world=None, instance_id=0, hit_group=0):
        m = np.eye(3, 4, dtype=np.float32) if object_to_world is None else np.asarray(object_to_world, np.float32)
        m = np.ascontiguousarray(m.reshape(12))
        _check(_lib.rtx_builder_add_instance(self._h, blas_id, _ptr(m), instance_id, hit_group),
               "rtx_builder_add_instance")

    def finish(self, threads=0, ctx=None, quality=BUILD_DEFAULT):
        """Build on the host (SAH), or with `ctx` the BLASes on that context's device.
        quality: BUILD_DEFAULT or BUILD_SAH_OPTIMAL (rtx_builder_set_quality; on the
        device path it applies to the TLAS only)."""
        nb, tl = ctypes.c_uint64(0), ctypes.c_uint64(0)
        _check(_lib.rtx_builder_set_quality(self._h, quality), "rtx_builder_set_quality")
        if ctx is None:
            _check(_lib.rtx_builder_finish(self._h, threads, ctypes.byref(nb), ctypes.byref(tl)), "rtx_builder_finish")
        else:
            _check(_lib.rtx_builder_finish_device(self._h, ctx._h, ctypes.byref(nb), ctypes.byref(tl)),
                   "rtx_builder_finish_device")
        s = Scene(self._h.value)
        s._capture_finished()
        self._h = ctypes.c_void_p(None)
        return s


def displaced_sphere(slices, stacks, radius=1.0, disp=0.05, seed=1):
    out = np.empty((slices * stacks * 2, 9), dtype=np.float32)
    _check(_lib.rtx_mesh_displaced_sphere(slices, stacks, radius, disp, seed, _ptr(out)), "rtx_mesh_displaced_sphere")
    return out


def workload_desc(kind=WORKLOAD_SPHERE_AND_QUAD, n_instances=2, slices=128, stacks=64, shared_blas=0, width=128,
                  height=128, seed=1, box_size=100.0, scale_min=0.5, scale_max=2.0, displacement=0.05,
                  ray_flags=RAY_FLAG_OPAQUE):
    return _WorkloadDesc(kind, n_instances, slices, stacks, shared_blas, width, height, seed, box_size, scale_min,
                         scale_max, displacement, ray_flags, 0)


def workload_scene(desc, threads=0, ctx=None):
    """Build a workload scene on the host, or with `ctx` its BLASes on that device."""
    h = ctypes.c_void_p()
    if ctx is None:
        _check(_lib.rtx_workload_scene(ctypes.byref(desc), threads, ctypes.byref(h)), "rtx_workload_scene")
    else:
        _check(_lib.rtx_workload_scene_device(ctypes.byref(desc), ctx._h, ctypes.byref(h)), "rtx_workload_scene_device")
    s = Scene(h.value)
    s._capture_finished()
    return s


def workload_builder(desc):
    """The workload's meshes and instances in an unfinished SceneBuilder."""
    b = SceneBuilder.__new__(SceneBuilder)
    h = ctypes.c_void_p()
    _check(_lib.rtx_workload_builder(ctypes.byref(desc), ctypes.byref(h)), "rtx_workload_builder")
    b._h = h
    return b


def primary_rays(desc):
    rays = np.zeros(desc.width * desc.height, dtype=RAY_DTYPE)
    _check(_lib.rtx_workload_primary_rays(ctypes.byref(desc), _ptr(rays)), "rtx_workload_primary_rays")
    return rays


# ---- reference formats ------------------------------------------------------------
def format_treelet_trace(offsets, records, roots, base_address=0, first_ray_id=1):
    """Per-ray treelet trace lines of traceRayWithTreelets ("RayID,<id>,0x<root>,...")."""
    offsets = np.ascontiguousarray(offsets, np.uint64)
    records = np.ascontiguousarray(records, RECORD_DTYPE)
    roots = np.ascontiguousarray(roots, np.uint64)
    n = len(offsets) - 1
    length = ctypes.c_size_t(0)
    args = (_ptr(records) if len(records) else None, _ptr(offsets), n, _ptr(roots) if len(roots) else None,
            len(roots), base_address, first_ray_id)
    _check(_lib.rtx_format_treelet_trace(*args, None, 0, ctypes.byref(length)), "rtx_format_treelet_trace")
    buf = ctypes.create_string_buffer(length.value + 1)
    _check(_lib.rtx_format_treelet_trace(*args, buf, length.value + 1, ctypes.byref(length)),
           "rtx_format_treelet_trace")
    return buf.raw[: length.value].decode()


def as_dump_write(arena, tlas_offset, prefix):
    """Write the split-file AS dump (.asmain/.asback/.asfront/.asmetadata) of an arena."""
    a = np.ascontiguousarray(arena, np.uint8)
    _check(_lib.rtx_as_dump_write(_ptr(a), a.nbytes, tlas_offset, str(prefix).encode()), "rtx_as_dump_write")


def as_dump_read(prefix):
    """Rebuild an arena from a split-file AS dump; returns (arena, tlas_offset)."""
    nb, tl = ctypes.c_uint64(0), ctypes.c_uint64(0)
    _check(_lib.rtx_as_dump_info(str(prefix).encode(), ctypes.byref(nb), ctypes.byref(tl)), "rtx_as_dump_info")
    a = np.zeros(nb.value, np.uint8)
    _check(_lib.rtx_as_dump_read(str(prefix).encode(), _ptr(a), a.nbytes), "rtx_as_dump_read")
    return a, tl.value


# ---- device context -------------------------------------------------------------
class DeviceBuffer:
    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = max(int(nbytes), 1)
        p = ctypes.c_void_p()
        _check(_lib.rtx_malloc(ctx._h, self.nbytes, ctypes.byref(p)), "rtx_malloc")
        self.ptr = p

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        _check(_lib.rtx_memcpy_h2d(self.ctx._h, self.ptr, _ptr(a), a.nbytes), "rtx_memcpy_h2d")

    def copy_from(self, other, nbytes):
        """Device-to-device copy of nbytes from another buffer of the same context."""
        assert nbytes <= min(self.nbytes, other.nbytes)
        _check(_lib.rtx_memcpy_d2d(self.ctx._h, self.ptr, other.ptr, nbytes), "rtx_memcpy_d2d")

    def download(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            _check(_lib.rtx_memcpy_d2h(self.ctx._h, _ptr(out), self.ptr, out.nbytes), "rtx_memcpy_d2h")
        return out

    def free(self):
        if self.ptr:
            _lib.rtx_free(self.ctx._h, self.ptr)
            self.ptr = ctypes.c_void_p(None)

    def __del__(self):
        try:
            if self.ctx._h:
                self.free()
        except Exception:
            pass


MLP_SAMPLE_DTYPE = np.dtype([("cur", "<u4"), ("octant", "<u4"), ("depth", "<f4"), ("hist", "<u4", 4),
                             ("target", "<u4")])


class MLP:
    """The next-treelet network (rtx_mlp): training and inference on the device."""

    def __init__(self, ctx, n_treelets, n_classes, seed=1):
        self.ctx = ctx
        self.n_treelets, self.n_classes = int(n_treelets), int(n_classes)
        h = ctypes.c_void_p()
        _check(_lib.rtx_mlp_create(ctx._h, self.n_treelets, self.n_classes, seed, ctypes.byref(h)), "rtx_mlp_create")
        self._h = h
        c = ctypes.c_uint64(0)
        _check(_lib.rtx_mlp_param_count(h, ctypes.byref(c)), "rtx_mlp_param_count")
        self.n_params = c.value
        self._samples = None

    def upload(self, samples):
        """Copy samples (MLP_SAMPLE_DTYPE) to the device; returns a DeviceBuffer."""
        s = np.ascontiguousarray(samples, MLP_SAMPLE_DTYPE)
        return self.ctx.to_device(s)

    def _buf(self, samples):
        return samples if isinstance(samples, DeviceBuffer) else self.upload(samples)

    def params(self):
        p = np.empty(self.n_params, np.float32)
        _check(_lib.rtx_mlp_get_params(self._h, _ptr(p), None), "rtx_mlp_get_params")
        return p

    def grads(self):
        g = np.empty(self.n_params, np.float32)
        _check(_lib.rtx_mlp_get_params(self._h, None, _ptr(g)), "rtx_mlp_get_params")
        return g

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert len(p) == self.n_params
        _check(_lib.rtx_mlp_set_params(self._h, _ptr(p)), "rtx_mlp_set_params")

    def loss_grad(self, samples, n=None):
        b = self._buf(samples)
        n = len(samples) if n is None else n
        loss = ctypes.c_float(0)
        _check(_lib.rtx_mlp_loss_grad(self._h, b.ptr, n, ctypes.byref(loss)), "rtx_mlp_loss_grad")
        return loss.value

    def train(self, samples, n=None, epochs=1, batch=1024, lr=1e-3):
        b = self._buf(samples)
        n = len(samples) if n is None else n
        loss = ctypes.c_float(0)
        _check(_lib.rtx_mlp_train(self._h, b.ptr, n, epochs, batch, lr, ctypes.byref(loss)), "rtx_mlp_train")
        return loss.value

    def predict(self, samples, n=None):
        """Top-4 classes and their probabilities, arrays [n, 4]."""
        b = self._buf(samples)
        n = len(samples) if n is None else n
        tb = self.ctx.buffer(n * 16)
        pb = self.ctx.buffer(n * 16)
        _check(_lib.rtx_mlp_predict(self._h, b.ptr, n, tb.ptr, pb.ptr), "rtx_mlp_predict")
        return tb.download(np.uint32, 4 * n).reshape(n, 4), pb.download(np.float32, 4 * n).reshape(n, 4)

    def close(self):
        if self._h:
            _lib.rtx_mlp_destroy(self._h)
            self._h = ctypes.c_void_p(None)

    def __del__(self):
        try:
            if self.ctx._h:
                self.close()
        except Exception:
            pass


class BatchGraph:
    """A captured multi-scene batch launch (rtx_batch_graph)."""

    def __init__(self, handle):
        self._h = handle

    def launch(self, replays=1):
        ms = ctypes.c_float(0)
        st = _lib.rtx_batch_graph_launch(self._h, replays, ctypes.byref(ms))
        return st, ms.value

    def close(self):
        if self._h:
            _lib.rtx_batch_graph_destroy(self._h)
            self._h = ctypes.c_void_p(None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One rtx_ctx on a device (VulkanRayTracing's static state, per context)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        _check(_lib.rtx_create(device, ctypes.byref(h)), "rtx_create")
        self._h = h
        self.summary = None

    def close(self):
        if self._h:
            _lib.rtx_destroy(self._h)
            self._h = ctypes.c_void_p(None)

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_info(self):
        n, lds = ctypes.c_int(0), ctypes.c_int(0)
        name = ctypes.create_string_buffer(128)
        _check(_lib.rtx_device_info(self._h, ctypes.byref(n), ctypes.byref(lds), name, 128), "rtx_device_info")
        return {"n_cus": n.value, "lds_per_cu": lds.value, "name": name.value.decode()}

    def buffer(self, nbytes):
        return DeviceBuffer(self, nbytes)

    def to_device(self, arr):
        b = DeviceBuffer(self, np.asarray(arr).nbytes)
        b.upload(arr)
        return b

    def set_scene(self, arena, tlas_offset=0):
        a = np.ascontiguousarray(arena, dtype=np.uint8)
        _check(_lib.rtx_set_scene(self._h, _ptr(a), a.nbytes, tlas_offset), "rtx_set_scene")
        self.arena_bytes = a.nbytes
        self.summary = None

    def validate_scene(self):
        n = ctypes.c_uint64(0)
        _check(_lib.rtx_validate_scene(self._h, ctypes.byref(n)), "rtx_validate_scene")
        return n.value

    def build_treelets(self, budget_bytes):
        s = _TreeletSummary()
        _check(_lib.rtx_build_treelets(self._h, budget_bytes, ctypes.byref(s)), "rtx_build_treelets")
        self.summary = {f: getattr(s, f) for f, _ in _TreeletSummary._fields_}
        return self.summary

    def get_treelets(self):
        if self.summary is None:
            raise RtxError(10, "get_treelets")
        n = self.summary["n_treelets"]
        roots = np.empty(n, np.uint64)
        lo = np.empty(n + 1, np.uint64)
        ln = np.empty(max(self.summary["n_list_entries"], 1), np.uint64)
        owner = np.empty(self.arena_bytes // 64, np.uint32)
        _check(_lib.rtx_get_treelets(self._h, _ptr(roots), _ptr(lo), _ptr(ln), _ptr(owner)), "rtx_get_treelets")
        return {"roots": roots, "list_offsets": lo, "list_nodes": ln[: self.summary["n_list_entries"]],
                "owner": owner}

    def build_treelet_layout(self, stride=0):
        """remapBVHToTreeletLayout of the current partition; returns the summary as a dict."""
        s = _LayoutSummary()
        _check(_lib.rtx_build_treelet_layout(self._h, stride, ctypes.byref(s)), "rtx_build_treelet_layout")
        self.layout = {f: getattr(s, f) for f, _ in _LayoutSummary._fields_}
        return self.layout

    def get_treelet_layout(self):
        """(remap[arena/64], metadata[n_treelets, budget/64]) of the last built layout."""
        n, per = self.summary["n_treelets"], self.layout["metadata_entries"]
        remap = np.empty(self.arena_bytes // 64, np.uint64)
        meta = np.empty((n, per), np.int32)
        _check(_lib.rtx_get_treelet_layout(self._h, _ptr(remap), _ptr(meta)), "rtx_get_treelet_layout")
        return remap, meta

    def set_treelets(self, budget, roots, list_offsets, list_nodes, owner):
        r = np.ascontiguousarray(roots, np.uint64)
        lo = np.ascontiguousarray(list_offsets, np.uint64)
        ln = np.ascontiguousarray(list_nodes, np.uint64)
        ow = np.ascontiguousarray(owner, np.uint32)
        if len(ln) == 0:
            ln = np.zeros(1, np.uint64)
        _check(_lib.rtx_set_treelets(self._h, budget, len(r), _ptr(r), _ptr(lo), _ptr(ln), _ptr(ow)),
               "rtx_set_treelets")
        self.summary = {"n_treelets": len(r), "budget_bytes": budget, "n_list_entries": int(lo[-1])}

    def trace_device(self, rays_buf, n, hits_buf, mode=MODE_PLAIN, prefetch=PREFETCH_NONE, threshold=0.5,
                     records_buf=None, record_capacity=0, offsets_buf=None, waves_per_cu=0, lds_stack=0,
                     stage_bytes=0, refill=0, slot_bytes=0, min_score=0.0, pred_fraction=0.0, lanes_per_ray=0,
                     claim_start=0, record_slab=0, record_layout=0, claim_order=None):
        """rtx_trace on device buffers; returns the rtx_status (0 = ok).  claim_order: a
        DeviceBuffer of uint32 ray indices (rtx_coherence_order) or None."""
        o = _TraceOptions(mode, prefetch, threshold, 1 if offsets_buf is not None else 0, waves_per_cu, lds_stack,
                          stage_bytes, refill, slot_bytes, min_score, pred_fraction, lanes_per_ray, claim_start,
                          record_layout, record_slab, claim_order.ptr.value if claim_order is not None else None)
        return _lib.rtx_trace(self._h, rays_buf.ptr, n, ctypes.byref(o), hits_buf.ptr,
                              records_buf.ptr if records_buf is not None else None, record_capacity,
                              offsets_buf.ptr if offsets_buf is not None else None)

    def trace_static_device(self, rays_buf, n, hits_buf, mode=MODE_PLAIN, waves_per_cu=0, lds_stack=0, refill=0,
                            lanes_per_ray=0, claim_start=0):
        """rtx_trace_static (hits only, no allocation) on device buffers; returns the rtx_status."""
        o = _TraceOptions(mode, PREFETCH_NONE, 0.0, 0, waves_per_cu, lds_stack, 0, refill, 0, 0.0, 0.0, lanes_per_ray,
                          claim_start, 0, 0, None)
        return _lib.rtx_trace_static(self._h, rays_buf.ptr, n, ctypes.byref(o), hits_buf.ptr)

    def trace_stats(self, hits_buf, n, records_buf=None, n_records=0):
        """rtx_trace_stats_of on device buffers -> dict (access_type as a list)."""
        s = _TraceStats()
        _check(_lib.rtx_trace_stats_of(self._h, hits_buf.ptr, n, records_buf.ptr if records_buf is not None else None,
                                       n_records, ctypes.byref(s)), "rtx_trace_stats_of")
        out = {f: getattr(s, f) for f, _ in _TraceStats._fields_}
        out["access_type"] = list(s.access_type)
        return out

    def coherence_order_device(self, rays_buf, n, order_buf, hash="d", world=COHERENCE_WORLD_REFERENCE,
                               francois_bits=2, grid_bits=4, sphere_bits=4, two_point_ratio=0.5, packet=0,
                               world_min=(0.0, 0.0, 0.0), world_max=(1.0, 1.0, 1.0), hashes_buf=None):
        """rtx_coherence_order on device buffers; returns the rtx_status."""
        c = _CoherenceConfig(hash.encode(), world, 0, francois_bits, grid_bits, sphere_bits, two_point_ratio, packet,
                             (ctypes.c_float * 3)(*world_min), (ctypes.c_float * 3)(*world_max))
        return _lib.rtx_coherence_order(self._h, rays_buf.ptr, n, ctypes.byref(c), order_buf.ptr,
                                        hashes_buf.ptr if hashes_buf is not None else None)

    def coherence_order(self, rays, **cfg):
        """Host rays in; (order, hashes) out."""
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        n = len(rays)
        rb, ob, hb = self.to_device(rays), self.buffer(max(n, 1) * 4), self.buffer(max(n, 1) * 4)
        _check(self.coherence_order_device(rb, n, ob, hashes_buf=hb, **cfg), "rtx_coherence_order")
        return ob.download(np.uint32, n), hb.download(np.uint32, n)

    def trace(self, rays, mode=MODE_PLAIN, prefetch=PREFETCH_NONE, threshold=0.5, records=False,
              records_per_ray=48, **kw):
        """Host arrays in and out.  Returns hits, or (hits, offsets, records) when records=True.

        With records, one rtx_trace call fills a buffer of records_per_ray * n
        records; when the rays produce more, the call reports the total and is
        repeated with a buffer of that size."""
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        n = len(rays)
        rb = self.to_device(rays)
        hb = self.buffer(max(n, 1) * HIT_DTYPE.itemsize)
        if not records:
            _check(self.trace_device(rb, n, hb, mode, prefetch, threshold, **kw), "rtx_trace")
            return hb.download(HIT_DTYPE, n)
        ob = self.buffer((n + 1) * 8)
        cap = max(int(records_per_ray) * n, 1)
        recb = self.buffer(cap * RECORD_DTYPE.itemsize)
        st = self.trace_device(rb, n, hb, mode, prefetch, threshold, recb, cap, ob, **kw)
        if st == 5:
            total = int(ob.download(np.uint64, n + 1)[-1])
            recb.free()
            cap = max(total, 1)
            recb = self.buffer(cap * RECORD_DTYPE.itemsize)
            st = self.trace_device(rb, n, hb, mode, prefetch, threshold, recb, cap, ob, **kw)
        _check(st, "rtx_trace")
        offsets = ob.download(np.uint64, n + 1)
        total = int(offsets[-1])
        return hb.download(HIT_DTYPE, n), offsets, recb.download(RECORD_DTYPE, total)

    def set_prefetch_table(self, table):
        """Install the learned-prefetch table of the current partition: uint32 [n_treelets * 8, 4]."""
        t = np.ascontiguousarray(table, np.uint32).reshape(-1, 4)
        _check(_lib.rtx_set_prefetch_table(self._h, _ptr(t), len(t)), "rtx_set_prefetch_table")

    def mlp(self, n_treelets, n_classes, seed=1):
        return MLP(self, n_treelets, n_classes, seed)

    def gemm(self, a, b, trans_a=False, trans_b=False):
        """The MLP's BF16 MFMA GEMM on host arrays: op(a) @ op(b) in FP32."""
        a = np.ascontiguousarray(a, np.float32)
        b = np.ascontiguousarray(b, np.float32)
        m, k = (a.shape[1], a.shape[0]) if trans_a else a.shape
        k2, n = (b.shape[1], b.shape[0]) if trans_b else b.shape
        assert k == k2
        ab, bb, cb = self.to_device(a), self.to_device(b), self.buffer(m * n * 4)
        _check(_lib.rtx_mlp_gemm(self._h, m, n, k, int(trans_a), int(trans_b), ab.ptr, bb.ptr, cb.ptr), "rtx_mlp_gemm")
        return cb.download(np.float32, m * n).reshape(m, n)

    def prefetch_stats(self):
        s = _PrefetchStats()
        _check(_lib.rtx_get_prefetch_stats(self._h, ctypes.byref(s)), "rtx_get_prefetch_stats")
        return {f: getattr(s, f) for f, _ in _PrefetchStats._fields_}

    # ---- multi-scene batches ----------------------------------------------------
    def add_scene(self, arena, tlas_offset=0):
        a = np.ascontiguousarray(arena, dtype=np.uint8)
        sid = ctypes.c_uint32(0)
        _check(_lib.rtx_add_scene(self._h, _ptr(a), a.nbytes, tlas_offset, ctypes.byref(sid)), "rtx_add_scene")
        return sid.value

    def build_scene_treelets(self, scene_id, budget_bytes):
        s = _TreeletSummary()
        _check(_lib.rtx_build_scene_treelets(self._h, scene_id, budget_bytes, ctypes.byref(s)),
               "rtx_build_scene_treelets")
        return {f: getattr(s, f) for f, _ in _TreeletSummary._fields_}

    @staticmethod
    def _jobs(jobs):
        """jobs: list of (scene_id, rays DeviceBuffer, n_rays, hits DeviceBuffer)."""
        arr = (_BatchJob * len(jobs))()
        for k, (sid, rb, n, hb) in enumerate(jobs):
            arr[k] = _BatchJob(sid, n, rb.ptr.value if n else None, hb.ptr.value if n else None)
        return arr

    @staticmethod
    def _opts(mode=MODE_PLAIN, waves_per_cu=0, lds_stack=0, refill=0):
        return _TraceOptions(mode, PREFETCH_NONE, 0.0, 0, waves_per_cu, lds_stack, 0, refill, 0, 0.0, 0.0, 0, 0, 0, 0, None)

    def trace_batch(self, jobs, **opts):
        arr = self._jobs(jobs)
        o = self._opts(**opts)
        return _lib.rtx_trace_batch(self._h, arr, len(jobs), ctypes.byref(o))

    def batch_graph(self, jobs, **opts):
        arr = self._jobs(jobs)
        o = self._opts(**opts)
        g = ctypes.c_void_p()
        _check(_lib.rtx_batch_graph_create(self._h, arr, len(jobs), ctypes.byref(o), ctypes.byref(g)),
               "rtx_batch_graph_create")
        return BatchGraph(g)

    def autotune_batch(self, jobs, mode=MODE_PLAIN):
        arr = self._jobs(jobs)
        o = self._opts(mode=mode)
        ms = ctypes.c_float(0)
        _check(_lib.rtx_autotune_batch(self._h, arr, len(jobs), ctypes.byref(o), ctypes.byref(ms)),
               "rtx_autotune_batch")
        return {"mode": o.mode, "waves_per_cu": o.waves_per_cu, "lds_stack": o.lds_stack, "refill": o.refill}, ms.value

    def autotune_batch_treelets(self, jobs, budgets=(4096, 8192, 16384, 32768, 65536)):
        """Treelet-order batch: pick the treelet budget and launch shape; scenes keep the winning partitions."""
        arr = self._jobs(jobs)
        o = self._opts(mode=MODE_TREELET)
        b = np.ascontiguousarray(budgets, np.uint32)
        win, ms = ctypes.c_uint32(0), ctypes.c_float(0)
        _check(_lib.rtx_autotune_batch_treelets(self._h, arr, len(jobs), _ptr(b), len(b), ctypes.byref(o),
                                                ctypes.byref(win), ctypes.byref(ms)), "rtx_autotune_batch_treelets")
        return ({"mode": o.mode, "waves_per_cu": o.waves_per_cu, "lds_stack": o.lds_stack, "refill": o.refill},
                win.value, ms.value)

    def last_kernel_ms(self):
        ms = ctypes.c_float(0)
        _check(_lib.rtx_last_kernel_ms(self._h, ctypes.byref(ms)), "rtx_last_kernel_ms")
        return ms.value

    def timer_start(self):
        _check(_lib.rtx_timer_start(self._h), "rtx_timer_start")

    def timer_stop(self):
        ms = ctypes.c_float(0)
        _check(_lib.rtx_timer_stop(self._h, ctypes.byref(ms)), "rtx_timer_stop")
        return ms.value

    def synchronize(self):
        _check(_lib.rtx_synchronize(self._h), "rtx_synchronize")

    def wait(self):
        """rtx_synchronize; returns the rtx_status (the first error of queued async traces)."""
        return _lib.rtx_synchronize(self._h)

    def trace_async_device(self, rays_buf, n, hits_buf, mode=MODE_PLAIN, waves_per_cu=0, lds_stack=0, refill=0,
                           lanes_per_ray=0, claim_start=0, claim_order=None):
        """rtx_trace_async (hits only, queued) on device buffers; returns the rtx_status of the call."""
        o = _TraceOptions(mode, PREFETCH_NONE, 0.0, 0, waves_per_cu, lds_stack, 0, refill, 0, 0.0, 0.0, lanes_per_ray,
                          claim_start, 0, 0, claim_order.ptr if claim_order is not None else None)
        return _lib.rtx_trace_async(self._h, rays_buf.ptr, n, ctypes.byref(o), hits_buf.ptr)

    def stream(self):
        st = ctypes.c_void_p()
        _check(_lib.rtx_get_stream(self._h, ctypes.byref(st)), "rtx_get_stream")
        return st.value
