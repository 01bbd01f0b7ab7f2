"""ctypes wrapper of liboracle.so — the CPU restatement used as the checker.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product path.  See oracle.h for the
reference functions each entry point restates.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built; run `make -C oracle`")
_lib = ctypes.CDLL(LIB_PATH)

RAY_DTYPE = np.dtype([("org", "<f4", 3), ("tmin", "<f4"), ("dir", "<f4", 3), ("tmax", "<f4"),
                      ("flags", "<u4"), ("cull_mask", "<u4"), ("sbt_record_offset", "<u4"),
                      ("sbt_record_stride", "<u4"), ("miss_index", "<u4"), ("reserved", "<u4", 3)])
HIT_DTYPE = np.dtype([("hit", "<u4"), ("primitive_index", "<u4"), ("geometry_index", "<u4"),
                      ("instance_index", "<u4"), ("t", "<f4"), ("p", "<f4", 3), ("bary", "<f4", 3),
                      ("n_records", "<u4"), ("n_nodes", "<u4"), ("fetch_bytes", "<u4"), ("n_any_hits", "<u4"),
                      ("max_level", "<u4"), ("instance_leaf", "<u8"), ("accessed_bytes", "<u4"),
                      ("path_level", "<u4")])
RECORD_DTYPE = np.dtype([("offset", "<u8"), ("treelet", "<u4"), ("size", "<u2"), ("type", "u1"),
                         ("level", "u1")])

_vp = ctypes.c_void_p
_u32, _u64, _i32, _f32 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_float
_P = ctypes.POINTER


def _sig(name, restype, *args):
    f = getattr(_lib, name)
    f.restype = restype
    f.argtypes = list(args)


_sig("orc_create_treelets", _i32, _vp, _u64, _u64, _u32, _i32, _P(_vp))
_sig("orc_free_partition", None, _vp)
_sig("orc_partition_count", _u32, _vp)
_sig("orc_partition_list_entries", _u64, _vp)
_sig("orc_partition_total_bytes", _u64, _vp)
_sig("orc_partition_copy", _i32, _vp, _vp, _vp, _vp, _vp)
_sig("orc_treelet_layout", _i32, _vp, _u64, _u64, _vp, _u32, _vp, _vp)
_sig("orc_trace", _i32, _vp, _u64, _u64, _vp, _i32, _vp, _u32, _vp, _vp, _u64, _vp)
_sig("orc_trace_parallel", _i32, _vp, _u64, _u64, _vp, _i32, _vp, _u32, _vp, _i32)
_sig("orc_ray_box_test", _i32, _vp, _vp, _vp, _vp, _f32, _f32, _P(_f32))
_sig("orc_calculate_idir", None, _vp, _vp)
_sig("orc_mt_ray_triangle_test", _i32, _vp, _vp, _vp, _vp, _vp, _P(_f32))
_sig("orc_barycentric", None, _vp, _vp, _vp, _vp, _vp)
_sig("orc_set_child_bounds", None, _vp, _i32, _vp, _vp)
_sig("orc_make_transformed_ray", None, _vp, _vp, _vp, _f32, _f32, _vp, _vp, _P(_f32), _P(_f32), _P(_f32))
_sig("orc_validate", _i32, _vp, _u64, _u64, _P(_u64))


class OracleError(RuntimeError):
    def __init__(self, status, what):
        self.status = status
        super().__init__(f"{what}: status {status}")


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _f3(v):
    return np.ascontiguousarray(v, dtype=np.float32)


class Partition:
    """createTreelets + buildNodeToRootMap output."""

    def __init__(self, arena, tlas, budget, expand_duplicates=False):
        self.arena = np.ascontiguousarray(arena, dtype=np.uint8)
        self.tlas = tlas
        self.budget = budget
        h = ctypes.c_void_p()
        st = _lib.orc_create_treelets(_p(self.arena), self.arena.nbytes, tlas, budget, 1 if expand_duplicates else 0,
                                      ctypes.byref(h))
        if st != 0:
            raise OracleError(st, "orc_create_treelets")
        self._h = h
        n = _lib.orc_partition_count(h)
        m = _lib.orc_partition_list_entries(h)
        self.roots = np.empty(n, np.uint64)
        self.list_offsets = np.empty(n + 1, np.uint64)
        self.list_nodes = np.empty(max(m, 1), np.uint64)
        self.owner = np.empty(self.arena.nbytes // 64, np.uint32)
        _lib.orc_partition_copy(h, _p(self.roots), _p(self.list_offsets), _p(self.list_nodes), _p(self.owner))
        self.list_nodes = self.list_nodes[:m]
        self.total_bytes = _lib.orc_partition_total_bytes(h)

    @property
    def n_treelets(self):
        return len(self.roots)

    def layout(self, stride=0):
        """remapBVHToTreeletLayout (vulkan_ray_tracing.cc:1473-1509) with -treelet_remap_stride
        `stride`: (remap[arena/64] of layout offsets, ~0 for nodes in no list; metadata
        [n_treelets, budget/64] of byte offsets from each treelet root's place, -1 past a list)."""
        remap = np.empty(self.arena.nbytes // 64, np.uint64)
        meta = np.empty((self.n_treelets, self.budget // 64), np.int32)
        st = _lib.orc_treelet_layout(_p(self.arena), self.arena.nbytes, self.tlas, self._h, stride, _p(remap),
                                     _p(meta))
        if st != 0:
            raise OracleError(st, "orc_treelet_layout")
        return remap, meta

    def nodes_of(self, t):
        return self.list_nodes[self.list_offsets[t]:self.list_offsets[t + 1]]

    def close(self):
        if self._h:
            _lib.orc_free_partition(self._h)
            self._h = ctypes.c_void_p(None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def trace(arena, tlas, rays, mode=0, partition=None, records=True):
    """orc_trace: returns hits, or (hits, offsets, records)."""
    arena = np.ascontiguousarray(arena, dtype=np.uint8)
    rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
    n = len(rays)
    hits = np.zeros(n, HIT_DTYPE)
    ph = partition._h if partition is not None else None
    offsets = np.zeros(n + 1, np.uint64)
    st = _lib.orc_trace(_p(arena), arena.nbytes, tlas, ph, mode, _p(rays), n, _p(hits), None, 0, _p(offsets))
    if st != 0:
        raise OracleError(st, "orc_trace")
    if not records:
        return hits
    total = int(offsets[-1])
    recs = np.zeros(max(total, 1), RECORD_DTYPE)
    st = _lib.orc_trace(_p(arena), arena.nbytes, tlas, ph, mode, _p(rays), n, _p(hits), _p(recs), total,
                        _p(offsets))
    if st != 0:
        raise OracleError(st, "orc_trace")
    return hits, offsets, recs[:total]


def trace_parallel(arena, tlas, rays, mode=0, partition=None, threads=1):
    arena = np.ascontiguousarray(arena, dtype=np.uint8)
    rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
    hits = np.zeros(len(rays), HIT_DTYPE)
    ph = partition._h if partition is not None else None
    st = _lib.orc_trace_parallel(_p(arena), arena.nbytes, tlas, ph, mode, _p(rays), len(rays), _p(hits), threads)
    if st != 0:
        raise OracleError(st, "orc_trace_parallel")
    return hits


def validate(arena, tlas):
    arena = np.ascontiguousarray(arena, dtype=np.uint8)
    n = ctypes.c_uint64(0)
    st = _lib.orc_validate(_p(arena), arena.nbytes, tlas, ctypes.byref(n))
    if st != 0:
        raise OracleError(st, "orc_validate")
    return n.value


def ray_box_test(lo, hi, idir, org, tmin, tmax):
    a = [_f3(lo), _f3(hi), _f3(idir), _f3(org)]
    th = ctypes.c_float(0)
    r = _lib.orc_ray_box_test(*(_p(x) for x in a), tmin, tmax, ctypes.byref(th))
    return bool(r), th.value


def calculate_idir(d):
    dd = _f3(d)
    out = np.zeros(3, np.float32)
    _lib.orc_calculate_idir(_p(dd), _p(out))
    return out


def mt_ray_triangle_test(p0, p1, p2, org, d):
    a = [_f3(p0), _f3(p1), _f3(p2), _f3(org), _f3(d)]
    th = ctypes.c_float(0)
    r = _lib.orc_mt_ray_triangle_test(*(_p(x) for x in a), ctypes.byref(th))
    return bool(r), th.value


def barycentric(p, a, b, c):
    args = [_f3(p), _f3(a), _f3(b), _f3(c)]
    out = np.zeros(3, np.float32)
    _lib.orc_barycentric(*(_p(x) for x in args), _p(out))
    return out


def set_child_bounds(node64, child):
    node = np.ascontiguousarray(node64, dtype=np.uint8)
    lo, hi = np.zeros(3, np.float32), np.zeros(3, np.float32)
    _lib.orc_set_child_bounds(_p(node), child, _p(lo), _p(hi))
    return lo, hi


def make_transformed_ray(rows12, org, d, tmin, tmax):
    rows = np.ascontiguousarray(rows12, dtype=np.float32)
    o3, d3 = _f3(org), _f3(d)
    oo, od = np.zeros(3, np.float32), np.zeros(3, np.float32)
    a, b, m = ctypes.c_float(0), ctypes.c_float(0), ctypes.c_float(0)
    _lib.orc_make_transformed_ray(_p(rows), _p(o3), _p(d3), tmin, tmax, _p(oo), _p(od), ctypes.byref(a),
                                  ctypes.byref(b), ctypes.byref(m))
    return oo, od, a.value, b.value, m.value
