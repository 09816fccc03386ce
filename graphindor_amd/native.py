"""ctypes bindings to the native libraries built in-tree.

libgm_hip.so   the C ABI of include/graphindor_amd.h (device engine)
libgraphset.so the host C++ library (graphs, Pattern, Schedule_IEP, plans)

Loading fails loudly (NativeLibraryMissing) when a library has not been
built; nothing here falls back to another implementation.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")

GM_OK = 0
GM_ERR_INVALID = -1
GM_ERR_HIP = -2
GM_ERR_OOM = -3
GM_ERR_UNSUPPORTED = -4
GM_MODE_PATTERN = 0
GM_MODE_CLIQUE = 1
GM_COUNT_ALG_BYTES = 1
GM_MAX_PATTERN = 7


class NativeLibraryMissing(RuntimeError):
    pass


class GmError(RuntimeError):
    pass


class GmPlanDesc(ctypes.Structure):
    _fields_ = [
        ("size", ctypes.c_int32), ("total_prefix_num", ctypes.c_int32), ("basic_prefix_num", ctypes.c_int32),
        ("total_restrict_num", ctypes.c_int32), ("in_exclusion_optimize_num", ctypes.c_int32),
        ("is_vertex_induced", ctypes.c_int32), ("mode", ctypes.c_int32),
        ("adj_mat", ctypes.POINTER(ctypes.c_int32)), ("father_prefix_id", ctypes.POINTER(ctypes.c_int32)),
        ("last", ctypes.POINTER(ctypes.c_int32)), ("next", ctypes.POINTER(ctypes.c_int32)),
        ("break_size", ctypes.POINTER(ctypes.c_int32)), ("loop_set_prefix_id", ctypes.POINTER(ctypes.c_int32)),
        ("restrict_last", ctypes.POINTER(ctypes.c_int32)), ("restrict_next", ctypes.POINTER(ctypes.c_int32)),
        ("restrict_index", ctypes.POINTER(ctypes.c_int32)), ("only_need_size", ctypes.POINTER(ctypes.c_uint8)),
        ("iep_vertex_num", ctypes.c_int32), ("iep_vertex_id", ctypes.POINTER(ctypes.c_int32)),
        ("iep_vertex_flag", ctypes.POINTER(ctypes.c_uint8)), ("iep_vertex_coef", ctypes.POINTER(ctypes.c_int32)),
        ("iep_array_size", ctypes.c_int32), ("iep_coef", ctypes.POINTER(ctypes.c_int32)),
        ("iep_flag", ctypes.POINTER(ctypes.c_uint8)), ("iep_ans_pos", ctypes.POINTER(ctypes.c_int32)),
    ]


class GmTuning(ctypes.Structure):
    _fields_ = [("waves_per_cu", ctypes.c_int32), ("chunk_edges", ctypes.c_int32),
                ("lds_set_cap", ctypes.c_int32), ("bitmap_max_degree", ctypes.c_int32),
                ("lane_tail", ctypes.c_int32), ("reserved", ctypes.c_int32 * 3)]


class GmStats(ctypes.Structure):
    _fields_ = [("kernel_ms", ctypes.c_double), ("alg_bytes", ctypes.c_uint64), ("root_tasks", ctypes.c_uint64),
                ("grid_waves", ctypes.c_int32), ("flags_used", ctypes.c_int32)]


_gm = None
_gs = None
VP = ctypes.c_void_p


def _load(name):
    path = os.path.join(LIB_DIR, name)
    if not os.path.exists(path):
        raise NativeLibraryMissing(f"{path} is not built; run python -m graphindor_amd.build")
    return ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)


def gm():
    """The device engine's C ABI (libgm_hip.so)."""
    global _gm
    if _gm is None:
        lib = _load("libgm_hip.so")
        lib.gm_last_error.restype = ctypes.c_char_p
        lib.gm_device_count.restype = ctypes.c_int
        lib.gm_graph_create.argtypes = [VP, VP, ctypes.c_int32, ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(VP)]
        lib.gm_graph_create.restype = ctypes.c_int
        lib.gm_graph_destroy.argtypes = [VP]
        lib.gm_graph_destroy.restype = None
        lib.gm_plan_create.argtypes = [ctypes.POINTER(GmPlanDesc), ctypes.POINTER(VP)]
        lib.gm_plan_create.restype = ctypes.c_int
        lib.gm_plan_destroy.argtypes = [VP]
        lib.gm_plan_destroy.restype = None
        lib.gm_count.argtypes = [VP, VP, ctypes.POINTER(GmTuning), ctypes.c_int32, VP,
                                 ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(GmStats)]
        lib.gm_count.restype = ctypes.c_int
        lib.gm_count_multi.argtypes = [VP, ctypes.POINTER(VP), ctypes.c_int32, ctypes.POINTER(GmTuning), ctypes.c_int32,
                                       VP, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(GmStats)]
        lib.gm_count_multi.restype = ctypes.c_int
        _gm = lib
    return _gm


def gs():
    """The host C++ library's scripting entry points (libgraphset.so)."""
    global _gs
    if _gs is None:
        gm()
        lib = _load("libgraphset.so")
        lib.gs_last_error.restype = ctypes.c_char_p
        for name, args, res in [
            ("gs_graph_load", [ctypes.c_char_p], VP),
            ("gs_graph_save", [VP, ctypes.c_char_p, ctypes.c_int32], ctypes.c_int),
            ("gs_graph_from_csr", [ctypes.c_int32, ctypes.c_int64, VP, VP, ctypes.c_int64], VP),
            ("gs_graph_rmat", [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                               ctypes.c_uint64, ctypes.c_int], VP),
            ("gs_graph_er", [ctypes.c_int32, ctypes.c_double, ctypes.c_uint64], VP),
            ("gs_graph_free", [VP], None),
            ("gs_graph_info", [VP, ctypes.POINTER(ctypes.c_int64)], None),
            ("gs_graph_csr", [VP, ctypes.POINTER(VP), ctypes.POINTER(VP)], None),
            ("gs_graph_set_tri_cnt", [VP, ctypes.c_int64], None),
            ("gs_graph_count_triangles", [VP], ctypes.c_int64),
            ("gs_graph_reduce_for_clique", [VP], None),
            ("gs_graph_orient_degree", [VP], VP),
            ("gs_set_max_intersection_size", [ctypes.c_int32], None),
            ("gs_set_restrict_validation", [ctypes.c_int], None),
            ("gs_schedule_new", [ctypes.c_int, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int32, ctypes.c_int64, ctypes.c_int64, ctypes.c_int], VP),
            ("gs_schedule_new_plain", [ctypes.c_int, ctypes.c_char_p], VP),
            ("gs_schedule_free", [VP], None),
            ("gs_schedule_set_redundancy", [VP, ctypes.c_int64], None),
            ("gs_schedule_redundancy", [VP], ctypes.c_int64),
            ("gs_schedule_multiplicity", [VP], ctypes.c_int),
            ("gs_schedule_print", [VP], None),
            ("gs_schedule_count_complete", [VP, ctypes.c_int], ctypes.c_int64),
            ("gs_restrictions_valid", [ctypes.c_int, ctypes.c_char_p, ctypes.c_int, VP], ctypes.c_int),
            ("gs_iep_head_multiplicity", [ctypes.c_int, ctypes.c_char_p, ctypes.c_int, VP, ctypes.c_int], ctypes.c_int),
            ("gs_schedule_dump", [VP, VP, ctypes.c_int], ctypes.c_int),
            ("gs_schedule_plan_desc", [VP, ctypes.c_int], ctypes.POINTER(GmPlanDesc)),
            ("gs_gpu_count", [VP, VP, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                              ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
            ("gs_motifs", [ctypes.c_int, ctypes.c_char_p, ctypes.c_int], ctypes.c_int),
        ]:
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = res
        _gs = lib
    return _gs


def device_count():
    return gm().gm_device_count()


def _gm_check(rc, what):
    if rc != GM_OK:
        raise GmError(f"{what} failed ({rc}): {gm().gm_last_error().decode()}")


class Graph:
    """A CSR graph owned by the host library."""

    def __init__(self, handle):
        if not handle:
            raise GmError(gs().gs_last_error().decode() or "graph creation failed")
        self.h = handle

    @classmethod
    def load(cls, path):
        return cls(gs().gs_graph_load(path.encode()))

    @classmethod
    def rmat(cls, scale, edge_factor, seed, a=0.57, b=0.19, c=0.19, permute=True):
        return cls(gs().gs_graph_rmat(scale, edge_factor, a, b, c, seed, 1 if permute else 0))

    @classmethod
    def erdos_renyi(cls, n, p, seed):
        return cls(gs().gs_graph_er(n, p, seed))

    @classmethod
    def from_csr(cls, vertex, edge, tri_cnt=0):
        vertex = np.ascontiguousarray(vertex, dtype=np.int64)
        edge = np.ascontiguousarray(edge, dtype=np.int32)
        return cls(gs().gs_graph_from_csr(len(vertex) - 1, len(edge), vertex.ctypes.data, edge.ctypes.data, tri_cnt))

    def free(self):
        if self.h:
            gs().gs_graph_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def info(self):
        out = (ctypes.c_int64 * 4)()
        gs().gs_graph_info(self.h, out)
        return {"v_cnt": out[0], "e_cnt": out[1], "tri_cnt": out[2], "max_isect": out[3]}

    def csr(self):
        """(vertex int64[n+1], edge int32[nnz]) views of the host arrays (valid while self lives)."""
        inf = self.info()
        v, e = VP(), VP()
        gs().gs_graph_csr(self.h, ctypes.byref(v), ctypes.byref(e))
        vertex = np.ctypeslib.as_array(ctypes.cast(v, ctypes.POINTER(ctypes.c_int64)), shape=(inf["v_cnt"] + 1,))
        if inf["e_cnt"] > 0:
            edge = np.ctypeslib.as_array(ctypes.cast(e, ctypes.POINTER(ctypes.c_int32)), shape=(inf["e_cnt"],))
        else:
            edge = np.zeros(0, dtype=np.int32)
        return vertex, edge

    def save(self, path, max_isect=None):
        mi = self.info()["max_isect"] if max_isect is None else max_isect
        if gs().gs_graph_save(self.h, path.encode(), mi) != 0:
            raise GmError(f"cannot write {path}")

    def reduce_for_clique(self):
        gs().gs_graph_reduce_for_clique(self.h)

    def orient_by_degree(self):
        """New graph: each edge once, from the smaller (degree, id) end."""
        return Graph(gs().gs_graph_orient_degree(self.h))

    def set_tri_cnt(self, t):
        gs().gs_graph_set_tri_cnt(self.h, t)


class Schedule:
    """A Schedule_IEP built by the host library's schedule compiler."""

    def __init__(self, handle):
        if not handle:
            raise GmError(gs().gs_last_error().decode() or "schedule creation failed")
        self.h = handle

    @classmethod
    def build(cls, n, adj, v_cnt, e_cnt, tri_cnt=0, perf=1, restricts=1, iep=True, vertex_induced=False):
        return cls(gs().gs_schedule_new(n, adj.encode(), perf, restricts, 1 if iep else 0, v_cnt, e_cnt, tri_cnt,
                                        1 if vertex_induced else 0))

    @classmethod
    def plain(cls, n, adj):
        return cls(gs().gs_schedule_new_plain(n, adj.encode()))

    def free(self):
        if self.h:
            gs().gs_schedule_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def redundancy(self):
        return gs().gs_schedule_redundancy(self.h)

    @redundancy.setter
    def redundancy(self, r):
        gs().gs_schedule_set_redundancy(self.h, r)

    def dump(self):
        buf = (ctypes.c_int32 * 4096)()
        n = gs().gs_schedule_dump(self.h, buf, 4096)
        if n < 0:
            raise GmError("schedule dump too large")
        return list(buf[:n])

    def plan_desc(self, mode=GM_MODE_PATTERN):
        return gs().gs_schedule_plan_desc(self.h, mode)

    def count_complete(self, m):
        return gs().gs_schedule_count_complete(self.h, m)

    def print(self):
        gs().gs_schedule_print(self.h)


def set_restrict_validation(mode):
    """'exact' (default) or 'reference'."""
    gs().gs_set_restrict_validation(1 if mode == "reference" else 0)


def restrictions_valid(n, adj, pairs):
    arr = (ctypes.c_int32 * (2 * len(pairs) + 1))(*[x for p in pairs for x in p])
    return bool(gs().gs_restrictions_valid(n, adj.encode(), len(pairs), arr))


def iep_head_multiplicity(n, adj, pairs, k):
    arr = (ctypes.c_int32 * (2 * len(pairs) + 1))(*[x for p in pairs for x in p])
    return gs().gs_iep_head_multiplicity(n, adj.encode(), len(pairs), arr, k)


def motifs(k):
    buf = ctypes.create_string_buffer(k * k * 256)
    n = gs().gs_motifs(k, buf, len(buf))
    raw = buf.raw[: n * k * k].decode()
    return [raw[i * k * k:(i + 1) * k * k] for i in range(n)]


class DeviceGraph:
    """A CSR uploaded through gm_graph_create."""

    def __init__(self, vertex, edge, max_isect=0):
        self.vertex = np.ascontiguousarray(vertex, dtype=np.int64)
        self.edge = np.ascontiguousarray(edge, dtype=np.int32)
        h = VP()
        _gm_check(gm().gm_graph_create(self.vertex.ctypes.data, self.edge.ctypes.data if len(self.edge) else None,
                                       len(self.vertex) - 1, len(self.edge), max_isect, ctypes.byref(h)),
                  "gm_graph_create")
        self.h = h

    def free(self):
        if self.h:
            gm().gm_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DevicePlan:
    def __init__(self, desc):
        h = VP()
        _gm_check(gm().gm_plan_create(desc, ctypes.byref(h)), "gm_plan_create")
        self.h = h

    def free(self):
        if self.h:
            gm().gm_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def count(dgraph, dplan, flags=0, waves_per_cu=0, bitmap_max_degree=0, lane_tail=0):
    """Raw sum (before the redundancy division) and the gm_stats, via gm_count."""
    raw = ctypes.c_uint64(0)
    st = GmStats()
    tun = GmTuning()
    tun.waves_per_cu = waves_per_cu
    tun.bitmap_max_degree = bitmap_max_degree
    tun.lane_tail = lane_tail
    _gm_check(gm().gm_count(dgraph.h, dplan.h, ctypes.byref(tun), flags, None, ctypes.byref(raw), ctypes.byref(st)),
              "gm_count")
    return raw.value, st
