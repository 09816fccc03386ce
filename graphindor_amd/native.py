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
GM_ERR_TIMEOUT = -5
GM_MODE_PATTERN = 0
GM_MODE_CLIQUE = 1
GM_COUNT_ALG_BYTES = 1
GM_COUNT_EAGER_DONATION = 2
GM_COUNT_DEEP_DONATION = 4
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
                ("lane_tail", ctypes.c_int32), ("root_stride", ctypes.c_int32), ("slice_len", ctypes.c_int32),
                ("bitmap_block_degree", ctypes.c_int32), ("time_limit_ms", ctypes.c_int32),
                ("hub_degree", ctypes.c_int32)]


class GmStats(ctypes.Structure):
    _fields_ = [("kernel_ms", ctypes.c_double), ("alg_bytes", ctypes.c_uint64), ("root_tasks", ctypes.c_uint64),
                ("grid_waves", ctypes.c_int32), ("flags_used", ctypes.c_int32), ("list_ms", ctypes.c_double),
                ("bitmap_ms", ctypes.c_double), ("bitmap_alg_bytes", ctypes.c_uint64),
                ("bitmap_block_ms", ctypes.c_double), ("bitmap_block_alg_bytes", ctypes.c_uint64),
                ("slices", ctypes.c_uint64), ("donations", ctypes.c_uint64), ("queue_retries", ctypes.c_uint64),
                ("slice_wave_ms", ctypes.c_double), ("idle_wave_ms", ctypes.c_double),
                ("hub_vertices", ctypes.c_int64),
                ("deep_donations", ctypes.c_uint64)]


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
        lib.gm_count_static.argtypes = [VP, VP, ctypes.POINTER(GmTuning), ctypes.c_int32, VP,
                                        ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(GmStats)]
        lib.gm_count_static.restype = ctypes.c_int
        lib.gm_count_multi.argtypes = [VP, ctypes.POINTER(VP), ctypes.c_int32, ctypes.POINTER(GmTuning), ctypes.c_int32,
                                       VP, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(GmStats)]
        lib.gm_count_multi.restype = ctypes.c_int
        lib.gm_count_batch.argtypes = [VP, VP, ctypes.c_int32, ctypes.POINTER(VP), ctypes.c_int32,
                                       ctypes.POINTER(GmTuning), ctypes.c_int32, VP, ctypes.POINTER(ctypes.c_uint64),
                                       ctypes.POINTER(GmStats)]
        lib.gm_count_batch.restype = ctypes.c_int
        lib.gm_batch_create.argtypes = [VP, VP, ctypes.c_int32, ctypes.POINTER(VP), ctypes.c_int32,
                                        ctypes.POINTER(GmTuning), ctypes.c_int32, ctypes.POINTER(VP)]
        lib.gm_batch_create.restype = ctypes.c_int
        lib.gm_batch_run.argtypes = [VP, VP, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(GmStats)]
        lib.gm_batch_run.restype = ctypes.c_int
        lib.gm_batch_destroy.argtypes = [VP]
        lib.gm_batch_destroy.restype = None
        lib.gm_mlp_create.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_uint64, ctypes.POINTER(VP)]
        lib.gm_mlp_destroy.argtypes = [VP]
        lib.gm_mlp_destroy.restype = None
        lib.gm_mlp_param_count.argtypes = [VP]
        lib.gm_mlp_param_count.restype = ctypes.c_int64
        lib.gm_mlp_get_params.argtypes = [VP, VP]
        lib.gm_mlp_set_params.argtypes = [VP, VP]
        lib.gm_mlp_get_grads.argtypes = [VP, VP]
        lib.gm_mlp_predict.argtypes = [VP, VP, ctypes.c_int32, VP]
        lib.gm_mlp_train.argtypes = [VP, VP, VP, VP, ctypes.c_int32, ctypes.c_int32, ctypes.c_float, ctypes.c_float,
                                     ctypes.POINTER(ctypes.c_float)]
        lib.gm_mlp_last_error.restype = ctypes.c_char_p
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
            ("gs_graph_batch", [ctypes.c_int, ctypes.c_uint64, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                ctypes.c_double, VP, VP], VP),
            ("gs_graph_free", [VP], None),
            ("gs_graph_info", [VP, ctypes.POINTER(ctypes.c_int64)], None),
            ("gs_graph_csr", [VP, ctypes.POINTER(VP), ctypes.POINTER(VP)], None),
            ("gs_graph_set_tri_cnt", [VP, ctypes.c_int64], None),
            ("gs_graph_count_triangles", [VP], ctypes.c_int64),
            ("gs_graph_reduce_for_clique", [VP], None),
            ("gs_graph_orient_degree", [VP], VP),
            ("gs_graph_orient_degree_rank", [VP], VP),
            ("gs_graph_relabel_degree", [VP, ctypes.c_int], VP),
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
            ("gs_autotune_batch", [VP, VP, ctypes.c_int32, ctypes.POINTER(VP), ctypes.c_int32, ctypes.c_char_p,
                                   ctypes.c_char_p, ctypes.c_int32, VP], ctypes.c_int),
            ("gs_degree_class", [VP, ctypes.c_int32, ctypes.c_char_p, ctypes.c_int], ctypes.c_int),
            ("gs_schedule_candidates", [ctypes.c_int, ctypes.c_char_p, ctypes.c_int, ctypes.c_int32, ctypes.c_int64,
                                        ctypes.c_int64, VP, ctypes.c_int, VP, ctypes.c_int], ctypes.c_int),
            ("gs_schedule_from_candidate", [ctypes.c_int, ctypes.c_char_p, VP, ctypes.c_int, VP, ctypes.c_int], VP),
            ("gs_encode_candidates", [ctypes.c_int, ctypes.c_char_p, ctypes.c_int, VP, ctypes.c_int32, ctypes.c_int64,
                                      VP, ctypes.c_int], ctypes.c_int),
            ("gs_feature_count", [], ctypes.c_int),
            ("gs_costmodel_new", [ctypes.c_int, ctypes.c_uint64], VP),
            ("gs_costmodel_load", [ctypes.c_char_p], VP),
            ("gs_costmodel_free", [VP], None),
            ("gs_costmodel_save", [VP, ctypes.c_char_p], ctypes.c_int),
            ("gs_costmodel_train", [VP, VP, VP, VP, ctypes.c_int32, ctypes.c_int32, ctypes.c_float, ctypes.c_float,
                                    ctypes.c_int, ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
            ("gs_costmodel_predict", [VP, VP, ctypes.c_int32, VP], ctypes.c_int),
            ("gs_costmodel_set_margin", [VP, ctypes.c_float], None),
            ("gs_schedule_select_measured", [ctypes.c_int, ctypes.c_char_p, ctypes.c_int, VP, VP, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, VP, ctypes.POINTER(ctypes.c_double)], VP),
            ("gs_costmodel_margin", [VP], ctypes.c_float),
            ("gs_schedule_select", [ctypes.c_int, ctypes.c_char_p, ctypes.c_int, VP, ctypes.c_int32, ctypes.c_int64, VP,
                                    VP, ctypes.POINTER(ctypes.c_double)], VP),
        ]:
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = res
        _gs = lib
    return _gs


def device_count():
    return gm().gm_device_count()


class GmTimeout(GmError):
    """gm_tuning.time_limit_ms passed before the count finished."""


def _gm_check(rc, what):
    if rc == GM_ERR_TIMEOUT:
        raise GmTimeout(f"{what}: {gm().gm_last_error().decode()}")
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
    def batch(cls, count, seed_base, log2_min=10.0, log2_max=17.0, deg_min=4.0, deg_max=64.0):
        """(union graph, vertex offsets, edge offsets) of a seeded batch."""
        voff = np.zeros(count + 1, dtype=np.int64)
        eoff = np.zeros(count + 1, dtype=np.int64)
        h = gs().gs_graph_batch(count, seed_base, log2_min, log2_max, deg_min, deg_max, voff.ctypes.data,
                                eoff.ctypes.data)
        return cls(h), voff, eoff

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

    def relabel_by_degree(self, ascending=True):
        """New graph: the same edges, vertices renumbered by (degree, id)
        ascending (or descending); every pattern count is unchanged."""
        return Graph(gs().gs_graph_relabel_degree(self.h, 1 if ascending else 0))

    def orient_by_degree_rank(self):
        """New graph: vertices renumbered by (degree, id) rank, each edge once
        from the lower rank to the higher (every row holds larger ids only)."""
        return Graph(gs().gs_graph_orient_degree_rank(self.h))

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

    @classmethod
    def from_candidate(cls, n, adj, cand):
        """Exactly the given Candidate (Schedule_IEP(pattern, candidate))."""
        order = (ctypes.c_int32 * n)(*cand.order)
        flat = [x for p in cand.pairs for x in p]
        pairs = (ctypes.c_int32 * (len(flat) + 1))(*flat)
        return cls(gs().gs_schedule_from_candidate(n, adj.encode(), order, len(cand.pairs), pairs,
                                                   1 if cand.use_iep else 0))

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
        """Pointer to the flattened plan, which the schedule owns; the pointer
        keeps the schedule alive, so a temporary schedule is safe to use."""
        d = gs().gs_schedule_plan_desc(self.h, mode)
        if d:
            d._owner = self
        return d

    def count_complete(self, m):
        return gs().gs_schedule_count_complete(self.h, m)

    def print(self):
        gs().gs_schedule_print(self.h)


class Candidate:
    """One (order, restriction set, IEP) candidate with its analytic cost."""

    __slots__ = ("order", "pairs", "use_iep", "analytic_cost")

    def __init__(self, order, pairs, use_iep, analytic_cost):
        self.order, self.pairs, self.use_iep, self.analytic_cost = tuple(order), tuple(pairs), use_iep, analytic_cost

    def __repr__(self):
        return f"Candidate(order={self.order}, pairs={self.pairs}, iep={self.use_iep}, cost={self.analytic_cost:.4g})"


def schedule_candidates(n, adj, v_cnt, e_cnt, tri_cnt, iep=True):
    """Every candidate the analytic selector weighs (Schedule_IEP::enumerate_candidates)."""
    cap, ccap = 1 << 20, 1 << 15
    buf = (ctypes.c_int32 * cap)()
    costs = (ctypes.c_double * ccap)()
    k = gs().gs_schedule_candidates(n, adj.encode(), 1 if iep else 0, v_cnt, e_cnt, tri_cnt, buf, cap, costs, ccap)
    if k < 0:
        raise GmError(gs().gs_last_error().decode() if k == -1 else "too many candidates")
    out, pos = [], 0
    for i in range(k):
        use_iep, npairs = bool(buf[pos]), buf[pos + 1]
        order = list(buf[pos + 2:pos + 2 + n])
        fl = list(buf[pos + 2 + n:pos + 2 + n + 2 * npairs])
        pos += 2 + n + 2 * npairs
        out.append(Candidate(order, [(fl[2 * j], fl[2 * j + 1]) for j in range(npairs)], use_iep, costs[i]))
    return out


def encode_candidates(n, adj, vertex, tri_cnt, iep=True):
    """Feature rows (float32 [candidates, F]) in the order of schedule_candidates."""
    v = np.ascontiguousarray(vertex, dtype=np.int64)
    F = gs().gs_feature_count()
    cap = 1 << 14
    out = np.empty((cap, F), dtype=np.float32)
    k = gs().gs_encode_candidates(n, adj.encode(), 1 if iep else 0, v.ctypes.data, len(v) - 1, tri_cnt,
                                  out.ctypes.data, cap)
    if k < 0:
        raise GmError(gs().gs_last_error().decode() if k == -1 else "too many candidates")
    return out[:k].copy()


class CostModel:
    """The learned schedule cost model (host CostModel over the gm_mlp kernels)."""

    def __init__(self, handle):
        if not handle:
            raise GmError(gs().gs_last_error().decode() or "cost model creation failed")
        self.h = handle

    @classmethod
    def new(cls, hidden=256, seed=6):
        return cls(gs().gs_costmodel_new(hidden, seed))

    @classmethod
    def load(cls, path):
        return cls(gs().gs_costmodel_load(str(path).encode()))

    def save(self, path):
        if gs().gs_costmodel_save(self.h, str(path).encode()) != 0:
            raise GmError(gs().gs_last_error().decode())

    def train(self, x, log_ms, group_off, steps, lr=1e-3, weight_decay=1e-4, fit_norm=True):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(log_ms, dtype=np.float32)
        go = np.ascontiguousarray(group_off, dtype=np.int32)
        loss = ctypes.c_float(0)
        if gs().gs_costmodel_train(self.h, x.ctypes.data, t.ctypes.data, go.ctypes.data, len(go) - 1, steps, lr,
                                   weight_decay, 1 if fit_norm else 0, ctypes.byref(loss)) != 0:
            raise GmError(gs().gs_last_error().decode())
        return loss.value

    @property
    def margin(self):
        return gs().gs_costmodel_margin(self.h)

    @margin.setter
    def margin(self, m):
        gs().gs_costmodel_set_margin(self.h, float(m))

    def predict(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.empty(x.shape[0], dtype=np.float32)
        if gs().gs_costmodel_predict(self.h, x.ctypes.data, x.shape[0], y.ctypes.data) != 0:
            raise GmError(gs().gs_last_error().decode())
        return y

    def free(self):
        if self.h:
            gs().gs_costmodel_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def select_schedule(n, adj, vertex, tri_cnt, model=None, iep=True):
    """(Schedule, info) chosen by `model` among the analytic selector's
    candidates, or the analytic pick when model is None."""
    v = np.ascontiguousarray(vertex, dtype=np.int64)
    info = (ctypes.c_int32 * 4)()
    ms = ctypes.c_double(0)
    h = gs().gs_schedule_select(n, adj.encode(), 1 if iep else 0, v.ctypes.data, len(v) - 1, tri_cnt,
                                model.h if model is not None else None, info, ctypes.byref(ms))
    s = Schedule(h)
    return s, {"candidates": info[0], "chosen": info[1], "analytic": info[2], "learned": bool(info[3]),
               "select_ms": ms.value}


def select_schedule_measured(n, adj, graph, model=None, iep=True, top_k=3, sample_roots=4096, time_limit_ms=200,
                             root_stride=0, lane_tail=0, hub_degree=0, waves_per_cu=0):
    """(Schedule, info): the model's top_k proposals and the analytic pick
    timed on a root sample of `graph` (a host Graph; its device copy is
    used) -- about sample_roots roots, or exactly the edges e % root_stride
    == 0 when root_stride > 0 -- under the given gm_tuning knobs, the
    fastest kept."""
    info = (ctypes.c_int32 * 4)()
    ms = ctypes.c_double(0)
    h = gs().gs_schedule_select_measured(n, adj.encode(), 1 if iep else 0, graph.h, model.h if model is not None else None,
                                         top_k, sample_roots, root_stride, time_limit_ms, lane_tail, hub_degree,
                                         waves_per_cu, info, ctypes.byref(ms))
    s = Schedule(h)
    return s, {"candidates": info[0], "chosen": info[1], "analytic": info[2], "learned": bool(info[3]),
               "select_ms": ms.value}


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
    """Connected k-vertex patterns in MotifGenerator order (n*n adjacency strings)."""
    cap = k * k * 256
    buf = ctypes.create_string_buffer(cap)
    n = gs().gs_motifs(k, buf, len(buf))
    if n < 0:  # the buffer was too small: -n bytes are needed
        buf = ctypes.create_string_buffer(-n)
        n = gs().gs_motifs(k, buf, len(buf))
    if n < 0:
        raise RuntimeError(f"gs_motifs({k}) failed ({n})")
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


def count(dgraph, dplan, flags=0, **tuning):
    """Raw sum (before the redundancy division) and the gm_stats, via gm_count.
    Keyword arguments are gm_tuning fields (see _tuning)."""
    raw = ctypes.c_uint64(0)
    st = GmStats()
    tun = _tuning(**tuning)
    _gm_check(gm().gm_count(dgraph.h, dplan.h, ctypes.byref(tun), flags, None, ctypes.byref(raw), ctypes.byref(st)),
              "gm_count")
    return raw.value, st


def count_static(dgraph, dplan, flags=0, **tuning):
    """gm_count through gm_count_static: the same raw sum, computed in the
    library's static per-device workspace without allocating device memory."""
    raw = ctypes.c_uint64(0)
    st = GmStats()
    tun = _tuning(**tuning)
    _gm_check(gm().gm_count_static(dgraph.h, dplan.h, ctypes.byref(tun), flags, None, ctypes.byref(raw),
                                   ctypes.byref(st)), "gm_count_static")
    return raw.value, st


def _tuning(waves_per_cu=0, bitmap_max_degree=0, lane_tail=0, chunk_edges=0, root_stride=0, slice_len=0,
            bitmap_block_degree=0, time_limit_ms=0, hub_degree=0, lds_set_cap=0):
    tun = GmTuning()
    tun.hub_degree = hub_degree
    tun.lds_set_cap = lds_set_cap
    tun.time_limit_ms = time_limit_ms
    tun.bitmap_block_degree = bitmap_block_degree
    tun.root_stride = root_stride
    tun.slice_len = slice_len
    tun.waves_per_cu = waves_per_cu
    tun.bitmap_max_degree = bitmap_max_degree
    tun.lane_tail = lane_tail
    tun.chunk_edges = chunk_edges
    return tun


def count_multi(dgraph, dplans, flags=0, **tuning):
    """Raw sums of several plans over one graph in one call (gm_count_multi)."""
    arr = (VP * len(dplans))(*[p.h for p in dplans])
    out = (ctypes.c_uint64 * len(dplans))()
    st = GmStats()
    tun = _tuning(**tuning)
    _gm_check(gm().gm_count_multi(dgraph.h, arr, len(dplans), ctypes.byref(tun), flags, None, out, ctypes.byref(st)),
              "gm_count_multi")
    return list(out), st


def count_batch(dgraph, edge_offsets, dplans, flags=0, **tuning):
    """Raw sums [graph][plan] over a batch held as one CSR (gm_count_batch)."""
    eoff = np.ascontiguousarray(edge_offsets, dtype=np.int64)
    ng = len(eoff) - 1
    arr = (VP * len(dplans))(*[p.h for p in dplans])
    out = (ctypes.c_uint64 * (ng * len(dplans)))()
    st = GmStats()
    tun = _tuning(**tuning)
    _gm_check(gm().gm_count_batch(dgraph.h, eoff.ctypes.data, ng, arr, len(dplans), ctypes.byref(tun), flags, None,
                                  out, ctypes.byref(st)), "gm_count_batch")
    vals = list(out)
    return [vals[i * len(dplans):(i + 1) * len(dplans)] for i in range(ng)], st


class Batch:
    """A batch count captured into a hipGraph (gm_batch_create / gm_batch_run)."""

    def __init__(self, dgraph, edge_offsets, dplans, flags=0, **tuning):
        self.eoff = np.ascontiguousarray(edge_offsets, dtype=np.int64)
        self.ng = len(self.eoff) - 1
        self.np = len(dplans)
        self._keep = (dgraph, dplans)
        arr = (VP * self.np)(*[p.h for p in dplans])
        tun = _tuning(**tuning)
        h = VP()
        _gm_check(gm().gm_batch_create(dgraph.h, self.eoff.ctypes.data, self.ng, arr, self.np, ctypes.byref(tun),
                                       flags, ctypes.byref(h)), "gm_batch_create")
        self.h = h

    def run(self):
        out = (ctypes.c_uint64 * (self.ng * self.np))()
        st = GmStats()
        _gm_check(gm().gm_batch_run(self.h, None, out, ctypes.byref(st)), "gm_batch_run")
        vals = list(out)
        return [vals[i * self.np:(i + 1) * self.np] for i in range(self.ng)], st

    def free(self):
        if self.h:
            gm().gm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def degree_class(vertex):
    v = np.ascontiguousarray(vertex, dtype=np.int64)
    buf = ctypes.create_string_buffer(64)
    gs().gs_degree_class(v.ctypes.data, len(v) - 1, buf, 64)
    return buf.value.decode()


def autotune_batch(dgraph, edge_offsets, dplans, key, cache_path=None, root_stride=1):
    """Fastest tuning on a sample of the batch (cached per key); returns a dict."""
    eoff = np.ascontiguousarray(edge_offsets, dtype=np.int64)
    arr = (VP * len(dplans))(*[p.h for p in dplans])
    out = (ctypes.c_int64 * 6)()
    rc = gs().gs_autotune_batch(dgraph.h, eoff.ctypes.data, len(eoff) - 1, arr, len(dplans), key.encode(),
                                cache_path.encode() if cache_path else None, root_stride, out)
    if rc != 0:
        raise GmError(gs().gs_last_error().decode())
    return {"waves_per_cu": out[0], "chunk_edges": out[1], "lane_tail": out[2], "sample_ms": out[3] / 1000.0,
            "from_cache": bool(out[4]), "tried": out[5]}


class Mlp:
    """The cost model's FP32 MLP on the GPU (gm_mlp_*): F inputs, 2 x H ReLU, 1 output."""

    def __init__(self, in_dim, hidden=256, seed=6):
        h = VP()
        rc = gm().gm_mlp_create(in_dim, hidden, seed, ctypes.byref(h))
        if rc != 0:
            raise GmError(gm().gm_mlp_last_error().decode())
        self.h, self.in_dim, self.hidden = h, in_dim, hidden

    def _check(self, rc, what):
        if rc != 0:
            raise GmError(f"{what}: {gm().gm_mlp_last_error().decode()}")

    @property
    def nparams(self):
        return gm().gm_mlp_param_count(self.h)

    def get_params(self):
        out = np.empty(self.nparams, dtype=np.float32)
        self._check(gm().gm_mlp_get_params(self.h, out.ctypes.data), "gm_mlp_get_params")
        return out

    def get_grads(self):
        out = np.empty(self.nparams, dtype=np.float32)
        self._check(gm().gm_mlp_get_grads(self.h, out.ctypes.data), "gm_mlp_get_grads")
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparams
        self._check(gm().gm_mlp_set_params(self.h, p.ctypes.data), "gm_mlp_set_params")

    def predict(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        assert x.ndim == 2 and x.shape[1] == self.in_dim
        y = np.empty(x.shape[0], dtype=np.float32)
        self._check(gm().gm_mlp_predict(self.h, x.ctypes.data, x.shape[0], y.ctypes.data), "gm_mlp_predict")
        return y

    def train(self, x, t, group_off, steps, lr=1e-3, weight_decay=0.0):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        go = np.ascontiguousarray(group_off, dtype=np.int32)
        assert x.shape == (go[-1], self.in_dim) and t.shape == (go[-1],)
        loss = ctypes.c_float(0)
        self._check(gm().gm_mlp_train(self.h, x.ctypes.data, t.ctypes.data, go.ctypes.data, len(go) - 1, steps, lr,
                                      weight_decay, ctypes.byref(loss)), "gm_mlp_train")
        return loss.value

    def free(self):
        if self.h:
            gm().gm_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
