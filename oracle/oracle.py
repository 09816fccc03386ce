"""ctypes binding of the CPU oracle (oracle/liboracle.so) and of the
reference build (oracle/_ref/libgsref.so, present only where the reference
sources were available at build time).

TEST INFRASTRUCTURE: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg only.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


class OrGraph(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("nnz", ctypes.c_int64), ("vertex", ctypes.c_void_p), ("edge", ctypes.c_void_p)]


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} not built")
        L = ctypes.CDLL(path)
        for name in ("oracle_count_vertex_rooted",):
            getattr(L, name).argtypes = [ctypes.POINTER(OrGraph), ctypes.c_void_p, ctypes.c_int]
            getattr(L, name).restype = ctypes.c_uint64
        L.oracle_count_vertex_rooted_strided.argtypes = [ctypes.POINTER(OrGraph), ctypes.c_void_p, ctypes.c_int,
                                                         ctypes.c_int32, ctypes.c_int32]
        L.oracle_count_vertex_rooted_strided.restype = ctypes.c_uint64
        for name in ("oracle_count_edge_rooted", "oracle_count_clique"):
            getattr(L, name).argtypes = [ctypes.POINTER(OrGraph), ctypes.c_void_p, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_uint64)]
            getattr(L, name).restype = ctypes.c_uint64
        L.oracle_count_edge_rooted_strided.argtypes = [ctypes.POINTER(OrGraph), ctypes.c_void_p, ctypes.c_int,
                                                       ctypes.c_int64, ctypes.POINTER(ctypes.c_uint64)]
        L.oracle_count_edge_rooted_strided.restype = ctypes.c_uint64
        L.oracle_count_batch_strided.argtypes = [ctypes.POINTER(OrGraph), ctypes.c_void_p, ctypes.c_int32,
                                                 ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_void_p]
        L.oracle_count_batch_strided.restype = None
        L.oracle_count_clique_hybrid.argtypes = [ctypes.POINTER(OrGraph), ctypes.c_void_p, ctypes.c_int,
                                                  ctypes.c_int32, ctypes.POINTER(ctypes.c_uint64)]
        L.oracle_count_clique_hybrid.restype = ctypes.c_uint64
        L.oracle_brute_force.argtypes = [ctypes.POINTER(OrGraph), ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                         ctypes.c_int]
        L.oracle_brute_force.restype = ctypes.c_uint64
        _lib = L
    return _lib


def ref_available():
    """True when the reference's CPU path was built here (checked without
    loading it: the library is only loaded by the tests that call into it)."""
    return os.path.exists(os.path.join(HERE, "_ref", "libgsref.so"))


def ref():
    """The reference's own CPU path, or None when it was not built here."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libgsref.so")
        if not os.path.exists(path):
            return None
        L = ctypes.CDLL(path)
        L.gsref_schedule_dump.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int32, ctypes.c_int64, ctypes.c_int64, ctypes.c_int,
                                          ctypes.c_void_p, ctypes.c_int]
        L.gsref_schedule_dump.restype = ctypes.c_int
        L.gsref_schedule_dump_plain.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int]
        L.gsref_schedule_dump_plain.restype = ctypes.c_int
        L.gsref_pattern_matching.argtypes = [ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_int32, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                                             ctypes.c_int64]
        L.gsref_pattern_matching.restype = ctypes.c_longlong
        L.gsref_clique.argtypes = [ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                   ctypes.c_int]
        L.gsref_clique.restype = ctypes.c_longlong
        L.gsref_motifs.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
        L.gsref_motifs.restype = ctypes.c_int
        _ref = L
    return _ref


class CSR:
    """Keeps numpy arrays alive for an OrGraph view."""

    def __init__(self, vertex, edge):
        self.vertex = np.ascontiguousarray(vertex, dtype=np.int64)
        self.edge = np.ascontiguousarray(edge, dtype=np.int32)
        self.g = OrGraph(len(self.vertex) - 1, len(self.edge), self.vertex.ctypes.data, self.edge.ctypes.data)

    @property
    def n(self):
        return len(self.vertex) - 1


def count_vertex_rooted(csr, desc_ptr, threads=0):
    return lib().oracle_count_vertex_rooted(ctypes.byref(csr.g), ctypes.cast(desc_ptr, ctypes.c_void_p), threads)


def count_vertex_rooted_strided(csr, desc_ptr, threads, stride, offset=0):
    return lib().oracle_count_vertex_rooted_strided(ctypes.byref(csr.g), ctypes.cast(desc_ptr, ctypes.c_void_p),
                                                    threads, stride, offset)


def count_edge_rooted(csr, desc_ptr, threads=0):
    b = ctypes.c_uint64(0)
    s = lib().oracle_count_edge_rooted(ctypes.byref(csr.g), ctypes.cast(desc_ptr, ctypes.c_void_p), threads,
                                       ctypes.byref(b))
    return s, b.value


def count_edge_rooted_strided(csr, desc_ptr, root_stride, threads=0):
    """Device rule over the root edges e with e % root_stride == 0: (raw sum, bytes)."""
    b = ctypes.c_uint64(0)
    s = lib().oracle_count_edge_rooted_strided(ctypes.byref(csr.g), ctypes.cast(desc_ptr, ctypes.c_void_p), threads,
                                               root_stride, ctypes.byref(b))
    return s, b.value


def count_batch_strided(csr, graph_edge_offsets, desc_ptr, root_stride, threads=0):
    """Per-graph raw sums of a batch CSR over its root edges e % root_stride == 0."""
    off = np.ascontiguousarray(graph_edge_offsets, dtype=np.int64)
    out = np.zeros(len(off) - 1, dtype=np.uint64)
    lib().oracle_count_batch_strided(ctypes.byref(csr.g), off.ctypes.data, len(off) - 1,
                                     ctypes.cast(desc_ptr, ctypes.c_void_p), threads, root_stride, out.ctypes.data)
    return [int(x) for x in out]


def count_clique(csr, desc_ptr, threads=0):
    b = ctypes.c_uint64(0)
    s = lib().oracle_count_clique(ctypes.byref(csr.g), ctypes.cast(desc_ptr, ctypes.c_void_p), threads,
                                  ctypes.byref(b))
    return s, b.value


def count_clique_hybrid(csr, desc_ptr, dmax, threads=0):
    b = ctypes.c_uint64(0)
    s = lib().oracle_count_clique_hybrid(ctypes.byref(csr.g), ctypes.cast(desc_ptr, ctypes.c_void_p), threads, dmax,
                                         ctypes.byref(b))
    return s, b.value


def brute_force(csr, n, adj, induced=False, threads=0):
    a = np.array([1 if ch == "1" else 0 for ch in adj], dtype=np.int32)
    a = np.maximum(a.reshape(n, n), a.reshape(n, n).T).reshape(-1).copy()
    return lib().oracle_brute_force(ctypes.byref(csr.g), n, a.ctypes.data, 1 if induced else 0, threads)


def ref_schedule_dump(n, adj, v_cnt, e_cnt, tri_cnt=0, perf=1, restricts=1, iep=True, vertex_induced=False):
    L = ref()
    buf = (ctypes.c_int32 * 4096)()
    k = L.gsref_schedule_dump(n, adj.encode(), perf, restricts, 1 if iep else 0, v_cnt, e_cnt, tri_cnt,
                              1 if vertex_induced else 0, buf, 4096)
    if k < 0:
        raise RuntimeError(f"reference schedule dump failed ({k})")
    return list(buf[:k])


def ref_pattern_matching(csr, max_isect, n, adj, iep=True, tri_cnt=0):
    return ref().gsref_pattern_matching(csr.n, len(csr.edge), csr.vertex.ctypes.data, csr.edge.ctypes.data,
                                        max_isect, n, adj.encode(), 1 if iep else 0, tri_cnt)


def ref_clique(csr, max_isect, k):
    return ref().gsref_clique(csr.n, len(csr.edge), csr.vertex.ctypes.data, csr.edge.ctypes.data, max_isect, k)
