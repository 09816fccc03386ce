"""The .g file format, the generators and the host graph utilities."""
import os
import struct

import numpy as np
import pytest

from graphindor_amd import native as N


def read_header(path):
    with open(path, "rb") as f:
        raw = f.read(32)
    words = struct.unpack("<8I", raw)
    v_cnt, _pad, e_cnt, mi, cs, tri = struct.unpack("<iIqIIq", raw)
    return words, v_cnt, e_cnt, mi, tri


def test_wiki_vote_header(wiki_vote_path):
    words, v, e, mi, tri = read_header(wiki_vote_path)
    assert sum(words) % (1 << 32) == 0
    # The shipped header carries Orkut's triangle constant (include/dataloader.h:31).
    assert (v, e, mi, tri) == (7115, 201524, 773, 627584181)
    g = N.Graph.load(wiki_vote_path)
    vertex, edge = g.csr()
    assert vertex[0] == 0 and vertex[-1] == len(edge) == 201524
    assert all(np.all(np.diff(edge[vertex[i]:vertex[i + 1]]) > 0) for i in range(0, 7115, 97))
    assert N.gs().gs_graph_count_triangles(g.h) == 608389


def test_round_trip(tmp_path, wiki_vote_path):
    g = N.Graph.load(wiki_vote_path)
    p = str(tmp_path / "copy.g")
    g.save(p, max_isect=773)
    with open(wiki_vote_path, "rb") as a, open(p, "rb") as b:
        ra, rb = a.read(), b.read()
    # Same payload; header identical up to the padding word the reference leaves undefined.
    assert ra[32:] == rb[32:]
    _, v, e, mi, tri = read_header(p)
    assert (v, e, mi, tri) == (7115, 201524, 773, 627584181)
    assert sum(struct.unpack("<8I", rb[:32])) % (1 << 32) == 0


def test_bad_files_are_rejected(tmp_path, wiki_vote_path):
    with open(wiki_vote_path, "rb") as f:
        raw = bytearray(f.read())
    bad = tmp_path / "bad.g"
    bad.write_bytes(bytes(raw[:1000]))            # truncated
    with pytest.raises(N.GmError):
        N.Graph.load(str(bad))
    raw2 = bytearray(raw)
    raw2[20] ^= 1                                 # checksum broken
    bad.write_bytes(bytes(raw2))
    with pytest.raises(N.GmError):
        N.Graph.load(str(bad))
    with pytest.raises(N.GmError):
        N.Graph.load(str(tmp_path / "missing.g"))


def check_simple(vertex, edge):
    n = len(vertex) - 1
    src = np.repeat(np.arange(n), np.diff(vertex))
    assert np.all(edge != src)
    for i in range(n):
        assert np.all(np.diff(edge[vertex[i]:vertex[i + 1]]) > 0)
    fwd = set(zip(src.tolist(), edge.tolist()))
    assert all((b, a) in fwd for a, b in fwd)


def test_rmat_is_simple_symmetric_and_deterministic():
    g1 = N.Graph.rmat(10, 16, 1)
    g2 = N.Graph.rmat(10, 16, 1)
    v1, e1 = g1.csr()
    v2, e2 = g2.csr()
    assert np.array_equal(v1, v2) and np.array_equal(e1, e2)
    check_simple(v1, e1)
    g3 = N.Graph.rmat(10, 16, 2)
    assert not np.array_equal(g3.csr()[1][:100], e1[:100])
    # Exact triangle count in the header field: trace(A^3)/6.
    n = len(v1) - 1
    A = np.zeros((n, n), dtype=np.int64)
    for u in range(n):
        A[u, e1[v1[u]:v1[u + 1]]] = 1
    assert g1.info()["tri_cnt"] == int(np.trace(A @ A @ A)) // 6
    assert g1.info()["max_isect"] == int(np.sort(np.diff(v1))[-2])


def test_rmat_scale14_matches_config_statistics():
    g = N.Graph.rmat(14, 16, 1)
    i = g.info()
    assert i["v_cnt"] == 16384
    assert 400000 < i["e_cnt"] < 450000       # ~213K undirected edges after cleaning
    assert i["max_isect"] > 1000              # a heavy power-law tail


def test_erdos_renyi_density():
    n, p = 1 << 14, 32.0 / (1 << 14)
    g = N.Graph.erdos_renyi(n, p, 3)
    v, e = g.csr()
    check_simple(v, e)
    mean = len(e) / n
    assert abs(mean - 32.0) < 0.5


def test_thread_count_does_not_change_the_graph():
    import subprocess, sys
    code = ("import sys; sys.path.insert(0, %r); from graphindor_amd import native as N; import hashlib;"
            "g = N.Graph.rmat(12, 16, 9); v, e = g.csr(); print(hashlib.sha1(e.tobytes()).hexdigest())"
            % os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    outs = set()
    for t in ("1", "5"):
        env = dict(os.environ, OMP_NUM_THREADS=t)
        outs.add(subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True).stdout.strip())
    assert len(outs) == 1


def test_reduce_for_clique_keeps_lower_neighbours():
    from tests.common import reduce_for_clique, random_graph
    v, e = random_graph(30, 0.3, 1)
    g = N.Graph.from_csr(v, e)
    g.reduce_for_clique()
    v2, e2 = g.csr()
    rv, re_ = reduce_for_clique(v, e)
    assert np.array_equal(v2, rv) and np.array_equal(e2, re_)


def test_degree_orientation():
    from tests.common import random_graph
    v, e = random_graph(50, 0.2, 4)
    g = N.Graph.from_csr(v, e)
    d = g.orient_by_degree()
    dv, de = d.csr()
    assert len(de) * 2 == len(e)
    deg = np.diff(v)
    src = np.repeat(np.arange(len(v) - 1), np.diff(dv))
    assert np.all((deg[src] < deg[de]) | ((deg[src] == deg[de]) & (src < de)))
    pairs = set(zip(src.tolist(), de.tolist()))
    full = set(zip(np.repeat(np.arange(len(v) - 1), deg).tolist(), e.tolist()))
    assert {tuple(sorted(p)) for p in pairs} == {tuple(sorted(p)) for p in full}
