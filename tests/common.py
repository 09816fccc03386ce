"""Shared test helpers: patterns, seeded small graphs, closed forms."""
import itertools
import math
import random

import numpy as np
import pytest

from graphindor_amd import native as N

# Row-major adjacency strings.
TRIANGLE = "011101110"
PATTERNS = {
    "triangle": (3, TRIANGLE),
    "wedge": (3, "011100100"),
    "k4": (4, "0111101111011110"),
    "diamond": (4, "0111101111001100"),
    "tailed_triangle": (4, "0111101011001000"),
    "star3": (4, "0111100010001000"),
    "c4": (4, "0110100110010110"),
    "path4": (4, "0110100110000100"),
    "house": (5, "0111010011100011100001100"),
    "house_cfg": (5, "0101110101010101010011000"),
    "c5": (5, "0100110100010100010110010"),
    "k4_pendant": (5, "0111110110110101110010000"),
    "k5": (5, "0111110111110111110111110"),
    "c6": (6, "010001101000010100001010000101100010"),
    "prism": (6, "011100101010110001100011010101001110"),
    # edge 0-1 with five triangles on it (vertices 2..6 joined to 0 and 1):
    # its five mutually non-adjacent tail vertices give an IEP depth of 5
    "book5": (7, "0111111101111111000001100000110000011000001100000"),
}
# MotifGenerator(4) order (src/motif_generator.cpp:5-44).
MOTIFS4 = ["0111101111011110", "0111101111001100", "0111101011001000", "0111100010001000",
           "0110100110010110", "0110100110000100"]


def clique_adj(k):
    return "".join("0" if i == j else "1" for i in range(k) for j in range(k))


def automorphisms(n, adj):
    A = [[adj[i * n + j] == "1" for j in range(n)] for i in range(n)]
    return sum(1 for p in itertools.permutations(range(n))
               if all(A[p[i]][p[j]] == A[i][j] for i in range(n) for j in range(n)))


def csr_from_adjsets(adj):
    n = len(adj)
    vertex = np.zeros(n + 1, dtype=np.int64)
    edges = []
    for i in range(n):
        nb = sorted(adj[i])
        edges += nb
        vertex[i + 1] = vertex[i] + len(nb)
    return vertex, np.array(edges, dtype=np.int32)


def random_graph(n, p, seed):
    rng = random.Random(seed)
    adj = [set() for _ in range(n)]
    for i in range(n):
        for j in range(i + 1, n):
            if rng.random() < p:
                adj[i].add(j)
                adj[j].add(i)
    return csr_from_adjsets(adj)


def complete_graph(n):
    return csr_from_adjsets([set(range(n)) - {i} for i in range(n)])


def hub_graph(n_leaves, seed, extra=0.05):
    """A star with random leaf-leaf edges: one hub of degree n_leaves."""
    rng = random.Random(seed)
    n = n_leaves + 1
    adj = [set() for _ in range(n)]
    for i in range(1, n):
        adj[0].add(i)
        adj[i].add(0)
    for i in range(1, n):
        for j in range(i + 1, n):
            if rng.random() < extra:
                adj[i].add(j)
                adj[j].add(i)
    return csr_from_adjsets(adj)


def copies_in_complete(n_host, k, adj):
    """Edge-induced copies of a k-vertex pattern in K_n: n!/((n-k)! |Aut|)."""
    return math.perm(n_host, k) // automorphisms(k, adj)


def require_gpu():
    if N.device_count() == 0:
        pytest.skip("no GPU visible")


def reduce_for_clique(vertex, edge):
    """Keep N(v) ∩ [0, v) (src/graph.cpp:710-735)."""
    n = len(vertex) - 1
    src = np.repeat(np.arange(n, dtype=np.int64), np.diff(vertex))
    keep = edge < src
    new_edge = edge[keep]
    counts = np.bincount(src[keep], minlength=n)
    new_vertex = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=new_vertex[1:])
    return new_vertex, new_edge.astype(np.int32)


def plan_fields(sched):
    """Named fields of a schedule's dump (layout of gs_schedule_dump)."""
    d = sched.dump()
    n = d[0]
    P = n * (n - 1) // 2 + 1
    I, A, R = d[7], d[8], d[9]
    out = {"n": n, "k": d[4], "redundancy": d[6]}
    off = 10
    for name, L in [("adj", n * n), ("father", P), ("last", n), ("next", P), ("break_size", P), ("loopset", n),
                    ("rlast", n), ("rnext", P), ("rindex", P), ("only", P), ("iep_id", I), ("iep_vflag", I),
                    ("iep_vcoef", I), ("iep_coef", A), ("iep_flag", A), ("iep_pos", A), ("pairs", 2 * R)]:
        out[name] = d[off:off + L]
        off += L
    return out


def root_edge_brute(v, e, sched, edge_ids):
    """Raw sum of the edge-rooted rule over the given root edges, restated
    from the plan alone (no prefixes, no IEP): for root e = (src(e), col[e]),
    the injective maps emb of the plan's vertex order with emb[0] = src(e),
    emb[1] = col[e], every plan edge present and every restriction
    emb[i] < emb[rindex] of depth i satisfied.  For plans built with
    iep=False this is the device's per-root contribution."""
    f = plan_fields(sched)
    n = f["n"]
    assert f["k"] == 0
    adj = [[f["adj"][i * n + j] for j in range(n)] for i in range(n)]
    nb = [set(e[v[u]:v[u + 1]].tolist()) for u in range(len(v) - 1)]
    restr = []
    for i in range(n):
        lst, r = [], f["rlast"][i]
        while r != -1:
            lst.append(f["rindex"][r])
            r = f["rnext"][r]
        restr.append(lst)
    src = np.repeat(np.arange(len(v) - 1), np.diff(v))
    total = 0

    def rec(emb):
        nonlocal total
        d = len(emb)
        if d == n:
            total += 1
            return
        anchors = [j for j in range(d) if adj[j][d]]
        cands = set.intersection(*(nb[emb[j]] for j in anchors)) if anchors else set(range(len(v) - 1))
        for x in cands:
            if x not in emb and all(x < emb[j] for j in restr[d]):
                rec(emb + [x])

    assert adj[0][1], "the plan's first two vertices form the root edge"
    for eid in edge_ids:
        v0, v1 = int(src[eid]), int(e[eid])
        if all(v1 < [v0][j] for j in restr[1]):
            rec([v0, v1])
    return total
