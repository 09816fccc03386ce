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
