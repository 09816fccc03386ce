"""The CPU oracle pinned against known answers, closed forms and brute force.

The oracle (oracle/gs_oracle.c) is what the GPU is compared with, so it is
pinned here first:
  * wiki-vote: 608389 triangles (reference include/dataloader.h:32) and
    2077903 4-cliques (independent brute force);
  * complete graphs: C(n,k) cliques, n!/((n-k)!|Aut|) copies of any pattern;
  * the six 4-vertex motifs against degree/common-neighbour closed forms;
  * a schedule-free brute force on seeded small graphs;
  * the reference's own Graph::pattern_matching, built from its sources
    (oracle/_ref), where that build exists;
and a planted mistake shows the checks have teeth.
"""
import ctypes
import math

import numpy as np
import pytest

from graphindor_amd import native as N
from oracle import oracle as O
from tests.common import (MOTIFS4, PATTERNS, automorphisms, clique_adj, complete_graph, copies_in_complete,
                          random_graph, reduce_for_clique)


def schedule_for(csr, n, adj, tri=0, **kw):
    return N.Schedule.build(n, adj, csr.n, len(csr.edge), tri, **kw)


def all_rules(csr, n, adj, tri=0, **kw):
    s = schedule_for(csr, n, adj, tri, **kw)
    d = s.plan_desc()
    r = s.redundancy
    return O.count_vertex_rooted(csr, d) // r, O.count_edge_rooted(csr, d)[0] // r


@pytest.fixture(scope="module")
def wiki(wiki_vote_path):
    g = N.Graph.load(wiki_vote_path)
    v, e = g.csr()
    return O.CSR(v.copy(), e.copy())


def test_wiki_vote_triangles(wiki):
    assert all_rules(wiki, 3, PATTERNS["triangle"][1]) == (608389, 608389)
    assert O.brute_force(wiki, 3, PATTERNS["triangle"][1]) == 608389


def test_wiki_vote_four_cliques(wiki):
    assert all_rules(wiki, 4, clique_adj(4)) == (2077903, 2077903)


def test_wiki_vote_clique_rule(wiki):
    v, e = reduce_for_clique(wiki.vertex, wiki.edge)
    red = O.CSR(v, e)
    for k, want in ((3, 608389), (4, 2077903)):
        s = N.Schedule.plain(k, clique_adj(k))
        s.redundancy = 1
        assert O.count_clique(red, s.plan_desc(N.GM_MODE_CLIQUE))[0] == want


@pytest.mark.parametrize("name", ["triangle", "k4", "diamond", "c4", "house", "c5", "k4_pendant", "k5", "c6", "prism"])
def test_complete_graph_closed_form(name):
    n, adj = PATTERNS[name]
    v, e = complete_graph(9)
    csr = O.CSR(v, e)
    want = copies_in_complete(9, n, adj)
    assert all_rules(csr, n, adj) == (want, want)
    assert O.brute_force(csr, n, adj) == want


def test_cliques_in_complete_graph():
    v, e = complete_graph(10)
    red = O.CSR(*reduce_for_clique(v, e))
    for k in range(3, 8):
        s = N.Schedule.plain(k, clique_adj(k))
        s.redundancy = 1
        assert O.count_clique(red, s.plan_desc(N.GM_MODE_CLIQUE))[0] == math.comb(10, k)


def motif_closed_forms(vertex, edge):
    """Edge-induced counts of the six 4-vertex motifs from degrees, triangles and
    common neighbours (K4 by direct enumeration)."""
    n = len(vertex) - 1
    A = np.zeros((n, n), dtype=np.int64)
    for u in range(n):
        A[u, edge[vertex[u]:vertex[u + 1]]] = 1
    d = A.sum(1)
    C = A @ A  # common neighbours
    tri_v = np.diag(A @ A @ A) // 2
    T = tri_v.sum() // 3
    iu = np.triu_indices(n, 1)
    star3 = int(sum(math.comb(int(x), 3) for x in d))
    eu, ev = np.nonzero(np.triu(A, 1))
    path4 = int(((d[eu] - 1) * (d[ev] - 1)).sum() - 3 * T)
    tailed = int((tri_v * (d - 2)).sum())
    c = C[iu]
    c4 = int((c * (c - 1) // 2).sum() // 2)
    diamond = int((C[eu, ev] * (C[eu, ev] - 1) // 2).sum())
    k4 = 0
    for a in range(n):
        for b in range(a + 1, n):
            if not A[a, b]:
                continue
            for cc in range(b + 1, n):
                if A[a, cc] and A[b, cc]:
                    k4 += int((A[a, cc + 1:] & A[b, cc + 1:] & A[cc, cc + 1:]).sum())
    return [k4, diamond, tailed, star3, c4, path4]


@pytest.mark.parametrize("seed", [3, 4])
def test_four_motifs_closed_forms(seed):
    v, e = random_graph(40, 0.2, seed)
    csr = O.CSR(v, e)
    want = motif_closed_forms(v, e)
    got = [all_rules(csr, 4, adj)[0] for adj in MOTIFS4]
    assert got == want
    assert [O.brute_force(csr, 4, adj) for adj in MOTIFS4] == want


@pytest.mark.parametrize("seed", [11, 12])
@pytest.mark.parametrize("name", ["house", "c5", "k4_pendant", "c6", "prism", "tailed_triangle"])
def test_rules_match_brute_force(seed, name):
    n, adj = PATTERNS[name]
    v, e = random_graph(26, 0.3, seed)
    csr = O.CSR(v, e)
    want = O.brute_force(csr, n, adj)
    for iep in (True, False):
        assert all_rules(csr, n, adj, iep=iep) == (want, want)


@pytest.mark.parametrize("name", ["c4", "house", "path4"])
def test_vertex_induced_matches_brute_force(name):
    n, adj = PATTERNS[name]
    v, e = random_graph(24, 0.3, 5)
    csr = O.CSR(v, e)
    want = O.brute_force(csr, n, adj, induced=True)
    got = all_rules(csr, n, adj, vertex_induced=True)
    assert got == (want, want)


def test_k4_pendant_closed_form():
    """K4 plus a pendant edge = sum over 4-cliques of sum over members of (d_v - 3)."""
    v, e = random_graph(30, 0.35, 8)
    n = len(v) - 1
    A = np.zeros((n, n), dtype=np.int64)
    for u in range(n):
        A[u, e[v[u]:v[u + 1]]] = 1
    d = A.sum(1)
    want = 0
    for q in __import__("itertools").combinations(range(n), 4):
        if all(A[a, b] for a, b in __import__("itertools").combinations(q, 2)):
            want += sum(int(d[x]) - 3 for x in q)
    csr = O.CSR(v, e)
    assert all_rules(csr, *PATTERNS["k4_pendant"]) == (want, want)


def test_planted_mistake_is_caught():
    """Dropping the restriction of a plan (a plausible bug) changes the count,
    and the brute force exposes it."""
    n, adj = PATTERNS["house"]
    v, e = random_graph(24, 0.3, 2)
    csr = O.CSR(v, e)
    s = schedule_for(csr, n, adj)
    d = s.plan_desc()
    good = O.count_vertex_rooted(csr, d) // s.redundancy
    assert good == O.brute_force(csr, n, adj)
    bad = N.GmPlanDesc.from_buffer_copy(d.contents)
    rl = (ctypes.c_int32 * n)(*([-1] * n))
    bad.restrict_last = ctypes.cast(rl, ctypes.POINTER(ctypes.c_int32))
    assert O.count_vertex_rooted(csr, ctypes.pointer(bad)) // s.redundancy != good
    # A wrong IEP coefficient is caught as well.
    if d.contents.in_exclusion_optimize_num > 0:
        bad2 = N.GmPlanDesc.from_buffer_copy(d.contents)
        A = bad2.iep_array_size
        coef = (ctypes.c_int32 * A)(*[-x for x in d.contents.iep_coef[:A]])
        bad2.iep_coef = ctypes.cast(coef, ctypes.POINTER(ctypes.c_int32))
        assert O.count_vertex_rooted(csr, ctypes.pointer(bad2)) // s.redundancy != good


def test_automorphism_orders():
    assert [automorphisms(*PATTERNS[k]) for k in ("house", "c5", "k4_pendant", "c6", "prism")] == [2, 10, 6, 12, 12]


# ---- against the reference's own CPU path, where it was built from source ----

ref_missing = O.ref() is None


@pytest.mark.skipif(ref_missing, reason="oracle/_ref not built (reference sources absent)")
@pytest.mark.parametrize("name", ["house", "k4_pendant", "c4", "diamond", "tailed_triangle", "path4"])
def test_reference_counts_agree(name):
    n, adj = PATTERNS[name]
    v, e = random_graph(22, 0.3, 21)
    csr = O.CSR(v, e)
    want = O.brute_force(csr, n, adj)
    assert O.ref_pattern_matching(csr, int(np.diff(v).max()), n, adj) == want
    assert all_rules(csr, n, adj)[0] == want


@pytest.mark.skipif(ref_missing, reason="oracle/_ref not built (reference sources absent)")
def test_reference_clique_test_agrees(wiki):
    assert O.ref_clique(wiki, 773, 3) == 608389


@pytest.mark.skipif(ref_missing, reason="oracle/_ref not built (reference sources absent)")
@pytest.mark.parametrize("name", ["c5", "prism"])
def test_reference_miscounts_are_not_inherited(name):
    """The reference's own count is wrong for these patterns (an invalid
    restriction set for the 5-cycle; a non-integer IEP multiplicity for the
    prism); the exact-mode schedule gives the brute-force count."""
    n, adj = PATTERNS[name]
    v, e = random_graph(22, 0.3, 21)
    csr = O.CSR(v, e)
    want = O.brute_force(csr, n, adj)
    assert O.ref_pattern_matching(csr, int(np.diff(v).max()), n, adj) != want
    assert all_rules(csr, n, adj) == (want, want)


def test_sampled_cpu_baseline_partitions_the_work():
    """The CPU baseline's depth-1 sampling: summed over all offsets it is the
    full count, for any stride."""
    v, e = random_graph(60, 0.3, 3)
    csr = O.CSR(v, e)
    for name in ("house", "k4"):
        n, adj = PATTERNS[name]
        s = schedule_for(csr, n, adj)
        full = O.count_vertex_rooted(csr, s.plan_desc())
        for stride in (2, 7):
            assert sum(O.count_vertex_rooted_strided(csr, s.plan_desc(), 4, stride, o) for o in range(stride)) == full
