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
                          random_graph, reduce_for_clique, root_edge_brute)


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


@pytest.mark.parametrize("name", ["house", "c5", "k4_pendant", "diamond"])
def test_rules_match_brute_force_on_a_hub(name):
    """A hub of degree 300 over leaves of degree ~10: intersections with the
    hub's list are 30x skewed and take the oracle's galloping search."""
    from tests.common import hub_graph
    n, adj = PATTERNS[name]
    v, e = hub_graph(300, 3, extra=0.03)
    csr = O.CSR(v, e)
    want = O.brute_force(csr, n, adj)
    assert want > 0
    assert all_rules(csr, n, adj) == (want, want)


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


def book_closed_form(v, e, pages):
    """Books of `pages` triangles on one edge: sum over edges uv of C(c_uv, pages),
    with c_uv the common neighbours of u and v."""
    n = len(v) - 1
    nb = [set(e[v[u]:v[u + 1]].tolist()) for u in range(n)]
    return sum(math.comb(len(nb[u] & nb[w]), pages) for u in range(n) for w in nb[u] if u < w)


def test_iep_depth_five_plan_and_count():
    """The 7-vertex book has five mutually non-adjacent trailing vertices: IEP
    depth 5 and Bell(6) - Bell(5) = 151 answer entries (every tail subset
    shares one prefix, so the vertex table has one row), within the plan's
    bounds, and both rules count the closed form."""
    n, adj = PATTERNS["book5"]
    v, e = random_graph(18, 0.6, 21)
    csr = O.CSR(v, e)
    s = schedule_for(csr, n, adj)
    hdr = s.dump()[:10]  # n, P, basic, restricts, k, induced, redundancy, I, A, pairs
    assert (hdr[4], hdr[8]) == (5, 151) and 1 <= hdr[7] <= 31
    want = book_closed_form(v, e, 5)
    assert want > 0
    assert all_rules(csr, n, adj) == (want, want)
    assert all_rules(csr, n, adj, iep=False) == (want, want)


def test_degree_rank_orientation():
    """Renumbering by (degree, id) rank keeps one direction per edge, from the
    smaller rank to the larger: every row is sorted and above its own id, the
    out-degrees are those of degree_orientation_init (the renumbering is an
    isomorphism), the edges are the original ones mapped through the rank,
    and the clique counts equal the id-reduced graph's."""
    g = N.Graph.rmat(11, 12, 9)
    v, e = (x.copy() for x in g.csr())
    n = len(v) - 1
    rg, dg = g.orient_by_degree_rank(), g.orient_by_degree()
    rv, re_ = (x.copy() for x in rg.csr())
    dv = dg.csr()[0].copy()
    assert len(re_) * 2 == len(e) and len(rv) == n + 1
    src = np.repeat(np.arange(n), np.diff(rv))
    assert (re_ > src).all()
    assert all((np.diff(re_[rv[u]:rv[u + 1]]) > 0).all() for u in range(n))
    deg = np.diff(v)
    by_rank = np.lexsort((np.arange(n), deg))  # the r-th vertex by (degree, id)
    assert (np.diff(rv) == np.diff(dv)[by_rank]).all()
    rank = np.empty(n, dtype=np.int64)
    rank[by_rank] = np.arange(n)
    osrc = np.repeat(np.arange(n), deg)
    lo, hi = np.minimum(rank[osrc], rank[e]), np.maximum(rank[osrc], rank[e])
    assert sorted(zip(src.tolist(), re_.tolist())) == sorted(set(zip(lo.tolist(), hi.tolist())))
    iv, ie = reduce_for_clique(v, e)
    for k in (3, 4, 5):
        d = N.Schedule.plain(k, clique_adj(k)).plan_desc(N.GM_MODE_CLIQUE)
        assert O.count_clique(O.CSR(rv, re_), d)[0] == O.count_clique(O.CSR(iv, ie), d)[0]


def test_clique_hybrid_rule_equals_clique_rule():
    """The bitmap split (roots of out-degree <= dmax counted on local bitmaps)
    changes neither the count nor, for dmax = 0, the byte tally of the plain
    clique rule; every threshold gives the clique rule's count."""
    v, e = random_graph(90, 0.25, 31)
    ov, oe = reduce_for_clique(v, e)
    csr = O.CSR(ov, oe)
    dmax_all = int(np.diff(ov).max())
    for k in (3, 4, 5, 6):
        s = N.Schedule.plain(k, clique_adj(k))
        d = s.plan_desc(N.GM_MODE_CLIQUE)
        want, wbytes = O.count_clique(csr, d)
        assert O.count_clique_hybrid(csr, d, 0) == (want, wbytes)
        for dmax in (1, 4, 8, 16, dmax_all):
            assert O.count_clique_hybrid(csr, d, dmax)[0] == want
    k4 = O.count_clique(csr, N.Schedule.plain(4, clique_adj(4)).plan_desc(N.GM_MODE_CLIQUE))[0]
    A = np.zeros((90, 90), dtype=np.int64)
    for u in range(90):
        A[u, e[v[u]:v[u + 1]]] = 1
    brute = sum(1 for q in __import__("itertools").combinations(range(90), 4)
                if all(A[a, b] for a, b in __import__("itertools").combinations(q, 2)))
    assert k4 == brute


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


@pytest.mark.parametrize("name", ["house", "c5", "k4_pendant", "tailed_triangle"])
def test_edge_rooted_stride_counts_exactly_the_multiples(name):
    """The strided device rule sums exactly the root edges e with e % S == 0:
    equal to a per-root brute force restated from the plan (IEP off), for
    several strides; stride 1 is the full count."""
    n, adj = PATTERNS[name]
    v, e = random_graph(22, 0.35, 41)
    csr = O.CSR(v, e)
    s = schedule_for(csr, n, adj, iep=False)
    d = s.plan_desc()
    full = O.count_edge_rooted(csr, d)
    assert O.count_edge_rooted_strided(csr, d, 1) == full
    assert full[0] // s.redundancy == O.brute_force(csr, n, adj)
    for S in (2, 5, 13, len(e) + 3):
        got, _ = O.count_edge_rooted_strided(csr, d, S)
        assert got == root_edge_brute(v, e, s, range(0, len(e), S)), S
    # IEP on: the strided tallies of the phases of S = 1 and the full tally agree
    si = schedule_for(csr, n, adj)
    di = si.plan_desc()
    assert O.count_edge_rooted_strided(csr, di, 1) == O.count_edge_rooted(csr, di)


@pytest.mark.parametrize("name", ["house", "c5", "k4_pendant", "c4"])
def test_split_hub_roots_keep_sums(name):
    """Roots whose depth-2 loop exceeds 64 candidates are split over threads;
    the sums equal the unsplit per-root rule (the batch driver) and brute force."""
    from tests.common import hub_graph
    n, adj = PATTERNS[name]
    v, e = hub_graph(400, 7, extra=0.02)
    csr = O.CSR(v, e)
    for iep in (True, False):
        s = schedule_for(csr, n, adj, iep=iep)
        d = s.plan_desc()
        got = O.count_edge_rooted_strided(csr, d, 1)[0]
        assert got == O.count_batch_strided(csr, [0, len(e)], d, 1)[0]
        assert got // s.redundancy == O.brute_force(csr, n, adj)
        for S in (3, 17):
            assert O.count_edge_rooted_strided(csr, d, S)[0] == O.count_batch_strided(csr, [0, len(e)], d, S)[0]


@pytest.mark.parametrize("stride", [1, 3, 64])
def test_batch_strided_per_graph_sums(stride):
    """Per-graph sums of a batch CSR (global ids, graph i owning the root
    edges [eoff[i], eoff[i+1])) over the absolute sample e % stride == 0:
    each equals the per-root brute force of that graph's sampled edges, and
    they add up to the strided count of the whole CSR."""
    g, voff, eoff = N.Graph.batch(6, 5, 5.0, 6.5, 4.0, 10.0)
    v, e = (x.copy() for x in g.csr())
    csr = O.CSR(v, e)
    for name in ("tailed_triangle", "c4", "house"):
        n, adj = PATTERNS[name]
        s = schedule_for(csr, n, adj, iep=False)
        d = s.plan_desc()
        sums = O.count_batch_strided(csr, eoff, d, stride)
        for i in range(len(eoff) - 1):
            first = (eoff[i] + stride - 1) // stride * stride
            assert sums[i] == root_edge_brute(v, e, s, range(first, eoff[i + 1], stride)), (name, i)
        assert sum(sums) == O.count_edge_rooted_strided(csr, d, stride)[0]


def test_automorphism_orders():
    assert [automorphisms(*PATTERNS[k]) for k in ("house", "c5", "k4_pendant", "c6", "prism")] == [2, 10, 6, 12, 12]


# ---- against the reference's own CPU path, where it was built from source ----

ref_missing = not O.ref_available()


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
