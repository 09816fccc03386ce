"""GPU parity: the HIP engine, called through the C ABI, against the oracle.

Bit-exact on the raw uint64 sum (before the redundancy division) and on the
algorithmic byte tally, for the pattern rule (gpu_graph / gpu_mc) and the
clique rule (gpu_clique), on seeded graphs the oracle finishes in seconds;
known answers on wiki-vote; closed forms on complete graphs; edge cases
(empty graph, isolated vertices, a hub of high degree, a graph whose hub
exceeds one 64-lane chunk many times over).
"""
import numpy as np
import pytest

from graphindor_amd import native as N
from oracle import oracle as O
from tests.common import (MOTIFS4, PATTERNS, clique_adj, complete_graph, copies_in_complete, csr_from_adjsets, hub_graph,
                          random_graph, reduce_for_clique, require_gpu)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    require_gpu()


BITMAP_DMAX = 1024  # the device's default local-bitmap bound (kBmMaxD)


def gpu_vs_oracle(v, e, sched, mode=N.GM_MODE_PATTERN, waves_per_cu=0):
    """GPU == oracle on raw sum and byte tally; for clique plans both with the
    local-bitmap engine (default) and without it (list engine only)."""
    print(f"case: n={len(v) - 1} nnz={len(e)} mode={mode}", flush=True)
    desc = sched.plan_desc(mode)
    dg = N.DeviceGraph(v, e)
    dp = N.DevicePlan(desc)
    csr = O.CSR(v, e)
    if mode == N.GM_MODE_CLIQUE:
        raw, st = N.count(dg, dp, flags=N.GM_COUNT_ALG_BYTES, waves_per_cu=waves_per_cu)
        oraw, ob = O.count_clique_hybrid(csr, desc, BITMAP_DMAX)
        assert (raw, st.alg_bytes) == (oraw, ob)
        raw_l, st_l = N.count(dg, dp, flags=N.GM_COUNT_ALG_BYTES, waves_per_cu=waves_per_cu, bitmap_max_degree=-1)
        oraw_l, ob_l = O.count_clique(csr, desc)
        assert (raw_l, st_l.alg_bytes) == (oraw_l, ob_l)
        assert raw_l == raw
        # the bitmap engine's share of the tally and time; none without it
        assert st.bitmap_alg_bytes <= st.alg_bytes
        assert st_l.bitmap_alg_bytes == 0 and st_l.bitmap_ms == 0.0
        assert st.list_ms + st.bitmap_ms <= st.kernel_ms * 1.001 + 0.01
    else:
        raw, st = N.count(dg, dp, flags=N.GM_COUNT_ALG_BYTES, waves_per_cu=waves_per_cu)
        oraw, ob = O.count_edge_rooted(csr, desc)
        assert raw == oraw
        assert st.alg_bytes == ob
    raw2, _ = N.count(dg, dp)          # without the byte tally: same sum
    assert raw2 == raw
    return raw // sched.redundancy


@pytest.fixture(scope="module")
def wiki(wiki_vote_path):
    g = N.Graph.load(wiki_vote_path)
    v, e = g.csr()
    return v.copy(), e.copy()


@pytest.mark.parametrize("name", ["triangle", "k4", "diamond", "k4_pendant", "c4", "tailed_triangle"])
def test_wiki_vote_patterns(wiki, name):
    v, e = wiki
    n, adj = PATTERNS[name]
    s = N.Schedule.build(n, adj, len(v) - 1, len(e), 608389)
    got = gpu_vs_oracle(v, e, s)
    if name == "triangle":
        assert got == 608389
    if name == "k4":
        assert got == 2077903


def test_wiki_vote_cliques(wiki):
    v, e = reduce_for_clique(*wiki)
    for k, want in ((3, 608389), (4, 2077903)):
        s = N.Schedule.plain(k, clique_adj(k))
        s.redundancy = 1
        assert gpu_vs_oracle(v, e, s, N.GM_MODE_CLIQUE) == want
    for k in (5, 6):
        s = N.Schedule.plain(k, clique_adj(k))
        s.redundancy = 1
        gpu_vs_oracle(v, e, s, N.GM_MODE_CLIQUE)


@pytest.mark.parametrize("name", ["house", "c5", "k4_pendant", "c6", "prism", "k5", "path4", "star3"])
@pytest.mark.parametrize("seed", [1, 2])
def test_random_graphs_match_brute_force(name, seed):
    n, adj = PATTERNS[name]
    v, e = random_graph(60, 0.15, seed)
    for iep in (True, False):
        s = N.Schedule.build(n, adj, len(v) - 1, len(e), 0, iep=iep)
        got = gpu_vs_oracle(v, e, s)
        assert got == O.brute_force(O.CSR(v, e), n, adj)


@pytest.mark.parametrize("name", ["c4", "house", "path4"])
def test_vertex_induced(name):
    n, adj = PATTERNS[name]
    v, e = random_graph(40, 0.25, 7)
    s = N.Schedule.build(n, adj, len(v) - 1, len(e), 0, vertex_induced=True)
    assert gpu_vs_oracle(v, e, s) == O.brute_force(O.CSR(v, e), n, adj, induced=True)


def test_motifs_on_rmat():
    g = N.Graph.rmat(11, 16, 5)
    v, e = g.csr()
    v, e = v.copy(), e.copy()
    inf = g.info()
    for adj in MOTIFS4:
        s = N.Schedule.build(4, adj, inf["v_cnt"], inf["e_cnt"], inf["tri_cnt"])
        gpu_vs_oracle(v, e, s)


def test_complete_graph_closed_forms():
    v, e = complete_graph(12)
    for name in ("house", "c5", "k4_pendant", "prism", "k5"):
        n, adj = PATTERNS[name]
        s = N.Schedule.build(n, adj, 12, len(e), 220)
        assert gpu_vs_oracle(v, e, s) == copies_in_complete(12, n, adj)
    rv, re_ = reduce_for_clique(v, e)
    for k in range(3, 8):
        s = N.Schedule.plain(k, clique_adj(k))
        s.redundancy = 1
        assert gpu_vs_oracle(rv, re_, s, N.GM_MODE_CLIQUE) == __import__("math").comb(12, k)


def test_empty_and_isolated():
    n, adj = PATTERNS["house"]
    v = np.zeros(11, dtype=np.int64)
    e = np.zeros(0, dtype=np.int32)
    s = N.Schedule.build(n, adj, 10, 0, 0)
    assert gpu_vs_oracle(v, e, s) == 0
    # a triangle plus isolated vertices
    v2 = np.array([0, 0, 2, 2, 4, 6, 6], dtype=np.int64)
    e2 = np.array([3, 4, 1, 4, 1, 3], dtype=np.int32)
    s3 = N.Schedule.build(3, PATTERNS["triangle"][1], 6, 6, 1)
    assert gpu_vs_oracle(v2, e2, s3) == 1


def test_hub_larger_than_a_chunk():
    """One vertex of degree 3000: its root edges span many 64-edge chunks and
    its neighbour list many 64-lane passes."""
    v, e = hub_graph(3000, 3, extra=0.002)
    for name in ("triangle", "diamond", "house", "tailed_triangle"):
        n, adj = PATTERNS[name]
        s = N.Schedule.build(n, adj, len(v) - 1, len(e), 0)
        got = gpu_vs_oracle(v, e, s)
        if name == "triangle":
            assert got == O.brute_force(O.CSR(v, e), n, adj)
    rv, re_ = reduce_for_clique(v, e)
    for k in (3, 4):
        s = N.Schedule.plain(k, clique_adj(k))
        s.redundancy = 1
        gpu_vs_oracle(rv, re_, s, N.GM_MODE_CLIQUE)


def test_tail_slices_of_hub_roots_keep_counts_and_bytes():
    """The lane tail at depth 2 (house) splits and donates a hub root's
    candidates in slices, like the DFS loop: the oracle's sums and byte
    tallies hold with donation off, lazy (default) and eager, and the eager
    runs did push slices."""
    # Two adjacent hubs with 5000 common neighbours (so N(h0) and N(h0) ∩ N(h1)
    # are both long), plus sparse leaf-leaf edges; the hubs carry the
    # highest ids in one labelling and the lowest in the other, since the
    # depth-1 restriction v1 < v0 decides which roots start at a hub.
    rng = np.random.default_rng(7)
    nl = 5000
    nv = nl + 2
    adj = [set() for _ in range(nv)]
    for h in (nl, nl + 1):
        for x in range(nl):
            adj[h].add(x)
            adj[x].add(h)
    adj[nl].add(nl + 1)
    adj[nl + 1].add(nl)
    for a, b in rng.integers(0, nl, size=(6000, 2)):
        if a != b:
            adj[a].add(int(b))
            adj[b].add(int(a))
    adj_r = [set(nv - 1 - x for x in adj[nv - 1 - a]) for a in range(nv)]
    pushed = False
    for v, e in (csr_from_adjsets(adj), csr_from_adjsets(adj_r)):
        csr = O.CSR(v, e)
        dg = N.DeviceGraph(v, e)
        for name in ("house", "tailed_triangle"):
            n, padj = PATTERNS[name]
            s = N.Schedule.build(n, padj, nv, len(e), 0)
            desc = s.plan_desc()
            dp = N.DevicePlan(desc)
            oraw, ob = O.count_edge_rooted(csr, desc)
            for sl, eager in ((-1, 0), (0, 0), (0, N.GM_COUNT_EAGER_DONATION), (1, N.GM_COUNT_EAGER_DONATION)):
                print(f"case: two hubs of 5000 {name}, hubs first {v[1] - v[0] > nl}, slice_len {sl}, eager {eager}",
                      flush=True)
                raw, st = N.count(dg, dp, flags=N.GM_COUNT_ALG_BYTES | eager, slice_len=sl)
                assert (raw, st.alg_bytes) == (oraw, ob), (name, sl, eager)
                if eager and name == "house":
                    pushed |= st.slices > 0
    assert pushed


def test_deep_donation_is_refused(wiki):
    """GM_COUNT_DEEP_DONATION is switched off: counts and batches refuse it
    with GM_ERR_UNSUPPORTED instead of running."""
    v, e = wiki
    n, adj = PATTERNS["house"]
    dg = N.DeviceGraph(v, e)
    dp = N.DevicePlan(N.Schedule.build(n, adj, len(v) - 1, len(e), 608389).plan_desc())
    with pytest.raises(N.GmError, match="switched off"):
        N.count(dg, dp, flags=N.GM_COUNT_DEEP_DONATION)
    with pytest.raises(N.GmError, match="switched off"):
        N.Batch(dg, [0, len(e)], [dp], flags=N.GM_COUNT_DEEP_DONATION)
    assert N.count(dg, dp)[0] == N.count(dg, dp, slice_len=-1)[0]


@pytest.mark.skip(reason="GM_COUNT_DEEP_DONATION is switched off (gm_count refuses it): with the 16-wave build "
                         "this test printed nothing for 180 s on the GPU and was stopped, and the 5-cycle sample on "
                         "R-MAT 2^22 stalled with depth-3 donation on; the cause is not shown")
def test_depth3_slices_keep_counts_and_bytes():
    """With GM_COUNT_DEEP_DONATION, below one depth-2 candidate a long
    depth-3 loop set (the DFS loop, or the lane tail when it sits at depth 3)
    is split and donated too: the
    popper rebuilds that candidate's prefixes without charging them, so sums
    equal the oracle's and byte tallies those of the same count without
    donation (the oracle's, with the lane tail on, whose byte model it
    follows), with donation off and eager at slice lengths 1 and 8, lane
    tail on and off; and depth-3 slices were pushed."""
    # Three mutually adjacent hubs over 2000 shared leaves: a root edge
    # (hub, hub) has the third hub among its depth-2 candidates, whose
    # neighbour list makes the depth-3 loop set long.
    rng = np.random.default_rng(11)
    nl = 2000
    hubs = (nl, nl + 1, nl + 2)
    adj = [set() for _ in range(nl + 3)]
    for h in hubs:
        for x in range(nl):
            adj[h].add(x)
            adj[x].add(h)
        for h2 in hubs:
            if h2 != h:
                adj[h].add(h2)
    for a, b in rng.integers(0, nl, size=(4000, 2)):
        if a != b:
            adj[a].add(int(b))
            adj[b].add(int(a))
    v, e = csr_from_adjsets(adj)
    csr = O.CSR(v, e)
    dg = N.DeviceGraph(v, e)
    deep = 0
    for name in ("house", "c5", "k4_pendant"):
        n, padj = PATTERNS[name]
        s = N.Schedule.build(n, padj, len(v) - 1, len(e), 0)
        desc = s.plan_desc()
        dp = N.DevicePlan(desc)
        want = O.count_edge_rooted(csr, desc)
        for tail in (0, -1):
            ref = None
            for sl, eager in ((-1, 0), (1, N.GM_COUNT_EAGER_DONATION), (8, N.GM_COUNT_EAGER_DONATION)):
                print(f"case: three hubs {name} lane_tail {tail} slice_len {sl} eager {eager}", flush=True)
                raw, st = N.count(dg, dp, flags=N.GM_COUNT_ALG_BYTES | N.GM_COUNT_DEEP_DONATION | eager, slice_len=sl,
                                  lane_tail=tail)
                ref = ref or (raw, st.alg_bytes)
                assert raw == want[0], (name, tail, sl, eager)
                assert st.alg_bytes == ref[1], (name, tail, sl, eager)
                if tail == 0:
                    assert st.alg_bytes == want[1], (name, sl, eager)
                deep += st.deep_donations
    assert deep > 0


@pytest.mark.parametrize("waves", [1, 4, 16])
def test_grid_size_does_not_change_counts(wiki, waves):
    v, e = wiki
    n, adj = PATTERNS["house"]
    s = N.Schedule.build(n, adj, len(v) - 1, len(e), 608389)
    gpu_vs_oracle(v, e, s, waves_per_cu=waves)


def test_rmat_scale14_triangles_and_cliques():
    """The smallest benchmark graph (R-MAT 2^14, seed 1) through the clique path."""
    g = N.Graph.rmat(14, 16, 1)
    inf = g.info()
    g.reduce_for_clique()
    v, e = g.csr()
    v, e = v.copy(), e.copy()
    for k in (3, 4, 5):
        s = N.Schedule.plain(k, clique_adj(k))
        s.redundancy = 1
        got = gpu_vs_oracle(v, e, s, N.GM_MODE_CLIQUE)
        if k == 3:
            assert got == inf["tri_cnt"]


def test_reference_mode_plans_match_oracle():
    """Plans carrying the reference's accumulated IEP rows are executed as
    written (same raw sum as the oracle following the same plan)."""
    g = N.Graph.rmat(12, 16, 3)
    v, e = g.csr()
    v, e = v.copy(), e.copy()
    N.set_restrict_validation("reference")
    try:
        for name in ("house", "k4_pendant", "c5"):
            n, adj = PATTERNS[name]
            s = N.Schedule.build(n, adj, len(v) - 1, len(e), g.info()["tri_cnt"])
            gpu_vs_oracle(v, e, s)
    finally:
        N.set_restrict_validation("exact")


def oriented(g, how):
    """One direction per edge: by (degree, id) ("degree"), by (degree, id) with
    the vertices renumbered in that order ("rank": every row above its own id),
    or toward smaller ids ("id", the reference's reduce_edges_for_clique)."""
    if how == "degree":
        d = g.orient_by_degree()
    elif how == "rank":
        d = g.orient_by_degree_rank()
    else:
        v, e = g.csr()
        return reduce_for_clique(v.copy(), e.copy())
    v, e = d.csr()
    return v.copy(), e.copy()


def test_degree_oriented_cliques():
    """The clique rule on degree-oriented graphs (reference
    degree_orientation_init, with and without renumbering by rank) counts
    what the id-oriented reduction counts."""
    for scale, seed in ((12, 4), (14, 1)):
        g = N.Graph.rmat(scale, 16, seed)
        graphs = [oriented(g, how) for how in ("degree", "rank", "id")]
        for k in (3, 4, 5, 6):
            s = N.Schedule.plain(k, clique_adj(k))
            s.redundancy = 1
            got = [gpu_vs_oracle(v, e, s, N.GM_MODE_CLIQUE) for v, e in graphs]
            assert got[0] == got[1] == got[2]


def test_clique_tail_with_large_loop_sets():
    """Two apex vertices over 2100 sparse leaves: the root edge between the
    apexes has a loop set larger than the LDS staging arena, which takes the
    wave-wide fallback; counts and bytes still match the oracle."""
    import random as _r
    rng = _r.Random(9)
    leaves = 2100
    adj = [set() for _ in range(leaves + 2)]
    for a in range(leaves):
        for b in range(a + 1, min(leaves, a + 40)):
            if rng.random() < 0.1:
                adj[a].add(b)
                adj[b].add(a)
    for apex in (leaves, leaves + 1):
        for a in range(leaves):
            adj[apex].add(a)
            adj[a].add(apex)
    adj[leaves].add(leaves + 1)
    adj[leaves + 1].add(leaves)
    from tests.common import csr_from_adjsets
    v, e = csr_from_adjsets(adj)
    rv, re_ = reduce_for_clique(v, e)
    for k in (4, 5):
        s = N.Schedule.plain(k, clique_adj(k))
        s.redundancy = 1
        gpu_vs_oracle(rv, re_, s, N.GM_MODE_CLIQUE)


@pytest.mark.parametrize("how", ["degree", "rank", "id"])
@pytest.mark.parametrize("dmax", [16, 64, 200, 1024])
def test_bitmap_threshold_does_not_change_counts(dmax, how):
    """Roots split between the local-bitmap and list engines at several
    out-degree bounds (word counts 1, 4 and 16 all exercised), on rows that
    are triangular in either direction or not at all."""
    g = N.Graph.rmat(13, 24, 6)
    v, e = oriented(g, how)
    dg = N.DeviceGraph(v, e)
    csr = O.CSR(v, e)
    for k in (3, 4, 5, 6, 7):
        s = N.Schedule.plain(k, clique_adj(k))
        s.redundancy = 1
        desc = s.plan_desc(N.GM_MODE_CLIQUE)
        want = O.count_clique_hybrid(csr, desc, dmax)
        # bitmap roots above the block degree are counted by a whole block
        # each (default 192, 8: almost all, -1: none)
        for block in (0, 8, -1):
            raw, st = N.count(dg, N.DevicePlan(desc), flags=N.GM_COUNT_ALG_BYTES, bitmap_max_degree=dmax,
                              bitmap_block_degree=block)
            assert (raw, st.alg_bytes) == want, (k, block)


@pytest.mark.parametrize("how", ["degree", "rank", "id"])
def test_bitmap_wide_rows(how):
    """Out-degrees between 256 and 1024 after orientation: rows of up to 16
    words held in global scratch."""
    v, e = random_graph(700, 0.5, 12)
    dv, de = oriented(N.Graph.from_csr(v, e), how)
    assert np.diff(dv).max() > 256
    dg = N.DeviceGraph(dv, de)
    for k in (3, 4, 5):
        s = N.Schedule.plain(k, clique_adj(k))
        s.redundancy = 1
        want = gpu_vs_oracle(dv, de, s, N.GM_MODE_CLIQUE)  # wide roots by a block each, rows in LDS
        # the same roots one wave each, rows in global scratch
        raw, _ = N.count(dg, N.DevicePlan(s.plan_desc(N.GM_MODE_CLIQUE)), bitmap_block_degree=-1)
        assert raw == want


@pytest.mark.parametrize("name", ["house", "house_cfg", "c5", "k4_pendant", "c6", "prism", "diamond", "tailed_triangle"])
def test_lane_tail_matches_wave_path(name):
    """The lane-parallel last candidate level gives the wave-wide path's sum.
    R-MAT 2^10 with edge factor 4 keeps the oracle to seconds for every
    pattern (on R-MAT 2^12 ef16 a 6-cycle is ~10 CPU-hours of set work)."""
    n, adj = PATTERNS[name]
    g = N.Graph.rmat(10, 4, 8)
    v, e = g.csr()
    v, e = v.copy(), e.copy()
    inf = g.info()
    for iep in (True, False):
        s = N.Schedule.build(n, adj, inf["v_cnt"], inf["e_cnt"], inf["tri_cnt"], iep=iep)
        dg = N.DeviceGraph(v, e)
        dp = N.DevicePlan(s.plan_desc())
        on, _ = N.count(dg, dp)
        off, _ = N.count(dg, dp, lane_tail=-1)
        assert on == off
        gpu_vs_oracle(v, e, s)


@pytest.mark.parametrize("name", ["house", "c5", "tailed_triangle"])
def test_lane_tail_unstaged_fathers(name):
    """Fathers longer than the LDS stage (a hub of degree 2500) are read from
    HBM by the lane-parallel tail; same sum as the wave path and the oracle."""
    n, adj = PATTERNS[name]
    v, e = hub_graph(2500, 5, extra=0.002)
    s = N.Schedule.build(n, adj, len(v) - 1, len(e), 0)
    dg = N.DeviceGraph(v, e)
    dp = N.DevicePlan(s.plan_desc())
    on, _ = N.count(dg, dp)
    off, _ = N.count(dg, dp, lane_tail=-1)
    assert on == off
    gpu_vs_oracle(v, e, s)


def test_count_multi_matches_single_plans():
    """All six 4-motifs in one gm_count_multi call equal six gm_count calls."""
    g = N.Graph.rmat(12, 16, 9)
    v, e = g.csr()
    v, e = v.copy(), e.copy()
    inf = g.info()
    dg = N.DeviceGraph(v, e)
    scheds = [N.Schedule.build(4, adj, inf["v_cnt"], inf["e_cnt"], inf["tri_cnt"]) for adj in MOTIFS4]
    plans = [N.DevicePlan(s.plan_desc()) for s in scheds]
    multi, st = N.count_multi(dg, plans, flags=N.GM_COUNT_ALG_BYTES)
    single = [N.count(dg, p)[0] for p in plans]
    assert multi == single
    csr = O.CSR(v, e)
    assert multi == [O.count_edge_rooted(csr, s.plan_desc())[0] for s in scheds]
    # A clique plan mixed in goes through its own launch.
    clique = N.Schedule.plain(4, clique_adj(4))
    clique.redundancy = 1
    d = N.Graph.from_csr(v, e).orient_by_degree()
    dv, de = d.csr()
    dgo = N.DeviceGraph(dv.copy(), de.copy())
    mixed, _ = N.count_multi(dgo, [N.DevicePlan(clique.plan_desc(N.GM_MODE_CLIQUE))])
    assert mixed[0] == multi[0] // scheds[0].redundancy  # K4 count, either rule


@pytest.mark.parametrize("chunk_edges", [0, 64, 512])
def test_count_batch_per_graph_sums(chunk_edges):
    """A batch of seeded graphs as one CSR: per-graph, per-motif sums equal
    counting each graph on its own (and the oracle)."""
    g, voff, eoff = N.Graph.batch(12, 5, 8.0, 11.0, 4.0, 24.0)
    v, e = g.csr()
    v, e = v.copy(), e.copy()
    inf = g.info()
    dg = N.DeviceGraph(v, e)
    scheds = [N.Schedule.build(4, adj, inf["v_cnt"], inf["e_cnt"], inf["tri_cnt"]) for adj in MOTIFS4]
    plans = [N.DevicePlan(s.plan_desc()) for s in scheds]
    sums, st = N.count_batch(dg, eoff, plans, flags=N.GM_COUNT_ALG_BYTES, chunk_edges=chunk_edges)
    assert len(sums) == 12
    for i in range(12):
        lv = v[voff[i]:voff[i + 1] + 1] - v[voff[i]]
        le = e[eoff[i]:eoff[i + 1]] - voff[i]
        csr = O.CSR(lv, le)
        want = [O.count_edge_rooted(csr, s.plan_desc())[0] for s in scheds]
        assert sums[i] == want, i
    # the union graph's per-plan totals equal the per-graph sums added up
    whole, _ = N.count_multi(dg, plans)
    assert whole == [sum(sums[i][j] for i in range(12)) for j in range(6)]


def test_graph_captured_batch_replays():
    """The hipGraph batch: several replays, with eager counts and syncs in
    between, each equal to the eager gm_count_batch result."""
    g, voff, eoff = N.Graph.batch(10, 40, 8.0, 11.0, 4.0, 24.0)
    v, e = g.csr()
    v, e = v.copy(), e.copy()
    inf = g.info()
    dg = N.DeviceGraph(v, e)
    scheds = [N.Schedule.build(4, adj, inf["v_cnt"], inf["e_cnt"], inf["tri_cnt"]) for adj in MOTIFS4]
    plans = [N.DevicePlan(s.plan_desc()) for s in scheds]
    eager, _ = N.count_batch(dg, eoff, plans)
    b = N.Batch(dg, eoff, plans)
    for _ in range(4):
        got, st = b.run()
        assert got == eager
        again, _ = N.count_batch(dg, eoff, plans)  # eager work between replays
        assert again == eager
    b.free()
    # a root-edge sub-range: graphs 2..5 only
    sub, _ = N.count_batch(dg, eoff[2:7], plans)
    assert sub == eager[2:6]
    # root sampling: the captured batch, the eager batch and per-graph
    # launches count the same sample
    sampled, _ = N.count_batch(dg, eoff, plans, root_stride=5)
    bs = N.Batch(dg, eoff, plans, root_stride=5)
    for _ in range(2):
        assert bs.run()[0] == sampled
    bs.free()
    for i in range(10):
        assert N.count_batch(dg, eoff[i:i + 2], plans, root_stride=5)[0][0] == sampled[i]


def test_autotuner_choice_keeps_counts(tmp_path):
    """The autotuner times its candidates on a sample of the batch, caches the
    choice per key, and the chosen tuning gives the default tuning's sums."""
    g, voff, eoff = N.Graph.batch(24, 77, 8.0, 11.0, 4.0, 24.0)
    v, e = g.csr()
    v, e = v.copy(), e.copy()
    inf = g.info()
    dg = N.DeviceGraph(v, e)
    scheds = [N.Schedule.build(4, adj, inf["v_cnt"], inf["e_cnt"], inf["tri_cnt"]) for adj in MOTIFS4]
    plans = [N.DevicePlan(s.plan_desc()) for s in scheds]
    cache = str(tmp_path / "tune.txt")
    key = "motifs4:" + N.degree_class(v)
    first = N.autotune_batch(dg, eoff, plans, key, cache)
    assert not first["from_cache"] and first["tried"] == 18
    again = N.autotune_batch(dg, eoff, plans, key, cache)
    assert again["from_cache"] and again["waves_per_cu"] == first["waves_per_cu"]
    want, _ = N.count_batch(dg, eoff, plans)
    got, _ = N.count_batch(dg, eoff, plans, waves_per_cu=first["waves_per_cu"], chunk_edges=first["chunk_edges"],
                           lane_tail=first["lane_tail"])
    assert got == want


@pytest.mark.parametrize("stride", [2, 7, 64])
def test_root_stride_counts_exactly_the_multiples(stride):
    """Root-edge sampling counts exactly the edges e with e % stride == 0:
    equal to the sum of single-edge ranges over those edges, and unchanged
    when the range is split into pieces."""
    g = N.Graph.rmat(9, 4, 21)
    v, e = g.csr()
    v, e = v.copy(), e.copy()
    inf = g.info()
    dg = N.DeviceGraph(v, e)
    s = N.Schedule.build(5, PATTERNS["house"][1], inf["v_cnt"], inf["e_cnt"], inf["tri_cnt"])
    dp = N.DevicePlan(s.plan_desc())
    nnz = len(e)
    sampled, _ = N.count_batch(dg, [0, nnz], [dp], root_stride=stride)
    singles = sum(N.count_batch(dg, [x, x + 1], [dp])[0][0][0] for x in range(0, nnz, stride))
    assert sampled[0][0] == singles
    cuts = [0, nnz // 3 + 1, nnz // 2 + 5, nnz]
    pieces, _ = N.count_batch(dg, cuts, [dp], root_stride=stride)
    assert sum(p[0] for p in pieces) == sampled[0][0]
    full, _ = N.count(dg, dp)
    assert N.count_batch(dg, [0, nnz], [dp], root_stride=1)[0][0][0] == full


@pytest.mark.parametrize("name", ["house", "c5", "k4_pendant", "c6"])
def test_slice_donation_keeps_sums_and_bytes(name):
    """Depth-2 slices donated to idle waves (slice lengths 1, 4 and the
    default, lazily and eagerly; eager slices of 1 overflow the 2^18-slot
    queue, so the kept-locally path runs too) give the undonated sum and byte
    tally, equal to the oracle's."""
    n, adj = PATTERNS[name]
    g = N.Graph.rmat(10, 8, 17)
    v, e = g.csr()
    v, e = v.copy(), e.copy()
    inf = g.info()
    s = N.Schedule.build(n, adj, inf["v_cnt"], inf["e_cnt"], inf["tri_cnt"])
    dg = N.DeviceGraph(v, e)
    dp = N.DevicePlan(s.plan_desc())
    # Undonated references: with the lane-parallel tail (the oracle's byte
    # model charges the tail depth's builds as size-only) and without it.
    oraw, ob = O.count_edge_rooted(O.CSR(v, e), s.plan_desc())
    ref = {}
    for tail in (0, -1):
        got, st = N.count(dg, dp, flags=N.GM_COUNT_ALG_BYTES, slice_len=-1, lane_tail=tail)
        ref[tail] = (got, st.alg_bytes)
    assert ref[0] == (oraw, ob)
    assert ref[-1][0] == oraw
    off = oraw
    for sl in (0, 1, 4, 32):  # 0: the default length
        for tail in (-1, 0):
            print(f"slice_len {sl} lane_tail {tail}", flush=True)
            for eager in (0, N.GM_COUNT_EAGER_DONATION):
                got, st = N.count(dg, dp, flags=N.GM_COUNT_ALG_BYTES | eager, slice_len=sl, lane_tail=tail)
                assert (got, st.alg_bytes) == ref[tail], (sl, tail, eager)
    # captured batch with donation, replayed
    b = N.Batch(dg, [0, len(e)], [dp], flags=N.GM_COUNT_EAGER_DONATION, slice_len=1)
    for _ in range(3):
        assert b.run()[0][0][0] == off
    b.free()


def test_time_limit_fails_cleanly_and_leaves_no_state():
    """A count that passes its time limit fails with GM_ERR_TIMEOUT (also
    with slices waiting in the donation queue), and the next counts on the
    same device are exact."""
    g = N.Graph.rmat(14, 16, 4)
    v, e = g.csr()
    v, e = v.copy(), e.copy()
    inf = g.info()
    dg = N.DeviceGraph(v, e)
    s = N.Schedule.build(5, PATTERNS["c5"][1], inf["v_cnt"], inf["e_cnt"], inf["tri_cnt"])
    dp = N.DevicePlan(s.plan_desc())
    for extra in ({}, {"slice_len": 4}):
        flags = N.GM_COUNT_EAGER_DONATION if extra else 0
        with pytest.raises(N.GmTimeout):
            N.count(dg, dp, flags=flags, time_limit_ms=5, **extra)
    small = N.Graph.rmat(10, 8, 17)
    sv, se = small.csr()
    sv, se = sv.copy(), se.copy()
    sinf = small.info()
    ss = N.Schedule.build(5, PATTERNS["c5"][1], sinf["v_cnt"], sinf["e_cnt"], sinf["tri_cnt"])
    want = O.count_edge_rooted(O.CSR(sv, se), ss.plan_desc())[0]
    sdg, sdp = N.DeviceGraph(sv, se), N.DevicePlan(ss.plan_desc())
    assert N.count(sdg, sdp)[0] == want
    assert N.count(sdg, sdp, flags=N.GM_COUNT_EAGER_DONATION, slice_len=4)[0] == want
    assert N.count(sdg, sdp, time_limit_ms=60000)[0] == want


def test_clique_time_limit_fails_cleanly():
    """The clique rule's list engine also honours the time limit: a count
    that passes it fails with GM_ERR_TIMEOUT, alone and beside a pattern plan
    (whose slice queue must then be left clear), and the next counts are exact."""
    g = N.Graph.rmat(14, 16, 4)
    v, e = (x.copy() for x in g.csr())
    inf = g.info()
    rv, re_ = reduce_for_clique(v, e)
    k6 = N.Schedule.plain(6, clique_adj(6))
    k6.redundancy = 1
    cplan = N.DevicePlan(k6.plan_desc(N.GM_MODE_CLIQUE))
    rdg = N.DeviceGraph(rv, re_)
    with pytest.raises(N.GmTimeout):
        N.count(rdg, cplan, time_limit_ms=1, bitmap_max_degree=-1)
    # a pattern plan and a clique plan in one call, on the symmetric graph
    dg = N.DeviceGraph(v, e)
    c5 = N.Schedule.build(5, PATTERNS["c5"][1], inf["v_cnt"], inf["e_cnt"], inf["tri_cnt"])
    with pytest.raises(N.GmTimeout):
        N.count_multi(dg, [N.DevicePlan(c5.plan_desc()), N.DevicePlan(k6.plan_desc(N.GM_MODE_CLIQUE))],
                      flags=N.GM_COUNT_EAGER_DONATION, time_limit_ms=2, slice_len=4, bitmap_max_degree=-1)
    small = N.Graph.rmat(10, 8, 17)
    sv, se = (x.copy() for x in small.csr())
    srv, sre = reduce_for_clique(sv, se)
    sdesc = k6.plan_desc(N.GM_MODE_CLIQUE)
    assert N.count(N.DeviceGraph(srv, sre), N.DevicePlan(sdesc), bitmap_max_degree=-1)[0] == \
        O.count_clique(O.CSR(srv, sre), sdesc)[0]
    sinf = small.info()
    ss = N.Schedule.build(5, PATTERNS["c5"][1], sinf["v_cnt"], sinf["e_cnt"], sinf["tri_cnt"])
    assert N.count(N.DeviceGraph(sv, se), N.DevicePlan(ss.plan_desc()), flags=N.GM_COUNT_EAGER_DONATION,
                   slice_len=4)[0] == O.count_edge_rooted(O.CSR(sv, se), ss.plan_desc())[0]


@pytest.mark.parametrize("hub", [8, 64, -1])
@pytest.mark.parametrize("name", ["house", "c5", "k4_pendant", "c6", "prism", "tailed_triangle", "c4", "diamond"])
def test_hub_bitmaps_keep_counts_and_bytes(name, hub):
    """Lookups in hub neighbour lists by bitmap (every vertex of degree at
    least `hub`; -1: none) give the oracle's sum and byte tally, and with the
    lane-parallel last level off (whose byte model the oracle does not
    follow) the sum and the tally of the same run without bitmaps.  R-MAT
    2^10 ef4 keeps the oracle to seconds for every pattern."""
    n, adj = PATTERNS[name]
    g = N.Graph.rmat(10, 4, 8)
    v, e = (x.copy() for x in g.csr())
    s = N.Schedule.build(n, adj, len(v) - 1, len(e), g.info()["tri_cnt"])
    desc = s.plan_desc()
    dg, dp = N.DeviceGraph(v, e), N.DevicePlan(desc)
    want = O.count_edge_rooted(O.CSR(v, e), desc)
    print(f"case: {name} hub={hub}", flush=True)
    raw, st = N.count(dg, dp, flags=N.GM_COUNT_ALG_BYTES, hub_degree=hub)
    assert (raw, st.alg_bytes) == want
    assert (st.hub_vertices > 0) == (hub > 0)
    off, st_off = N.count(dg, dp, flags=N.GM_COUNT_ALG_BYTES, hub_degree=hub, lane_tail=-1)
    plain, st_plain = N.count(dg, dp, flags=N.GM_COUNT_ALG_BYTES, hub_degree=-1, lane_tail=-1)
    assert off == plain == want[0]
    assert st_off.alg_bytes == st_plain.alg_bytes


@pytest.mark.parametrize("cap", [-1, 16, 0])
def test_tail_staging_cap_keeps_counts(wiki, cap):
    """The last level's fathers staged in LDS up to lds_set_cap elements (-1:
    never) give the same sums and bytes."""
    v, e = wiki
    for name in ("house", "k4_pendant"):
        n, adj = PATTERNS[name]
        s = N.Schedule.build(n, adj, len(v) - 1, len(e), 608389)
        desc = s.plan_desc()
        raw, st = N.count(N.DeviceGraph(v, e), N.DevicePlan(desc), flags=N.GM_COUNT_ALG_BYTES, lds_set_cap=cap)
        assert (raw, st.alg_bytes) == O.count_edge_rooted(O.CSR(v, e), desc)


def test_hub_bitmaps_in_captured_batch():
    """A captured batch whose graph has hub bitmaps replays the per-graph sums."""
    g, voff, eoff = N.Graph.batch(12, 9, 8.0, 11.0, 4.0, 24.0)
    inf = g.info()
    v, e = (x.copy() for x in g.csr())
    dg = N.DeviceGraph(v, e)
    plans = [N.DevicePlan(N.Schedule.build(4, adj, inf["v_cnt"], inf["e_cnt"], inf["tri_cnt"]).plan_desc())
             for adj in MOTIFS4]
    want, _ = N.count_batch(dg, eoff, plans, hub_degree=-1)
    got, st = N.count_batch(dg, eoff, plans, hub_degree=16)
    assert got == want and st.hub_vertices > 0
    b = N.Batch(dg, eoff, plans, hub_degree=16)
    for _ in range(3):
        assert b.run()[0] == want
    b.free()
