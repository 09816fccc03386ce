"""Learned schedule selection, CPU side: candidate enumeration, the candidate
constructor, the feature encoder and schedule invariance of the counts."""
import numpy as np
import pytest

from graphindor_amd import costmodel as CM
from graphindor_amd import native as N
from oracle import oracle as O
from tests.common import PATTERNS, random_graph

CONFIG5 = {"house": "0101110101010101010011000", "c5": "0100110100010100010110010",
           "k4_pendant": "0111110110110101110010000"}


def connected(n, adj):
    seen, stack = {0}, [0]
    while stack:
        u = stack.pop()
        for v in range(n):
            if adj[u * n + v] == "1" and v not in seen:
                seen.add(v)
                stack.append(v)
    return len(seen) == n


def test_random_patterns_are_seeded_connected_non_cliques():
    a = CM.random_connected_patterns(20, seed=6)
    assert a == CM.random_connected_patterns(20, seed=6)
    assert a != CM.random_connected_patterns(20, seed=7)
    for n, adj in a:
        assert len(adj) == n * n and connected(n, adj)
        assert all(adj[i * n + j] == adj[j * n + i] for i in range(n) for j in range(n))
        assert adj.count("1") < n * (n - 1)


@pytest.mark.parametrize("name", list(CONFIG5))
def test_analytic_argmin_is_the_constructors_choice(name):
    adj = CONFIG5[name]
    for v, e, t in [(1 << 20, 31400000, 2 * 10 ** 8), (1 << 14, 425994, 10 ** 6)]:
        cs = N.schedule_candidates(5, adj, v, e, t)
        best = min(range(len(cs)), key=lambda i: (cs[i].analytic_cost, i))
        assert N.Schedule.from_candidate(5, adj, cs[best]).dump() == N.Schedule.build(5, adj, v, e, t).dump()


def test_every_candidate_breaks_each_automorphism_once():
    for name, adj in CONFIG5.items():
        for c in N.schedule_candidates(5, adj, 1 << 16, 1 << 20, 1 << 18):
            reordered = "".join(adj[c.order[i] * 5 + c.order[j]] for i in range(5) for j in range(5))
            assert N.restrictions_valid(5, reordered, list(c.pairs)), (name, c)


def test_invalid_candidate_is_rejected():
    c = N.schedule_candidates(5, CONFIG5["c5"], 1 << 16, 1 << 20, 1 << 18)[0]
    bad = N.Candidate(c.order, [], c.use_iep, 0.0)
    with pytest.raises(N.GmError):
        N.Schedule.from_candidate(5, CONFIG5["c5"], bad)


@pytest.mark.parametrize("name", ["house", "c5", "k4_pendant", "c6", "prism"])
def test_schedule_invariance_over_candidates(name):
    """Every candidate counts the same on the oracle, equal to brute force
    (the count does not depend on which candidate a selector picks)."""
    n, adj = PATTERNS[name]
    v, e = random_graph(60, 0.18, seed=11)
    csr = O.CSR(v, e)
    want = O.brute_force(csr, n, adj)
    cands = N.schedule_candidates(n, adj, len(v) - 1, len(e), 1000)
    rng = np.random.default_rng(1)
    pick = sorted(set([0, len(cands) - 1] + rng.choice(len(cands), size=min(12, len(cands)), replace=False).tolist()))
    for i in pick:
        s = N.Schedule.from_candidate(n, adj, cands[i])
        raw, _ = O.count_edge_rooted(csr, s.plan_desc())
        assert raw % s.redundancy == 0
        assert raw // s.redundancy == want, (name, cands[i])


def test_encoder_shapes_and_fields():
    g = N.Graph.rmat(10, 8, 3)
    v, e = g.csr()
    inf = g.info()
    x = N.encode_candidates(5, CONFIG5["house"], v, inf["tri_cnt"])
    cs = N.schedule_candidates(5, CONFIG5["house"], inf["v_cnt"], len(e), inf["tri_cnt"])
    assert x.shape == (len(cs), 80) and np.isfinite(x).all()
    best = min(range(len(cs)), key=lambda i: cs[i].analytic_cost)
    assert x[best, 59] == 0.0 and (x[:, 59] >= 0).all()      # log2(cost / cheapest)
    assert np.allclose(x[:, 57], 5 / 7)
    assert np.allclose(x[:, 68:80].sum(axis=1), 1.0)          # degree histogram
    # adjacency bits: house has 6 edges
    assert (x[:, :21].sum(axis=1) == 6).all()


def test_candidates_of_random_patterns():
    """Every random training pattern has candidates, each one valid and
    buildable, whatever the labelling of the pattern."""
    for n, adj in CM.random_connected_patterns(12, seed=6):
        cs = N.schedule_candidates(n, adj, 1 << 16, 1 << 20, 1 << 18)
        assert cs
        for c in cs[:8]:
            reordered = "".join(adj[c.order[i] * n + c.order[j]] for i in range(n) for j in range(n))
            assert N.restrictions_valid(n, reordered, list(c.pairs))
            N.Schedule.from_candidate(n, adj, c)
        g = N.Graph.rmat(9, 4, 1)
        v, e = g.csr()
        x = N.encode_candidates(n, adj, v, g.info()["tri_cnt"])
        assert x.shape == (len(cs), 80)
