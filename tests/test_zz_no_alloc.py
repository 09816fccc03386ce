"""gm_count_static, the count that allocates nothing, against gm_count.

Through the C ABI on the GPU: two calls in a row on each rule (the second on
a smaller graph, so that stale contents of the shared static workspace from
the first call would show), each equal to gm_count's raw sum and byte tally;
and a graph whose prefix slots do not fit the workspace is refused before
anything runs.
"""
import numpy as np
import pytest

from graphindor_amd import native as N
from tests.common import PATTERNS, clique_adj, require_gpu

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    require_gpu()


def both(v, e, sched, mode=N.GM_MODE_PATTERN, **tuning):
    print(f"case: n={len(v) - 1} nnz={len(e)} mode={mode} {tuning}", flush=True)
    dg = N.DeviceGraph(v, e)
    dp = N.DevicePlan(sched.plan_desc(mode))
    want, st = N.count(dg, dp, flags=N.GM_COUNT_ALG_BYTES, **tuning)
    got, st2 = N.count_static(dg, dp, flags=N.GM_COUNT_ALG_BYTES, **tuning)
    assert (got, st2.alg_bytes) == (want, st.alg_bytes)
    again, _ = N.count_static(dg, dp, **tuning)
    assert again == want
    return want


def test_pattern_rule_two_calls_second_smaller(wiki_vote_path):
    g = N.Graph.load(wiki_vote_path)
    v, e = (x.copy() for x in g.csr())
    n, adj = PATTERNS["house"]
    big = both(v, e, N.Schedule.build(n, adj, len(v) - 1, len(e), 608389))
    small_g = N.Graph.rmat(10, 8, 3)
    sv, se = (x.copy() for x in small_g.csr())
    n, adj = PATTERNS["k4_pendant"]
    small = both(sv, se, N.Schedule.build(n, adj, len(sv) - 1, len(se), small_g.info()["tri_cnt"]))
    assert big > 0 and small > 0


def test_clique_rule_two_calls_second_smaller():
    for scale in (14, 10):
        og = N.Graph.rmat(scale, 16, 1).orient_by_degree_rank()
        v, e = (x.copy() for x in og.csr())
        for k in (4, 5):
            s = N.Schedule.plain(k, clique_adj(k))
            s.redundancy = 1
            both(v, e, s, N.GM_MODE_CLIQUE)
            # the list engine inside the static workspace (bitmap engine off)
            both(v, e, s, N.GM_MODE_CLIQUE, bitmap_max_degree=-1)


def test_slots_beyond_the_workspace_are_refused():
    # two hubs sharing 10^6 leaves: second-largest degree 10^6, so one block
    # of four waves needs 4 waves x P prefixes x 10^6 x 4 B, over 64 MiB for
    # any plan with P >= 5 (house has 5)
    leaves = 1_000_000
    n = leaves + 2
    vertex = np.zeros(n + 1, dtype=np.int64)
    deg = np.full(n, 2, dtype=np.int64)
    deg[leaves:] = leaves
    vertex[1:] = np.cumsum(deg)
    edge = np.empty(vertex[-1], dtype=np.int32)
    edge[:2 * leaves:2] = leaves
    edge[1:2 * leaves:2] = leaves + 1
    edge[2 * leaves:3 * leaves] = np.arange(leaves)
    edge[3 * leaves:] = np.arange(leaves)
    dg = N.DeviceGraph(vertex, edge)
    nn, adj = PATTERNS["house"]
    s = N.Schedule.build(nn, adj, n, len(edge), 0)
    with pytest.raises(N.GmError, match="static workspace"):
        N.count_static(dg, N.DevicePlan(s.plan_desc()))
