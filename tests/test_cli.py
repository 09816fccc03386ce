"""gpu_graph / gpu_clique / gpu_mc: the reference's argv and the stdout lines
its harness parses (reproduce/utils.py:7: `Counting time cost: (\\d+\\.\\d+) s`)."""
import os
import re
import subprocess

import pytest

from graphindor_amd import native as N
from oracle import oracle as O
from tests.common import MOTIFS4, PATTERNS, require_gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "graphindor_amd", "lib")
TIME_RE = re.compile(r"Counting time cost: (\d+\.\d+) s")


def run(*args, env=None):
    e = dict(os.environ)
    if env:
        e.update(env)
    return subprocess.run([os.path.join(BIN, args[0]), *args[1:]], capture_output=True, text=True, timeout=300, env=e)


def test_usage_without_arguments():
    for exe in ("gpu_graph", "gpu_mc"):
        r = run(exe)
        assert r.returncode == 1 and "Usage" in r.stderr
    r = run("gpu_clique")
    assert "usage" in r.stdout


@pytest.mark.gpu
def test_gpu_graph_wiki_vote(wiki_vote_path):
    require_gpu()
    r = run("gpu_graph", wiki_vote_path, "3", PATTERNS["triangle"][1])
    assert r.returncode == 0, r.stderr
    assert "Pattern count: 608389" in r.stdout
    assert TIME_RE.search(r.stdout)
    n, adj = PATTERNS["house"]
    r = run("gpu_graph", wiki_vote_path, str(n), adj, "1")
    g = N.Graph.load(wiki_vote_path)
    v, e = g.csr()
    csr = O.CSR(v.copy(), e.copy())
    # The header's tri_cnt drives the schedule, as in the reference.
    s = N.Schedule.build(n, adj, len(v) - 1, len(e), g.info()["tri_cnt"])
    want = O.count_edge_rooted(csr, s.plan_desc())[0] // s.redundancy
    assert f"Pattern count: {want}" in r.stdout


@pytest.mark.gpu
def test_gpu_clique_wiki_vote(wiki_vote_path):
    require_gpu()
    for orient in ("degree", "id"):
        for k, want in ((3, 608389), (4, 2077903)):
            r = run("gpu_clique", wiki_vote_path, str(k), env={"GRAPHSET_CLIQUE_ORIENTATION": orient})
            assert r.returncode == 0, r.stderr
            assert f"count {want}" in r.stdout
            assert TIME_RE.search(r.stdout)


@pytest.mark.gpu
def test_gpu_mc_four_motifs(tmp_path):
    require_gpu()
    g = N.Graph.rmat(11, 16, 5)
    p = str(tmp_path / "rmat11.g")
    g.save(p)
    r = run("gpu_mc", p, "4")
    assert r.returncode == 0, r.stderr
    counts = [int(x) for x in re.findall(r"Pattern count: (\d+)", r.stdout)]
    assert "Motifs number = 6" in r.stdout and TIME_RE.search(r.stdout)
    v, e = g.csr()
    csr = O.CSR(v.copy(), e.copy())
    inf = g.info()
    want = []
    for adj in MOTIFS4:
        s = N.Schedule.build(4, adj, inf["v_cnt"], inf["e_cnt"], inf["tri_cnt"])
        want.append(O.count_edge_rooted(csr, s.plan_desc())[0] // s.redundancy)
    assert counts == want


@pytest.mark.gpu
def test_gpu_graph_with_cost_model(wiki_vote_path, tmp_path):
    """GRAPHSET_COST_MODEL selects the schedule with a cost model (here an
    untrained one: any pick is a valid schedule) and the count is unchanged;
    an unusable file falls back to the analytic pick."""
    require_gpu()
    n, adj = PATTERNS["house"]
    base = run("gpu_graph", wiki_vote_path, str(n), adj)
    assert base.returncode == 0, base.stderr
    want = [l for l in base.stdout.splitlines() if l.startswith("Pattern count:")]
    path = tmp_path / "cm.gscm"
    N.CostModel.new(64, 11).save(path)
    r = run("gpu_graph", wiki_vote_path, str(n), adj, env={"GRAPHSET_COST_MODEL": str(path)})
    assert r.returncode == 0, r.stderr
    assert "Schedule: learned pick" in r.stderr
    assert [l for l in r.stdout.splitlines() if l.startswith("Pattern count:")] == want
    bad = tmp_path / "bad.gscm"
    bad.write_bytes(b"not a model")
    r = run("gpu_graph", wiki_vote_path, str(n), adj, env={"GRAPHSET_COST_MODEL": str(bad)})
    assert r.returncode == 0 and "analytic pick" in r.stderr
    assert [l for l in r.stdout.splitlines() if l.startswith("Pattern count:")] == want
