"""Learned schedule cost model: timing harness, training and evaluation.

The schedule compiler's analytic selector (src/schedule_IEP.cpp:307-382)
ranks candidate schedules -- vertex order, restriction set, IEP on/off -- by
formula.  Here the same candidates (Schedule_IEP::enumerate_candidates) are
timed on the GPU on synthetic R-MAT graphs, and an FP32 MLP (HIP training and
inference kernels behind gm_mlp_*, wrapped by the host CostModel) learns to
rank them from the features of CostModel's encoder.  Every candidate is a
validated schedule, so a choice changes the run time only, never the count.

Timing protocol (time_group): all candidates of one (pattern, graph) group
count the same uniform sample of root edges -- every S-th edge of the CSR,
S chosen so that the analytic pick takes about `target_ms` -- and each
candidate's time is the minimum over `reps` runs.  Training targets are
log(ms); the loss is centred per group, so only within-group differences
are learned.

    python -m graphindor_amd.costmodel train --out cost_model.gscm
    python -m graphindor_amd.costmodel eval --model cost_model.gscm --scale 20
"""
import argparse
import json
import math
import sys
import time

import numpy as np

from . import native as N


def random_connected_patterns(count, sizes=(5, 6), seed=6):
    """`count` random connected, non-clique patterns (row-major 0/1 strings),
    alternating over `sizes`, drawn with numpy's PCG64 from `seed`: a random
    spanning tree (each vertex i > 0 joins a uniform earlier vertex) plus each
    other pair with probability 0.35.  Duplicated strings are skipped."""
    rng = np.random.default_rng(seed)
    out, seen = [], set()
    i = 0
    while len(out) < count:
        n = sizes[i % len(sizes)]
        i += 1
        A = np.zeros((n, n), dtype=np.int8)
        for v in range(1, n):
            u = int(rng.integers(0, v))
            A[u, v] = A[v, u] = 1
        for u in range(n):
            for v in range(u + 1, n):
                if not A[u, v] and rng.random() < 0.35:
                    A[u, v] = A[v, u] = 1
        if A.sum() == n * (n - 1):
            continue  # clique: a single trivial schedule
        perm = rng.permutation(n)
        A = A[np.ix_(perm, perm)]
        s = "".join(str(int(x)) for x in A.ravel())
        if s in seen:
            continue
        seen.add(s)
        out.append((n, s))
    return out


def _time_plan(dg, nnz, dp, stride, reps, limit_ms=0):
    """Best of `reps` sampled runs in ms; None if one passes limit_ms (the
    kernel stops itself at the limit, see gm_tuning.time_limit_ms)."""
    best = math.inf
    for _ in range(reps):
        try:
            _, st = N.count_batch(dg, [0, nnz], [dp], root_stride=stride, time_limit_ms=int(limit_ms))
        except N.GmTimeout:
            return None
        best = min(best, st.kernel_ms)
    return best


def time_group(dg, nnz, n, adj, cands, target_ms=4.0, reps=2, max_cands=96, rng=None, log=None, budget_s=30.0,
               skip_ms=1000.0):
    """Times candidates of one pattern on one device graph on a shared root
    sample.  Returns (indices, ms, stride), the analytic pick first, or None
    when the group is skipped.  The sample starts at ~1024 roots; a group
    whose analytic pick already takes more than skip_ms there is skipped (a
    few hub roots dominate it, so its timings would measure them alone).
    Otherwise the sample grows (by 4x) until the analytic pick takes at least
    target_ms / 4, and candidates are timed until `budget_s` is spent."""
    analytic = min(range(len(cands)), key=lambda i: (cands[i].analytic_cost, i))
    idx = list(range(len(cands)))
    if max_cands and len(idx) > max_cands:
        rng = rng or np.random.default_rng(0)
        others = [i for i in idx if i != analytic]
        idx = [analytic] + sorted(rng.choice(others, size=max_cands - 1, replace=False).tolist())
    else:
        idx.remove(analytic)
        idx = [analytic] + idx
    sched = N.Schedule.from_candidate(n, adj, cands[analytic])
    dp = N.DevicePlan(sched.plan_desc())
    stride = max(1, nnz // 1024)
    t0 = _time_plan(dg, nnz, dp, stride, 1, skip_ms)
    if t0 is None:
        if log:
            log(f"    {n} {adj}: skipped, analytic pick over {skip_ms:.0f} ms on {nnz // stride} roots")
        return None
    while t0 < target_ms / 4 and stride > 1:
        stride = max(1, stride // 4)
        t0 = _time_plan(dg, nnz, dp, stride, 1, skip_ms)
        if t0 is None:
            return None
    # A candidate slower than 10x the analytic pick is stopped there and
    # recorded at that bound.
    limit = max(50.0, 10.0 * t0)
    ms, start = [], time.time()
    for i in idx:
        if ms and time.time() - start > budget_s:
            break
        sched = N.Schedule.from_candidate(n, adj, cands[i])
        dp = N.DevicePlan(sched.plan_desc())
        t = _time_plan(dg, nnz, dp, stride, reps, limit)
        ms.append(limit if t is None else t)
    idx = idx[:len(ms)]
    if log:
        log(f"    {n} {adj}: {len(idx)} candidates, stride {stride}, analytic {ms[0]:.3f} ms, "
            f"best {min(ms):.3f} ms, worst {max(ms):.3f} ms")
    return idx, ms, stride


def collect(graph_specs, patterns, target_ms=4.0, reps=2, max_cands=96, seed=6, log=None, budget_s=30.0):
    """Dataset over R-MAT graphs (scale, edge_factor, seed) x patterns."""
    log = log or (lambda *a, **k: print(*a, flush=True))
    rng = np.random.default_rng(seed)
    xs, ts, goff, meta = [], [], [0], []
    for scale, ef, gseed in graph_specs:
        g = N.Graph.rmat(scale, ef, gseed)
        v, e = g.csr()
        v, e = v.copy(), e.copy()
        inf = g.info()
        dg = N.DeviceGraph(v, e)
        log(f"  graph rmat{scale} ef{ef} seed{gseed}: nnz {len(e)}")
        for n, adj in patterns:
            cands = N.schedule_candidates(n, adj, inf["v_cnt"], len(e), inf["tri_cnt"])
            if len(cands) < 2:
                continue
            x = N.encode_candidates(n, adj, v, inf["tri_cnt"])
            got = time_group(dg, len(e), n, adj, cands, target_ms, reps, max_cands, rng, log, budget_s)
            if got is None or len(got[0]) < 2:
                continue
            idx, ms, stride = got
            xs.append(x[idx])
            ts.append(np.log(np.maximum(np.asarray(ms, dtype=np.float64), 1e-4)).astype(np.float32))
            goff.append(goff[-1] + len(idx))
            meta.append({"graph": [scale, ef, gseed], "n": n, "adj": adj, "stride": stride,
                         "analytic_ms": ms[0], "best_ms": float(min(ms))})
        del dg
    return {"x": np.concatenate(xs), "log_ms": np.concatenate(ts), "goff": np.asarray(goff, dtype=np.int32),
            "meta": meta}


def train(ds, hidden=256, steps=3000, lr=1e-3, weight_decay=1e-4, seed=6):
    m = N.CostModel.new(hidden, seed)
    loss = m.train(ds["x"], ds["log_ms"], ds["goff"], steps, lr, weight_decay, fit_norm=True)
    return m, loss


def group_regret(model, ds):
    """Per group: time of the model's pick / best time, and of the analytic pick (row 0)."""
    y = model.predict(ds["x"])
    out = []
    for g in range(len(ds["goff"]) - 1):
        a, b = ds["goff"][g], ds["goff"][g + 1]
        t = np.exp(ds["log_ms"][a:b])
        out.append((float(t[int(np.argmin(y[a:b]))] / t.min()), float(t[0] / t.min())))
    return out


def evaluate(model, scale, ef, seed, patterns, log=print, root_stride=1):
    """Analytic vs learned pick on one R-MAT graph.  root_stride 1: full
    counts, which must agree; S > 1: both schedules timed on the root sample
    e % S == 0 (sampled sums are not counts and differ between schedules)."""
    g = N.Graph.rmat(scale, ef, seed)
    v, e = g.csr()
    v, e = v.copy(), e.copy()
    inf = g.info()
    dg = N.DeviceGraph(v, e)
    rows = []
    for name, n, adj in patterns:
        res = {}
        for kind, mdl in (("analytic", None), ("learned", model)):
            s, info = N.select_schedule(n, adj, v, inf["tri_cnt"], mdl)
            dp = N.DevicePlan(s.plan_desc())
            raw, st = N.count_batch(dg, [0, len(e)], [dp], root_stride=root_stride)
            res[kind] = {"raw": raw[0][0], "redundancy": s.redundancy, "ms": st.kernel_ms, "choice": info}
        if root_stride == 1:
            ca = res["analytic"]["raw"] // res["analytic"]["redundancy"]
            cl = res["learned"]["raw"] // res["learned"]["redundancy"]
            assert ca == cl, name
        rows.append({"pattern": name, **res})
        log(f"  {name}: analytic {res['analytic']['ms']:.1f} ms learned {res['learned']['ms']:.1f} ms "
            f"(candidate {res['learned']['choice']['chosen']} vs {res['analytic']['choice']['analytic']} of "
            f"{res['analytic']['choice']['candidates']}, stride {root_stride})")
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m graphindor_amd.costmodel")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--out", required=True)
    t.add_argument("--patterns", type=int, default=24)
    t.add_argument("--scales", default="16,18,20")
    t.add_argument("--steps", type=int, default=3000)
    t.add_argument("--target-ms", type=float, default=4.0)
    t.add_argument("--max-cands", type=int, default=96)
    t.add_argument("--dataset", default=None, help="also write the timing dataset (.npz)")
    t.add_argument("--budget-s", type=float, default=30.0, help="timing budget per (pattern, graph) group")
    e = sub.add_parser("eval")
    e.add_argument("--model", required=True)
    e.add_argument("--scale", type=int, default=20)
    e.add_argument("--ef", type=int, default=16)
    e.add_argument("--seed", type=int, default=4)
    e.add_argument("--stride", type=int, default=1)
    a = ap.parse_args(argv)
    if a.cmd == "train":
        t0 = time.time()
        pats = random_connected_patterns(a.patterns, seed=6)
        specs = [(int(s), 16, 6 + i) for i, s in enumerate(a.scales.split(","))]
        ds = collect(specs, pats, a.target_ms, 2, a.max_cands, budget_s=a.budget_s)
        print(f"dataset: {len(ds['log_ms'])} timings in {len(ds['goff']) - 1} groups, {time.time() - t0:.1f} s",
              flush=True)
        m, loss = train(ds, steps=a.steps)
        m.save(a.out)
        reg = group_regret(m, ds)
        print(json.dumps({"loss": loss, "train_regret_learned": float(np.mean([r[0] for r in reg])),
                          "train_regret_analytic": float(np.mean([r[1] for r in reg]))}), flush=True)
        if a.dataset:
            np.savez(a.dataset, x=ds["x"], log_ms=ds["log_ms"], goff=ds["goff"], meta=json.dumps(ds["meta"]))
    else:
        m = N.CostModel.load(a.model)
        pats = [("house", 5, "0101110101010101010011000"), ("c5", 5, "0100110100010100010110010"),
                ("k4_pendant", 5, "0111110110110101110010000")]
        pats += [("c6", 6, "010001101000010100001010000101100010"), ("prism", 6, "011100101010110001100011010101001110")]
        evaluate(m, a.scale, a.ef, a.seed, pats, root_stride=a.stride)
    return 0


if __name__ == "__main__":
    sys.exit(main())
