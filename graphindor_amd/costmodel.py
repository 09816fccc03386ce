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


def subset(ds, groups):
    """The dataset restricted to the given group indices."""
    xs, ts, goff = [], [], [0]
    for g in groups:
        a, b = ds["goff"][g], ds["goff"][g + 1]
        xs.append(ds["x"][a:b])
        ts.append(ds["log_ms"][a:b])
        goff.append(goff[-1] + (b - a))
    return {"x": np.concatenate(xs), "log_ms": np.concatenate(ts), "goff": np.asarray(goff, dtype=np.int32)}


def pick(y, margin):
    """The model's pick in one group (row 0 = the analytic pick): its argmin
    if that is predicted faster than row 0 by more than `margin` (log time),
    else row 0 -- the rule CostModel::choose applies."""
    i = int(np.argmin(y))
    return i if y[i] < y[0] - margin else 0


def group_regret(model, ds, margin=0.0):
    """Per group: time of the model's pick / best time, and of the analytic pick (row 0)."""
    y = model.predict(ds["x"])
    out = []
    for g in range(len(ds["goff"]) - 1):
        a, b = ds["goff"][g], ds["goff"][g + 1]
        t = np.exp(ds["log_ms"][a:b])
        out.append((float(t[pick(y[a:b], margin)] / t.min()), float(t[0] / t.min())))
    return out


def train(ds, hidden=256, steps=3000, lr=1e-3, weight_decay=1e-4, seed=6, val_frac=0.25, chunk=250, tmp=None,
          log=print):
    """Trains on 1 - val_frac of the groups; keeps the checkpoint and the
    trust margin with the lowest mean regret (time of the pick / best time)
    on the held-out groups.  Returns (model, info)."""
    import os
    import tempfile
    ng = len(ds["goff"]) - 1
    rng = np.random.default_rng(seed)
    perm = rng.permutation(ng)
    nval = max(1, int(round(ng * val_frac))) if ng > 3 else 0
    val, tr = sorted(perm[:nval].tolist()), sorted(perm[nval:].tolist())
    dtr = subset(ds, tr)
    dval = subset(ds, val) if val else None
    margins = [0.0, math.log(1.05), math.log(1.1), math.log(1.25), math.log(1.5), 1e9]
    m = N.CostModel.new(hidden, seed)
    path = tmp or os.path.join(tempfile.mkdtemp(), "best.gscm")
    best = (math.inf, None, None, 0)
    loss = None
    for it in range(max(1, steps // chunk)):
        loss = m.train(dtr["x"], dtr["log_ms"], dtr["goff"], chunk, lr, weight_decay, fit_norm=(it == 0))
        if dval is None:
            continue
        for mg in margins:
            reg = float(np.mean([r[0] for r in group_regret(m, dval, mg)]))
            if reg < best[0] - 1e-9:
                m.margin = mg
                m.save(path)
                best = (reg, mg, (it + 1) * chunk, loss)
    info = {"groups_train": len(tr), "groups_val": len(val), "loss": loss}
    if dval is not None and best[1] is not None:
        m = N.CostModel.load(path)
        analytic = float(np.mean([r[1] for r in group_regret(m, dval, 1e9)]))
        info.update({"val_regret_learned": best[0], "val_regret_analytic": analytic, "margin": best[1],
                     "steps": best[2]})
    return m, info


def evaluate(model, scale, ef, seed, patterns, log=print, root_stride=1, limit_ms=60000):
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
            if mdl is None:
                s, info = N.select_schedule(n, adj, v, inf["tri_cnt"], None)
            else:
                s, info = N.select_schedule_measured(n, adj, g, mdl)
            dp = N.DevicePlan(s.plan_desc())
            try:
                raw, st = N.count_batch(dg, [0, len(e)], [dp], root_stride=root_stride, time_limit_ms=limit_ms)
                res[kind] = {"raw": raw[0][0], "redundancy": s.redundancy, "ms": st.kernel_ms, "choice": info}
            except N.GmTimeout:
                res[kind] = {"raw": None, "redundancy": s.redundancy, "ms": float(limit_ms), "choice": info,
                             "timed_out": True}
        if any(r.get("timed_out") for r in res.values()):
            rows.append({"pattern": name, **res})
            log(f"  {name}: over the {limit_ms} ms limit (analytic {res['analytic']['ms']:.1f} ms, "
                f"learned {res['learned']['ms']:.1f} ms)")
            continue
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
    t.add_argument("--ef", default="16", help="edge factors, one per scale or one for all")
    t.add_argument("--from-dataset", default=None, help="train on a saved timing dataset (.npz) instead")
    e = sub.add_parser("eval")
    e.add_argument("--model", required=True)
    e.add_argument("--scale", type=int, default=20)
    e.add_argument("--ef", type=int, default=16)
    e.add_argument("--seed", type=int, default=4)
    e.add_argument("--stride", type=int, default=1)
    a = ap.parse_args(argv)
    if a.cmd == "train":
        t0 = time.time()
        if a.from_dataset:
            z = np.load(a.from_dataset)
            ds = {"x": z["x"], "log_ms": z["log_ms"], "goff": z["goff"], "meta": json.loads(str(z["meta"]))}
        else:
            pats = random_connected_patterns(a.patterns, seed=6)
            scales = [int(s) for s in a.scales.split(",")]
            efs = [int(x) for x in a.ef.split(",")]
            efs = efs * len(scales) if len(efs) == 1 else efs
            specs = [(sc, ef, 6 + i) for i, (sc, ef) in enumerate(zip(scales, efs))]
            ds = collect(specs, pats, a.target_ms, 2, a.max_cands, budget_s=a.budget_s)
        print(f"dataset: {len(ds['log_ms'])} timings in {len(ds['goff']) - 1} groups, {time.time() - t0:.1f} s",
              flush=True)
        m, info = train(ds, steps=a.steps)
        m.save(a.out)
        reg = group_regret(m, ds, m.margin)
        info.update({"all_regret_learned": float(np.mean([r[0] for r in reg])),
                     "all_regret_analytic": float(np.mean([r[1] for r in reg]))})
        print(json.dumps(info), flush=True)
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
