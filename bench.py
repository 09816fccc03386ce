"""Benchmark: k-clique counting time on one MI355X (exact).

Workload (one GPU): R-MAT 2^20 vertices, edge factor 16, (a,b,c,d) =
(0.57,0.19,0.19,0.05), seed 2, symmetrised, cleaned, ids permuted; reduced
to one direction per edge and counted with the gpu_clique rule
(reference gpu/gpu_clique.cu).  `value` is the 5-clique counting time
(kernel + sync + 8-byte readback, graph resident in HBM); the 4-clique time
and an Erdos-Renyi G(2^20, 32/2^20) control (seed 3) are reported beside it.

A step is one full count.  The first warm-up step also tallies the
algorithmic byte model (4 B per list element read or written, per set
operation), from which `roofline.achieved` = bytes / average kernel time.

cpu_baseline: the oracle's restatement of the reference CPU engine
(Graph::pattern_matching, vertex-rooted, OpenMP dynamic) with the same
schedule the reference's clique_test builds, timed on a strided sample of
root vertices (ids are randomly permuted, so the sample is unbiased) and
scaled by the stride.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the dominant kernel on the default workload (R-MAT
# 2^20 ef16 seed 2, degree-oriented), from two rocprofv3 PMC runs (FETCH_SIZE
# and WRITE_SIZE, each scaled by what it reads for a 1 GiB device copy in the
# same run); see profiles/clique5_rmat20_hbm_traffic.md.
TRAFFIC = {(20, "clique_bitmap_big_kernel<5>"): 9.191e9}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def clique_adj(k):
    return "".join("0" if i == j else "1" for i in range(k) for j in range(k))


def time_counts(N, dg, plan, steps, warmup):
    """Warm-up (first one with the byte tally), then `steps` timed counts.
    Returns the count, the byte tally (total and the bitmap engine's share),
    the mean wall time per count and the mean HIP-event times of the whole
    region, the list-engine launches and the bitmap-engine launches."""
    tally = None
    raw = None
    for i in range(max(1, warmup)):
        r, st = N.count(dg, plan, flags=N.GM_COUNT_ALG_BYTES if i == 0 else 0)
        if i == 0:
            tally = (st.alg_bytes, st.bitmap_alg_bytes, st.bitmap_block_alg_bytes)
        raw = r if raw is None else raw
        assert r == raw
    walls, kms, lms, bms, kbs = [], [], [], [], []
    for _ in range(steps):
        t0 = time.perf_counter()
        r, st = N.count(dg, plan)
        walls.append(time.perf_counter() - t0)
        kms.append(st.kernel_ms)
        lms.append(st.list_ms)
        bms.append(st.bitmap_ms)
        kbs.append(st.bitmap_block_ms)
        assert r == raw, "count changed between steps"
    mean = lambda x: sum(x) / len(x)
    return dict(count=raw, alg_bytes=tally[0], bitmap_alg_bytes=tally[1], block_alg_bytes=tally[2],
                wall_s=mean(walls), kernel_ms=mean(kms), list_ms=mean(lms), bitmap_ms=mean(bms), block_ms=mean(kbs))


def cpu_baseline(N, O, v, e, k, target_s=20.0):
    """Reference-CPU-engine restatement (vertex-rooted, OpenMP dynamic) on the
    graph the reference's clique_test counts (id-reduced, src/clique_test.cpp:49-76),
    timed on an even sample of the work: every root, but only its depth-1
    candidates with index = 1 mod S; the time is scaled by S.  A first pass at
    S = 4096 sizes S for ~target_s."""
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    csr = O.CSR(v, e)
    sched = N.Schedule.build(k, clique_adj(k), len(v) - 1, len(e), 0)
    desc = sched.plan_desc()
    t0 = time.perf_counter()
    O.count_vertex_rooted_strided(csr, desc, threads, 4096, 0)
    full = (time.perf_counter() - t0) * 4096
    stride = max(2, min(4096, int(full / target_s)))
    t0 = time.perf_counter()
    O.count_vertex_rooted_strided(csr, desc, threads, stride, 1)
    t = time.perf_counter() - t0
    return {"value": round(t * stride, 3), "unit": "s", "cores": threads, "kind": "port",
            "sample": f"{k}-clique on the id-reduced graph (clique_test rule), every root, depth-1 "
                      f"candidates i = 1 mod {stride}: {t:.1f} s measured, scaled x{stride}"}


MOTIFS4 = ["0111101111011110", "0111101111001100", "0111101011001000", "0111100010001000", "0110100110010110",
           "0110100110000100"]


def batch_leg(N, count=512, stride=256):
    """Batched 4-motif counting over the 512-graph batch (graph i from seed
    5 + i; R-MAT or Erdos-Renyi by coin flip; n log-uniform in [2^10, 2^17];
    average degree U[4, 64]).  Plain form: one launch per (graph, pattern),
    as running gpu_mc per graph does (minus its re-uploads).  Batched form:
    all 6 patterns x 512 graphs in one hipGraph-captured persistent launch
    with the autotuned tuning.  Both count the same root sample (edges e with
    e % stride == 0; the full batch is ~10^2 TB of set work) and must agree
    on every (graph, pattern) sum."""
    t0 = time.time()
    g, voff, eoff = N.Graph.batch(count, 5, 10.0, 17.0, 4.0, 64.0)
    inf = g.info()
    v, e = g.csr()
    v, e = v.copy(), e.copy()
    gen_s = time.time() - t0
    dg = N.DeviceGraph(v, e)
    scheds = [N.Schedule.build(4, adj, inf["v_cnt"], inf["e_cnt"], inf["tri_cnt"]) for adj in MOTIFS4]
    plans = [N.DevicePlan(s.plan_desc()) for s in scheds]
    # plain: warm once, then time every (graph, pattern) launch
    N.count_batch(dg, eoff[:2], plans[:1], root_stride=stride)
    plain = [[0] * len(plans) for _ in range(count)]
    t0 = time.perf_counter()
    for i in range(count):
        for j, p in enumerate(plans):
            r, _ = N.count_batch(dg, eoff[i:i + 2], [p], root_stride=stride)
            plain[i][j] = r[0][0]
    plain_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    tc = N.autotune_batch(dg, eoff, plans, "motifs4:" + N.degree_class(v), None, root_stride=stride)
    tune_s = time.perf_counter() - t0
    tun = dict(waves_per_cu=tc["waves_per_cu"], chunk_edges=tc["chunk_edges"], lane_tail=tc["lane_tail"])
    b = N.Batch(dg, eoff, plans, root_stride=stride, **tun)
    b.run()
    walls = []
    for _ in range(3):
        t0 = time.perf_counter()
        got, st = b.run()
        walls.append(time.perf_counter() - t0)
        assert got == plain, "batched and per-graph sums differ"
    b.free()
    batched_s = min(walls)
    log(f"batch: {count} graphs, {len(e)} CSR entries (gen {gen_s:.1f} s), stride {stride}: plain {plain_s:.3f} s, "
        f"batched {batched_s:.3f} s, tune {tune_s:.2f} s {tc}")
    return {"graphs": count, "csr_entries": int(len(e)), "root_sample": f"edges e with e % {stride} == 0",
            "plain_per_pattern_launches_s": round(plain_s, 4), "launches_plain": count * len(plans),
            "batched_hipgraph_s": round(batched_s, 4), "speedup": round(plain_s / batched_s, 3),
            "autotune": {k: tc[k] for k in ("waves_per_cu", "chunk_edges", "lane_tail", "tried")},
            "autotune_s": round(tune_s, 3), "sums_equal": True}


PM5 = [("house", "0101110101010101010011000"), ("c5", "0100110100010100010110010"),
       ("k4_pendant", "0111110110110101110010000")]
MODEL_PATH = os.path.join(ROOT, "graphindor_amd", "data", "cost_model.gscm")


def pm_leg(N, scale=22, seed=4, budget_s=8.0):
    """5-vertex pattern matching with IEP on R-MAT 2^scale ef16 (house, 5-cycle,
    4-clique + pendant): the plain analytic schedule and, when a trained cost
    model is present, the learned pick.  A full count on R-MAT 2^22 is far
    beyond a bench run (hub neighbourhoods of 10^5 vertices), so each pattern
    is timed on the root sample e % S == 0, S halving from 2^16 until the
    analytic run takes about budget_s / 8 (or doubling up to 2^20 while it
    exceeds budget_s); both schedules use the same S.  The analytic schedule
    is also timed with the lane-parallel tail off (`analytic_tail_off_ms`):
    with depth-2 slice donation that is faster on several patterns.
    Sampled sums differ between schedules (a root edge maps to different
    embeddings), so the comparison is of times; full-count parity is in the
    GPU tests."""
    t0 = time.time()
    g = N.Graph.rmat(scale, 16, seed)
    inf = g.info()
    v, e = g.csr()
    v, e = v.copy(), e.copy()
    gen_s = time.time() - t0
    dg = N.DeviceGraph(v, e)
    model = N.CostModel.load(MODEL_PATH) if os.path.exists(MODEL_PATH) else None
    out = {"graph": f"R-MAT 2^{scale} ef16 seed {seed}", "csr_entries": int(len(e)),
           "max_degree": int((v[1:] - v[:-1]).max()), "model": os.path.basename(MODEL_PATH) if model else None,
           "patterns": {}}
    for name, adj in PM5:
        row = {}
        sa, ia = N.select_schedule(5, adj, v, inf["tri_cnt"], None)
        pa = N.DevicePlan(sa.plan_desc())
        limit = int(budget_s * 1000)
        S = 1 << 16

        def timed(plan, **tuning):
            try:
                return N.count_batch(dg, [0, len(e)], [plan], root_stride=S, time_limit_ms=limit,
                                     **tuning)[1].kernel_ms
            except N.GmTimeout:
                return None
        t = timed(pa)
        while t is None and S < (1 << 20):
            S *= 4
            t = timed(pa)
        grown = S > (1 << 16)
        while not grown and t is not None and t / 1e3 < budget_s / 8 and S > 1:
            S //= 2
            t = timed(pa)
        row.update({"root_stride": S, "analytic_candidate": ia["analytic"], "candidates": ia["candidates"],
                    "select_ms_analytic": round(ia["select_ms"], 2)})
        if t is None:
            # a few hub roots of the smallest sample already exceed the budget
            row["analytic_ms"] = f"over {limit} ms"
            out["patterns"][name] = row
            log(f"pm {name}: stride {S}: over the {limit} ms budget")
            continue
        t = timed(pa)
        row["analytic_ms"] = round(t, 3) if t is not None else f"over {limit} ms"
        t_off = timed(pa, lane_tail=-1)
        row["analytic_tail_off_ms"] = round(t_off, 3) if t_off is not None else f"over {limit} ms"
        if model is not None and t is not None:
            sl, il = N.select_schedule_measured(5, adj, g, model)
            t2 = timed(N.DevicePlan(sl.plan_desc()))
            row.update({"learned_ms": round(t2, 3) if t2 is not None else f"over {limit} ms",
                        "learned_candidate": il["chosen"], "select_ms_learned": round(il["select_ms"], 2)})
            if t2 is not None:
                row["learned_speedup"] = round(t / max(t2, 1e-6), 3)
        out["patterns"][name] = row
        log(f"pm {name}: stride {S}: {row}")
    out["gen_s"] = round(gen_s, 1)
    return out


def calibration_copy(mb):
    """One device-to-device copy of `mb` MiB (HIP's own copy kernel), so that
    a PMC run sees a kernel of known traffic next to the counting kernels."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = mb << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"calibration copy: {mb} MiB device to device")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-control", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the 512-graph batched motif leg")
    ap.add_argument("--batch-stride", type=int, default=256)
    ap.add_argument("--no-pm", action="store_true", help="skip the 5-vertex pattern-matching leg")
    ap.add_argument("--pm-scale", type=int, default=22)
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="also run one device copy of this many MiB (PMC traffic calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the counting path runs on one GPU; several would each take a range of the root edges)", file=sys.stderr)
        sys.exit(2)

    from graphindor_amd import native as N
    if N.device_count() < 1:
        sys.exit("no GPU visible")

    if args.calib_copy_mb > 0:
        calibration_copy(args.calib_copy_mb)

    t0 = time.time()
    g = N.Graph.rmat(args.scale, 16, 2)
    info = g.info()
    t1 = time.time()
    og = g.orient_by_degree()
    orient_s = time.time() - t1
    v, e = og.csr()
    v, e = v.copy(), e.copy()
    log(f"graph: R-MAT 2^{args.scale} ef16 seed 2: {info['v_cnt']} vertices, {info['e_cnt']} CSR entries, "
        f"{len(e)} after degree orientation ({time.time() - t0:.1f} s, orientation {orient_s:.2f} s)")
    dg = N.DeviceGraph(v, e)

    results = {}
    for k in (4, 5):
        s = N.Schedule.plain(k, clique_adj(k))
        s.redundancy = 1
        plan = N.DevicePlan(s.plan_desc(N.GM_MODE_CLIQUE))
        steps = args.steps if k == 5 else 1
        r = time_counts(N, dg, plan, steps, args.warmup if k == 5 else 1)
        results[k] = r
        log(f"{k}-clique: {r['count']} in {r['wall_s']:.4f} s (region {r['kernel_ms']:.2f} ms: bitmap engine "
            f"{r['bitmap_ms']:.2f} ms of which block-per-root {r['block_ms']:.2f} ms, list engine {r['list_ms']:.2f} ms), "
            f"alg bytes {r['alg_bytes']} (bitmap {r['bitmap_alg_bytes']}, block-per-root {r['block_alg_bytes']})")

    # The reference's own orientation (edges kept toward smaller ids,
    # src/graph.cpp:710-735), 4-cliques only: its 5-clique count is far
    # slower and the count is the same.
    g.reduce_for_clique()
    iv, ie = g.csr()
    dgi = N.DeviceGraph(iv.copy(), ie.copy())
    s4 = N.Schedule.plain(4, clique_adj(4))
    s4.redundancy = 1
    ri = time_counts(N, dgi, N.DevicePlan(s4.plan_desc(N.GM_MODE_CLIQUE)), 1, 1)
    raw_i, wall_i = ri["count"], ri["wall_s"]
    assert raw_i == results[4]["count"], "orientation changed the 4-clique count"
    log(f"id-oriented 4-clique: {raw_i} in {wall_i:.4f} s")
    del dgi

    control = None
    if not args.no_control:
        ge = N.Graph.erdos_renyi(1 << args.scale, 32.0 / (1 << args.scale), 3)
        ge = ge.orient_by_degree()
        ve, ee = ge.csr()
        dge = N.DeviceGraph(ve.copy(), ee.copy())
        control = {}
        for k in (4, 5):
            s = N.Schedule.plain(k, clique_adj(k))
            s.redundancy = 1
            plan = N.DevicePlan(s.plan_desc(N.GM_MODE_CLIQUE))
            rc = time_counts(N, dge, plan, 1, 1)
            control[f"k{k}"] = {"count": rc["count"], "time_s": round(rc["wall_s"], 6)}
            log(f"ER control {k}-clique: {rc['count']} in {rc['wall_s']:.4f} s")

    r5 = results[5]
    # Dominant kernel of the 5-clique count: the one with the larger share of
    # the region; its algorithmic bytes over its own HIP-event time.
    parts = [
        {"kernel": "clique_bitmap_kernel<16,5>", "ms": r5["bitmap_ms"] - r5["block_ms"],
         "bytes": r5["bitmap_alg_bytes"] - r5["block_alg_bytes"]},
        {"kernel": "clique_bitmap_big_kernel<5>", "ms": r5["block_ms"], "bytes": r5["block_alg_bytes"]},
        {"kernel": "enumerate_kernel<1,5>", "ms": r5["list_ms"], "bytes": r5["alg_bytes"] - r5["bitmap_alg_bytes"]},
    ]
    dom = max(parts, key=lambda x: x["ms"])
    achieved = dom["bytes"] / (dom["ms"] / 1e3) / 1e9
    out = {
        "metric": "5-vertex PM & 5-clique time on 1 MI355X (exact); learned vs plain schedule",
        "value": round(r5["wall_s"], 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r5["wall_s"] * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32/u64",
        "data": "synthetic (seeded R-MAT / Erdos-Renyi, generated in-process)",
        "config": {"workload": f"5-clique count, gpu_clique rule on the degree-oriented R-MAT 2^{args.scale} ef16 seed 2 "
                               "(a,b,c,d)=(0.57,0.19,0.19,0.05), plain schedule",
                   "vertices": info["v_cnt"], "csr_entries": info["e_cnt"], "reduced_edges": int(len(e)),
                   "count_5clique": r5["count"], "count_4clique": results[4]["count"],
                   "time_4clique_s": round(results[4]["wall_s"], 6),
                   "orientation": "degree (reference degree_orientation_init), host preprocessing "
                                  f"{orient_s:.2f} s outside the timed region",
                   "time_4clique_id_orientation_s": round(wall_i, 6),
                   "er_control": control},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": TRAFFIC.get((args.scale, dom["kernel"])),
                     "kernel": dom["kernel"], "kernel_ms": round(dom["ms"], 3), "alg_bytes": dom["bytes"],
                     "region_ms": round(r5["kernel_ms"], 3),
                     "kernels_ms": {x["kernel"]: round(x["ms"], 3) for x in parts}},
    }
    if not args.no_batch:
        out["config"]["batch_4motif"] = batch_leg(N, 512, args.batch_stride)
    if not args.no_pm:
        out["config"]["pm_5vertex"] = pm_leg(N, args.pm_scale, 4)
    if not args.no_cpu:
        from oracle import oracle as O
        out["cpu_baseline"] = cpu_baseline(N, O, iv.copy(), ie.copy(), 5)
        log(f"cpu baseline: {out['cpu_baseline']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
