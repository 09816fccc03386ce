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

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the dominant kernel, keyed by (scale, orientation,
# kernel), from two rocprofv3 PMC runs of this bench (FETCH_SIZE and
# WRITE_SIZE, each scaled by what it reads for a 1 GiB device copy in the
# same run).  None where no measurement of the current kernel exists.
TRAFFIC = {
    # FETCH_SIZE 3,554,610 KiB x 2.000 + WRITE_SIZE 1,939 KiB x 1.000 per
    # dispatch (mean of 4); the 1 GiB copy read 524,306 KiB and wrote
    # 1,048,576 KiB.
    (20, "rank", "clique_bitmap_big_kernel<5>"): 7281826304,
}


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


def host_cpu():
    """CPU model, sockets, cores and threads of this host (lscpu), and the
    physical core count."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, val = line.partition(":")
            info[k.strip()] = val.strip()
    except Exception:
        pass
    sockets = int(info.get("Socket(s)", "1") or 1)
    cores = int(info.get("Core(s) per socket", "0") or 0) * sockets
    smt = int(info.get("Thread(s) per core", "1") or 1)
    # a container may see fewer CPUs than the host has cores
    visible = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    physical = max(1, min(cores or visible, visible))
    return {"model": info.get("Model name", "unknown"), "sockets": sockets, "cores": cores, "threads_per_core": smt,
            "visible_cpus": visible}, physical


def cpu_baseline(N, O, graphs, k, host, target_s=20.0):
    """The oracle's restatement of the reference CPU engine (vertex-rooted,
    OpenMP dynamic; src/graph.cpp:305-566) counting k-cliques, on every
    physical core with OMP_PROC_BIND=close.  graphs: [(label, vertex, edge)]; the first is the
    graph the GPU counts (degree-oriented), the second the reference's own
    id orientation.  Each is counted in full when a first sample says that
    fits the budget; otherwise on an even sample of the work (every root,
    depth-1 candidates i = 1 mod S), reported as a partial-progress rate
    with the extrapolated full time labelled as such.  host: host_cpu()'s
    result, taken before any OpenMP runtime loaded (with OMP_PROC_BIND set,
    libgomp binds the initial thread to one CPU when it starts, after which
    the process's affinity mask shows one CPU)."""
    cpu, threads = host
    out = {"unit": "s", "cores": threads, "kind": "port", "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
           "cpu": cpu, "graphs": {}}
    for label, v, e in graphs:
        csr = O.CSR(v, e)
        if label == "id_oriented":
            # clique_test's schedule (src/clique_test.cpp:49-76): its id
            # restrictions agree with edges kept toward smaller ids
            sched = N.Schedule.build(k, clique_adj(k), len(v) - 1, len(e), 0)
        else:
            # a degree orientation is a DAG in another order: no restrictions
            sched = N.Schedule.plain(k, clique_adj(k))
            sched.redundancy = 1
        desc = sched.plan_desc()
        # Probe on ever larger samples (1/4096 of the depth-1 candidates,
        # then 1/512, 1/64) until one takes a second, to estimate the full
        # count's time without running past the budget.
        for probe in (4096, 512, 64):
            t0 = time.perf_counter()
            O.count_vertex_rooted_strided(csr, desc, threads, probe, 0)
            dt = time.perf_counter() - t0
            est = dt * probe
            log(f"cpu baseline {label}: 1/{probe} sample in {dt:.2f} s (full count ~{est:.0f} s)")
            if dt >= 1.0:
                break
        if est <= target_s * 1.5:
            t0 = time.perf_counter()
            raw = O.count_vertex_rooted(csr, desc, threads)
            t = time.perf_counter() - t0
            rec = {"time_s": round(t, 3), "count": int(raw), "measured": "full count, 1 run"}
        else:
            stride = max(2, min(4096, int(est / target_s)))
            t0 = time.perf_counter()
            O.count_vertex_rooted_strided(csr, desc, threads, stride, 1)
            t = time.perf_counter() - t0
            rec = {"time_s": round(t * stride, 3), "measured": f"partial: 1/{stride} of the depth-1 candidates of every "
                   f"root in {t:.1f} s; time_s extrapolates that rate to the full count"}
        out["graphs"][label] = rec
        log(f"cpu baseline {label}: {rec}")
    first = next(iter(out["graphs"].values()))
    out["value"] = first["time_s"]
    out["sample"] = (f"{k}-clique, clique_test schedule, on the degree-oriented graph the GPU counts: "
                     f"{first['measured']}; id-oriented (reference reduce_edges_for_clique) alongside")
    return out


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
    on every (graph, pattern) sum (and with the oracle's fixture at stride
    256).  The plain form is reported as wall time (3,072 synchronous calls,
    as gpu_mc's per-motif launches) and as the sum of its kernel times."""
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
    plain_kernel_ms = 0.0
    t0 = time.perf_counter()
    for i in range(count):
        for j, p in enumerate(plans):
            r, st = N.count_batch(dg, eoff[i:i + 2], [p], root_stride=stride)
            plain[i][j] = r[0][0]
            plain_kernel_ms += st.kernel_ms
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
    fix = fixture("batch512_s5")
    parity = None
    if fix and fix["stride"] == stride and count == fix["graphs"]:
        parity = plain == fix["sums"]
        assert parity, "batch sums differ from the oracle fixture"
    return {"graphs": count, "csr_entries": int(len(e)), "root_sample": f"edges e with e % {stride} == 0",
            "plain_per_pattern_launches_s": round(plain_s, 4), "launches_plain": count * len(plans),
            "plain_kernel_time_s": round(plain_kernel_ms / 1e3, 4),
            "batched_hipgraph_s": round(batched_s, 4), "speedup": round(plain_s / batched_s, 3),
            "speedup_kernel_time": round(plain_kernel_ms / 1e3 / batched_s, 3), "oracle_fixture_equal": parity,
            "autotune": {k: tc[k] for k in ("waves_per_cu", "chunk_edges", "lane_tail", "tried")},
            "autotune_s": round(tune_s, 3), "sums_equal": True}


PM5 = [("house", "0101110101010101010011000"), ("c5", "0100110100010100010110010"),
       ("k4_pendant", "0111110110110101110010000")]
PM6 = [("c6", "010001101000010100001010000101100010"), ("prism", "011100101010110001100011010101001110")]
MODEL_PATH = os.path.join(ROOT, "graphindor_amd", "data", "cost_model.gscm")
FIXTURES = os.path.join(ROOT, "tests", "golden", "config_counts.json")
# Root samples e % S == 0 of the pattern legs: those of the parity fixtures.
PM_STRIDES = {"house": 4096, "c5": 4194304, "k4_pendant": 256, "c6": 134217728, "prism": 16777216}
# The autotuned pattern-rule knobs: the lane-parallel last level on or off,
# and hub bitmaps for vertices of degree >= 4096 or none (library default:
# last level on, no hub bitmaps).
HUB_DEGREE = 4096
KNOBS = [{"lane_tail": 0, "hub_degree": -1}, {"lane_tail": -1, "hub_degree": -1},
         {"lane_tail": 0, "hub_degree": HUB_DEGREE}, {"lane_tail": -1, "hub_degree": HUB_DEGREE}]


def knob_label(kn):
    return f"tail{'off' if kn['lane_tail'] < 0 else 'on'}_hub{'off' if kn['hub_degree'] < 0 else 'on'}"


def fixture(key):
    if not os.path.exists(FIXTURES):
        return None
    return json.load(open(FIXTURES)).get(key)


def pm_patterns(N, g, patterns, n, fix, budget_ms=60000, full=(), learned=True):
    """Each pattern on its fixed root sample: the plain analytic schedule with
    the library defaults (its raw sum and byte tally checked against the
    oracle's fixture), the four settings of the lane-parallel last level and
    the hub bitmaps (autotuning, on the same sample), and the schedule the
    learned cost model proposes and a measurement on the same sample
    confirms.  Patterns named in `full` are then counted over every root
    edge with the tuned knobs (stopped cleanly after 170 s)."""
    inf = g.info()
    v, e = (x.copy() for x in g.csr())
    dg = N.DeviceGraph(v, e)
    model = N.CostModel.load(MODEL_PATH) if learned and os.path.exists(MODEL_PATH) else None
    out = {}
    for name, adj in patterns:
        S = PM_STRIDES[name]
        row = {"root_stride": S, "sampled_roots": (len(e) + S - 1) // S}

        def timed(plan, flags=0, **tun):
            try:
                raw, st = N.count_batch(dg, [0, len(e)], [plan], flags=flags, root_stride=S, time_limit_ms=budget_ms,
                                        **tun)
                return raw[0][0], st
            except N.GmTimeout:
                return None, None
        sa, ia = N.select_schedule(n, adj, v, inf["tri_cnt"], None)
        pa = N.DevicePlan(sa.plan_desc())
        log(f"pm {name}: {row['sampled_roots']} roots, plain analytic schedule (candidate {ia['analytic']} "
            f"of {ia['candidates']})")
        # A run past the budget is recorded as such and the others go on: on
        # the 3-prism sample the library defaults pass 60 s, the tuned knobs not.
        over = f"over {budget_ms} ms"
        f = (fix or {}).get("patterns", {}).get(name)
        raw, st = timed(pa, N.GM_COUNT_ALG_BYTES)
        if raw is None:
            log(f"pm {name}: library defaults with the byte tally {over}")
        elif f and f["stride"] == S:
            row["parity"] = {"raw": raw, "alg_bytes": st.alg_bytes,
                             "oracle_equal": (raw, st.alg_bytes) == (f["raw"], f["alg_bytes"])}
            assert row["parity"]["oracle_equal"], f"{name}: device differs from the oracle fixture"

        def check(r):
            """r: a raw sum of the analytic schedule on the sample, which
            every setting of the knobs must reproduce.  (Another schedule
            assigns the embeddings to other root edges, so its sums over a
            root sample differ; schedules agree on full counts, which the
            oracle tests check.)"""
            nonlocal raw
            if r is None:
                return
            if raw is None:
                raw = r
                if f and f["stride"] == S:
                    row["parity"] = {"raw": raw, "oracle_equal": raw == f["raw"]}
                    assert row["parity"]["oracle_equal"], f"{name}: device differs from the oracle fixture"
            assert r == raw
        times = {}
        for kn in KNOBS:
            r2, st2 = timed(pa, **kn)
            check(r2)
            times[knob_label(kn)] = round(st2.kernel_ms, 3) if st2 else over
            log(f"pm {name}: analytic {knob_label(kn)} {times[knob_label(kn)]} ms")
        done = [kn for kn in KNOBS if not isinstance(times[knob_label(kn)], str)]
        row["plain_ms"] = times["tailon_huboff"]  # library defaults
        row["analytic_knobs_ms"] = times
        row["analytic_candidate"] = ia["analytic"]
        row["candidates"] = ia["candidates"]
        if not done:
            out[name] = row
            log(f"pm {name}: {row}")
            continue
        best_knob = min(done, key=lambda kn: times[knob_label(kn)])
        best_ms = times[knob_label(best_knob)]
        # The same knobs with 8 resident waves per CU, the build that keeps
        # all its state in registers (the default runs 16, spilling some).
        kn8 = dict(best_knob, waves_per_cu=8)
        r8, st8 = timed(pa, **kn8)
        check(r8)
        times[knob_label(best_knob) + "_w8"] = round(st8.kernel_ms, 3) if st8 else over
        log(f"pm {name}: analytic {knob_label(best_knob)} at 8 waves per CU {times[knob_label(best_knob) + '_w8']} ms")
        if st8 and st8.kernel_ms < best_ms:
            best_knob, best_ms = kn8, st8.kernel_ms
        row["tuned_knobs"] = knob_label(best_knob) + ("_w8" if "waves_per_cu" in best_knob else "")
        row["analytic_tuned_ms"] = round(best_ms, 3)
        plain = row["plain_ms"] if not isinstance(row["plain_ms"], str) else None
        if model is not None:
            # The candidates are timed under the tuned knobs, each cut off at
            # twice the tuned time and at 35 s, so that the four timed
            # candidates of one selection call end within 140 s.  A sample
            # whose tuned count takes over 30 s is subsampled for the
            # selection (e % 4S == 0: a quarter of the same roots).
            sel_stride = S if best_ms < 30000 else 4 * S
            sl, il = N.select_schedule_measured(
                n, adj, g, model, root_stride=sel_stride,
                time_limit_ms=int(min(35000, max(200, 2 * best_ms * S / sel_stride))), **best_knob)
            row["select_root_stride"] = sel_stride
            log(f"pm {name}: learned selection chose candidate {il['chosen']} in {il['select_ms'] / 1e3:.1f} s")
            pl = N.DevicePlan(sl.plan_desc())
            r3, st3 = timed(pl)
            r4, st4 = timed(pl, **best_knob)
            if r3 is not None and r4 is not None:
                assert r3 == r4, f"{name}: the learned schedule's sum changed with the knobs"
            row.update({"learned_candidate": il["chosen"], "select_ms_learned": round(il["select_ms"], 1),
                        "learned_ms": round(st3.kernel_ms, 3) if st3 else over,
                        "learned_tuned_ms": round(st4.kernel_ms, 3) if st4 else over})
            if st4 and plain:
                row["learned_tuned_speedup_vs_plain"] = round(plain / max(st4.kernel_ms, 1e-6), 3)
            if st4:
                row["learned_tuned_speedup_vs_analytic_tuned"] = round(best_ms / max(st4.kernel_ms, 1e-6), 3)
        if name in full:
            log(f"pm {name}: full count over {len(e)} root edges with {knob_label(best_knob)} "
                f"(sample rate predicts {row['analytic_tuned_ms'] * S / 1e3:.0f} s)")
            try:
                rf, stf = N.count(dg, pa, time_limit_ms=170000, **best_knob)
                row["full"] = {"count": rf // sa.redundancy, "raw": rf, "s": round(stf.kernel_ms / 1e3, 3),
                               "knobs": knob_label(best_knob)}
            except N.GmTimeout:
                row["full"] = "over 170 s"
            log(f"pm {name}: full count {row['full']}")
        out[name] = row
        log(f"pm {name}: {row}")
    dg.free()
    return out


def pm_leg(N, scale=22, seed=4, full=(), only=None, learned=True):
    """5-vertex pattern matching with IEP on R-MAT 2^scale ef16 seed 4 (house,
    5-cycle, 4-clique + pendant), each on the fixed root sample of its parity
    fixture.  Full counts at this size are beyond a bench run (the oracle's
    byte model puts the 5-cycle near 3e18 bytes of set work); the 4-clique +
    pendant full count (minutes) is timed when `full` names it."""
    log(f"pm: generating R-MAT 2^{scale} ef16 seed {seed}")
    t0 = time.time()
    g = N.Graph.rmat(scale, 16, seed)
    gen_s = time.time() - t0
    fix = fixture(f"rmat{scale}_s{seed}")
    out = {"graph": f"R-MAT 2^{scale} ef16 seed {seed}", "csr_entries": int(g.info()["e_cnt"]), "gen_s": round(gen_s, 1),
           "patterns": pm_patterns(N, g, [p for p in PM5 if not only or p[0] in only], 5, fix, full=full,
                                   learned=learned)}
    return out


def pm6_leg(N, only=None):
    """The learned-model evaluation workload: 6-cycle and 3-prism on R-MAT
    2^23 ef32 seed 7, each on the fixed root sample of its parity fixture."""
    log("pm6: generating R-MAT 2^23 ef32 seed 7")
    t0 = time.time()
    g = N.Graph.rmat(23, 32, 7)
    gen_s = time.time() - t0
    fix = fixture("rmat23_ef32_s7")
    return {"graph": "R-MAT 2^23 ef32 seed 7", "csr_entries": int(g.info()["e_cnt"]), "gen_s": round(gen_s, 1),
            "patterns": pm_patterns(N, g, [p for p in PM6 if not only or p[0] in only], 6, fix)}


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


def intersection_leg(N, dg, v, e):
    """The intersection kernel in isolation, on the operand pairs of the real
    workload: every depth-1 pair N+(v0) ∩ N+(v1) of the oriented graph -- one
    per oriented edge -- intersected for its size (their sum is the triangle
    count).  Bytes by
    the size-only model, 4 (|A| + |B|) per pair.  Timed with the bitmap
    engine (rows of N+(v0) in LDS, every N+(v1) streamed once) and with the
    list engine (lane-parallel binary searches, bitmap engine off)."""
    deg = np.diff(v).astype(np.int64)
    pair_bytes = 4 * int((deg * deg).sum() + deg[e].sum())
    s = N.Schedule.plain(3, clique_adj(3))
    s.redundancy = 1
    plan = N.DevicePlan(s.plan_desc(N.GM_MODE_CLIQUE))
    out = {"pairs": int(len(e)), "alg_bytes": pair_bytes}
    for label, kn in (("bitmap_engine", {}), ("list_engine", {"bitmap_max_degree": -1})):
        raw, _ = N.count(dg, plan, **kn)
        best = min(N.count(dg, plan, **kn)[1].kernel_ms for _ in range(3))
        gbs = pair_bytes / (best / 1e3) / 1e9
        out[label] = {"triangles": int(raw), "ms": round(best, 3), "achieved_GBs": round(gbs, 1),
                      "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 3)}
        log(f"intersection ({label}): {len(e)} pairs, {pair_bytes / 1e9:.2f} GB in {best:.3f} ms = {gbs:.0f} GB/s")
    assert out["bitmap_engine"]["triangles"] == out["list_engine"]["triangles"], "the two engines count differently"
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=20)
    ap.add_argument("--orientation", choices=("rank", "degree"), default="rank",
                    help="rank: degree orientation with vertices renumbered by (degree, id); degree: as "
                         "degree_orientation_init leaves them")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-control", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the 512-graph batched motif leg")
    ap.add_argument("--batch-stride", type=int, default=64)
    ap.add_argument("--no-pm", action="store_true", help="skip the 5-vertex pattern-matching leg")
    ap.add_argument("--pm6", action="store_true",
                    help="also run the 6-vertex learned-model leg (R-MAT 2^23 ef32; minutes)")
    ap.add_argument("--pm-full", action="store_true",
                    help="also count 4-clique + pendant on R-MAT 2^22 in full (minutes)")
    ap.add_argument("--pm-scale", type=int, default=22)
    ap.add_argument("--pm-patterns", default="", help="comma-separated subset of house,c5,k4_pendant (and c6,prism "
                                                       "for --pm6)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-model selection in the PM legs")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="also run one device copy of this many MiB (PMC traffic calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the counting path runs on one GPU; several would each take a range of the root edges)", file=sys.stderr)
        sys.exit(2)

    # The CPU baseline's OpenMP threads stay on neighbouring cores (read by
    # libgomp when the host library first loads).  The host's CPUs are read
    # before that: libgomp then binds this thread to one of them.
    host = host_cpu()
    os.environ.setdefault("OMP_PROC_BIND", "close")
    from graphindor_amd import native as N
    if N.device_count() < 1:
        sys.exit("no GPU visible")

    if args.calib_copy_mb > 0:
        calibration_copy(args.calib_copy_mb)

    t0 = time.time()
    g = N.Graph.rmat(args.scale, 16, 2)
    info = g.info()
    t1 = time.time()
    og = g.orient_by_degree_rank() if args.orientation == "rank" else g.orient_by_degree()
    orient_s = time.time() - t1
    v, e = og.csr()
    v, e = v.copy(), e.copy()
    log(f"graph: R-MAT 2^{args.scale} ef16 seed 2: {info['v_cnt']} vertices, {info['e_cnt']} CSR entries, "
        f"{len(e)} after {args.orientation} orientation ({time.time() - t0:.1f} s, orientation {orient_s:.2f} s)")
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
    fix = fixture(f"rmat{args.scale}_s2")
    if fix:
        for k in (4, 5):
            assert (results[k]["count"], results[k]["alg_bytes"]) == \
                (fix[f"k{k}"]["count"], fix[f"k{k}"]["hybrid1024_bytes"]), f"{k}-clique differs from the oracle fixture"
    isect = intersection_leg(N, dg, v, e)

    # The same counts on the orientation degree_orientation_init leaves (ids
    # not renumbered) and on the reference's own (edges kept toward smaller
    # ids, src/graph.cpp:710-735; 4-cliques only, its 5-clique count is far
    # slower).  Counts are identical.
    other = {}
    og2 = g.orient_by_degree() if args.orientation == "rank" else g.orient_by_degree_rank()
    dv, de = (x.copy() for x in og2.csr())
    del og2
    dgo = N.DeviceGraph(dv, de)
    for k in (4, 5):
        s = N.Schedule.plain(k, clique_adj(k))
        s.redundancy = 1
        ro = time_counts(N, dgo, N.DevicePlan(s.plan_desc(N.GM_MODE_CLIQUE)), 1, 1)
        assert ro["count"] == results[k]["count"]
        other[f"time_{k}clique_{'degree' if args.orientation == 'rank' else 'rank'}_orientation_s"] = round(ro["wall_s"], 6)
    dgo.free()
    g.reduce_for_clique()
    iv, ie = g.csr()
    dgi = N.DeviceGraph(iv.copy(), ie.copy())
    s4 = N.Schedule.plain(4, clique_adj(4))
    s4.redundancy = 1
    ri = time_counts(N, dgi, N.DevicePlan(s4.plan_desc(N.GM_MODE_CLIQUE)), 1, 1)
    raw_i, wall_i = ri["count"], ri["wall_s"]
    assert raw_i == results[4]["count"], "orientation changed the 4-clique count"
    log(f"id-oriented 4-clique: {raw_i} in {wall_i:.4f} s; other orientation: {other}")
    dgi.free()

    control = None
    if not args.no_control:
        ge = N.Graph.erdos_renyi(1 << args.scale, 32.0 / (1 << args.scale), 3)
        ge = ge.orient_by_degree_rank()
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
        dge.free()

    r5 = results[5]
    # Dominant kernel of the 5-clique count: the one with the larger share of
    # the region; its algorithmic bytes over its own HIP-event time.
    parts = [
        {"kernel": "clique_bitmap_small_kernel<5>", "ms": r5["bitmap_ms"] - r5["block_ms"],
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
                   "oracle_fixture_equal": bool(fix),
                   "time_4clique_s": round(results[4]["wall_s"], 6),
                   "orientation": (f"{args.orientation}: degree orientation (reference degree_orientation_init)"
                                   + (" with vertices renumbered by (degree, id) rank" if args.orientation == "rank" else "")
                                   + f", host preprocessing {orient_s:.2f} s outside the timed region"),
                   **other,
                   "time_4clique_id_orientation_s": round(wall_i, 6),
                   "er_control": control,
                   "intersection": isect},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": TRAFFIC.get((args.scale, args.orientation, dom["kernel"])),
                     "kernel": dom["kernel"], "kernel_ms": round(dom["ms"], 3), "alg_bytes": dom["bytes"],
                     "region_ms": round(r5["kernel_ms"], 3),
                     "kernels_ms": {x["kernel"]: round(x["ms"], 3) for x in parts}},
    }
    if not args.no_batch:
        out["config"]["batch_4motif"] = batch_leg(N, 512, args.batch_stride)
    if not args.no_pm:
        out["config"]["pm_5vertex"] = pm_leg(N, args.pm_scale, 4, full=("k4_pendant",) if args.pm_full else (),
                                             only=[x for x in args.pm_patterns.split(",") if x],
                                             learned=not args.no_learned)
    if args.pm6:
        out["config"]["pm_6vertex_learned"] = pm6_leg(N, only=[x for x in args.pm_patterns.split(",") if x])
    if not args.no_cpu:
        from oracle import oracle as O
        out["cpu_baseline"] = cpu_baseline(N, O, [("degree_oriented", v, e), ("id_oriented", iv.copy(), ie.copy())], 5,
                                           host)
        log(f"cpu baseline: {out['cpu_baseline']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
