"""Writes tests/golden/config_counts.json: oracle counts for the benchmark
workloads at their full sizes, so that the GPU tests can check the device at
those sizes without running the oracle there.

Each section generates its graph with the package's seeded generators (the
same calls the GPU tests make), counts with the CPU oracle
(oracle/gs_oracle.c, test infrastructure) and records the result:

  rmat20   4- and 5-cliques of R-MAT 2^20 ef16 seed 2 and of the Erdos-Renyi
           control G(2^20, 32/2^20) seed 3 (clique rule on the degree-rank
           oriented graph; raw count and the hybrid byte tally at the
           device's default bitmap bound 1024)
  batch    per-graph sums of the 6 connected 4-motifs over the 512-graph batch
           (graph i from seed 5 + i), root sample e % 256 == 0
  rmat22   house, 5-cycle and 4-clique + pendant on R-MAT 2^22 ef16 seed 4,
           analytic schedule, edge-rooted rule on a root sample e % S == 0
  rmat23   6-cycle and 3-prism on R-MAT 2^23 ef32 seed 7, likewise

Usage: python tests/golden/make_config_fixtures.py [section ...] [--threads T]
Sections already in the file are kept unless named; within a named pattern
section, patterns already recorded at the same root stride are kept.  The
3-prism sample on R-MAT 2^23 took 43 minutes of oracle time on 8 threads.
"""
import argparse
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from graphindor_amd import native as N  # noqa: E402
from oracle import oracle as O  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "config_counts.json")

MOTIFS4 = ["0111101111011110", "0111101111001100", "0111101011001000", "0111100010001000", "0110100110010110",
           "0110100110000100"]
PM5 = {"house": "0101110101010101010011000", "c5": "0100110100010100010110010",
       "k4_pendant": "0111110110110101110010000"}
PM6 = {"prism": "011100101010110001100011010101001110", "c6": "010001101000010100001010000101100010"}
# Root samples: the strides keep each sampled count within minutes of oracle
# time here and seconds of device time.
STRIDES22 = {"house": 4096, "c5": 4194304, "k4_pendant": 256}
STRIDES23 = {"prism": 16777216, "c6": 134217728}


def clique_adj(k):
    return "".join("0" if i == j else "1" for i in range(k) for j in range(k))


def schedule_digest(sched):
    return hashlib.sha256(json.dumps(sched.dump()).encode()).hexdigest()[:16]


def section_rmat20(threads):
    out = {}
    for key, make in (("rmat20_s2", lambda: N.Graph.rmat(20, 16, 2)),
                      ("er20_s3", lambda: N.Graph.erdos_renyi(1 << 20, 32.0 / (1 << 20), 3))):
        g = make()
        og = g.orient_by_degree_rank()
        v, e = (x.copy() for x in og.csr())
        csr = O.CSR(v, e)
        rec = {"vertices": int(len(v) - 1), "oriented_edges": int(len(e))}
        for k in (4, 5):
            d = N.Schedule.plain(k, clique_adj(k)).plan_desc(N.GM_MODE_CLIQUE)
            t = time.time()
            raw, b = O.count_clique_hybrid(csr, d, 1024, threads)
            rec[f"k{k}"] = {"count": int(raw), "hybrid1024_bytes": int(b)}
            print(f"{key} k{k}: {raw} bytes {b} ({time.time() - t:.1f} s)", flush=True)
        out[key] = rec
    return out


def section_batch(threads, stride=256):
    g, voff, eoff = N.Graph.batch(512, 5, 10.0, 17.0, 4.0, 64.0)
    inf = g.info()
    v, e = (x.copy() for x in g.csr())
    csr = O.CSR(v, e)
    rec = {"graphs": 512, "csr_entries": int(len(e)), "stride": stride, "motifs": MOTIFS4, "schedules": [],
           "sums": [[0] * len(MOTIFS4) for _ in range(512)]}
    for j, adj in enumerate(MOTIFS4):
        s = N.Schedule.build(4, adj, inf["v_cnt"], inf["e_cnt"], inf["tri_cnt"])
        rec["schedules"].append(schedule_digest(s))
        t = time.time()
        sums = O.count_batch_strided(csr, eoff, s.plan_desc(), stride, threads)
        for i, x in enumerate(sums):
            rec["sums"][i][j] = x
        print(f"batch motif {j}: total {sum(sums)} ({time.time() - t:.1f} s)", flush=True)
    return {"batch512_s5": rec}


def section_pm(threads, key, make, n, patterns, strides, save, have=None):
    """Patterns already in `have` (the section as the file holds it) at the
    same stride are kept, not counted again."""
    g = make()
    inf = g.info()
    v, e = (x.copy() for x in g.csr())
    csr = O.CSR(v, e)
    rec = {"vertices": int(inf["v_cnt"]), "csr_entries": int(inf["e_cnt"]), "tri_cnt": int(inf["tri_cnt"]),
           "patterns": {}}
    if have and have.get("csr_entries") == rec["csr_entries"]:
        rec["patterns"] = {k: x for k, x in have.get("patterns", {}).items() if x.get("stride") == strides.get(k)}
    for name, adj in patterns.items():
        if name in rec["patterns"]:
            print(f"{key} {name}: kept from the file", flush=True)
            continue
        save({key: rec})  # each pattern is kept as soon as it is counted
        s, _ = N.select_schedule(n, adj, v, inf["tri_cnt"], None)
        S = strides[name]
        t = time.time()
        raw, b = O.count_edge_rooted_strided(csr, s.plan_desc(), S, threads)
        rec["patterns"][name] = {"adj": adj, "stride": S, "raw": int(raw), "alg_bytes": int(b),
                                 "schedule": schedule_digest(s)}
        print(f"{key} {name} stride {S}: raw {raw} bytes {b} ({time.time() - t:.1f} s)", flush=True)
    save({key: rec})
    return {key: rec}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("sections", nargs="*", default=["rmat20", "batch", "rmat22", "rmat23"])
    ap.add_argument("--threads", type=int, default=0)
    a = ap.parse_args()
    data = json.load(open(OUT)) if os.path.exists(OUT) else {}

    def save(part):
        data.update(part)
        with open(OUT + ".tmp", "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
            f.write("\n")
        os.replace(OUT + ".tmp", OUT)

    for sec in a.sections:
        if sec == "rmat20":
            data.update(section_rmat20(a.threads))
        elif sec == "batch":
            data.update(section_batch(a.threads))
        elif sec == "rmat22":
            data.update(section_pm(a.threads, "rmat22_s4", lambda: N.Graph.rmat(22, 16, 4), 5, PM5, STRIDES22, save,
                                   data.get("rmat22_s4")))
        elif sec == "rmat23":
            data.update(section_pm(a.threads, "rmat23_ef32_s7", lambda: N.Graph.rmat(23, 32, 7), 6, PM6, STRIDES23,
                                   save, data.get("rmat23_ef32_s7")))
        else:
            raise SystemExit(f"unknown section {sec}")
        save({})
    print("wrote", OUT)


if __name__ == "__main__":
    main()
