"""Writes tests/golden/reference_schedules.json: the schedules the reference's
own Schedule_IEP constructor builds (oracle/_ref/libgsref.so, compiled from
the reference sources in this container) for the benchmark patterns at the
graph statistics the schedule model sees.  Rerun after changing the
generator:  python tests/golden/make_fixtures.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from graphindor_amd import native as N  # noqa: E402
from oracle import oracle as O  # noqa: E402

PATTERNS = {
    "triangle": (3, "011101110"),
    "k4": (4, "0111101111011110"), "k5": (5, "0111110111110111110111110"),
    "house": (5, "0111010011100011100001100"), "house_cfg": (5, "0101110101010101010011000"),
    "c5": (5, "0100110100010100010110010"), "k4_pendant": (5, "0111110110110101110010000"),
    "diamond": (4, "0111101111001100"), "tailed_triangle": (4, "0111101011001000"),
    "star3": (4, "0111100010001000"), "c4": (4, "0110100110010110"), "path4": (4, "0110100110000100"),
    "c6": (6, "010001101000010100001010000101100010"),
    "prism": (6, "011100101010110001100011010101001110"),
}


def graph_stats():
    stats = {"wiki_vote_header": (7115, 201524, 627584181), "pm_test_tri0": (7115, 201524, 0)}
    for name, (scale, seed) in {"rmat14_s1": (14, 1), "rmat20_s2": (20, 2), "rmat22_s4": (22, 4)}.items():
        g = N.Graph.rmat(scale, 16, seed)
        i = g.info()
        stats[name] = (i["v_cnt"], i["e_cnt"], i["tri_cnt"])
        g.free()
    g = N.Graph.erdos_renyi(1 << 20, 32.0 / (1 << 20), 3)
    i = g.info()
    stats["er20_s3"] = (i["v_cnt"], i["e_cnt"], i["tri_cnt"])
    g.free()
    return stats


def main():
    if O.ref() is None:
        sys.exit("oracle/_ref is not built; run python -m graphindor_amd.build where the reference sources exist")
    stats = graph_stats()
    out = {"graph_stats": stats, "schedules": []}
    for sname, (v, e, t) in stats.items():
        for pname, (n, adj) in PATTERNS.items():
            for iep in (True, False):
                dump = O.ref_schedule_dump(n, adj, v, e, t, iep=iep)
                out["schedules"].append({"graph": sname, "pattern": pname, "n": n, "adj": adj, "iep": iep,
                                         "v_cnt": v, "e_cnt": e, "tri_cnt": t, "dump": dump})
    with open(os.path.join(HERE, "reference_schedules.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print("wrote", len(out["schedules"]), "schedules")


if __name__ == "__main__":
    main()
