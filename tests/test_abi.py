"""The C ABI library loads, exports every declared entry point and rejects
malformed input with an error code (no GPU needed: validation precedes any
device call)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from graphindor_amd import native as N
from tests.common import PATTERNS

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "graphindor_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gm_[a-z_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert {"gm_graph_create", "gm_plan_create", "gm_count", "gm_count_multi", "gm_last_error"} <= set(names)
    lib = N.gm()
    for name in names:
        assert hasattr(lib, name), name


def create_graph(vertex, edge):
    h = ctypes.c_void_p()
    v = np.ascontiguousarray(vertex, dtype=np.int64)
    e = np.ascontiguousarray(edge, dtype=np.int32)
    rc = N.gm().gm_graph_create(v.ctypes.data, e.ctypes.data if len(e) else None, len(v) - 1, len(e), 0,
                                ctypes.byref(h))
    if rc == 0:
        N.gm().gm_graph_destroy(h)
    return rc, N.gm().gm_last_error().decode()


@pytest.mark.parametrize("vertex,edge,msg", [
    ([0, 2, 3], [1, 1, 0], "strictly increasing"),      # duplicate neighbour
    ([0, 1, 2], [0, 0], "self-loop"),
    ([0, 1, 2], [1, 5], "out of range"),
    ([0, 2, 1], [1, 1], "row_ptr"),
    ([0, 1, 3], [1, 0], "row_ptr"),                      # row_ptr[n] != nnz
])
def test_bad_graphs_are_rejected(vertex, edge, msg):
    rc, err = create_graph(vertex, edge)
    assert rc == N.GM_ERR_INVALID
    assert msg in err


def test_bad_plans_are_rejected():
    n, adj = PATTERNS["house"]
    s = N.Schedule.build(n, adj, 1000, 10000, 100)
    good = s.plan_desc().contents
    h = ctypes.c_void_p()

    def rc_of(desc):
        r = N.gm().gm_plan_create(ctypes.byref(desc), ctypes.byref(h))
        if r == 0:
            N.gm().gm_plan_destroy(h)
        return r

    big = N.GmPlanDesc.from_buffer_copy(good)
    big.size = 8
    assert rc_of(big) == -4                      # GM_ERR_UNSUPPORTED
    bad = N.GmPlanDesc.from_buffer_copy(good)
    pos = (ctypes.c_int32 * 64)(*([99] * 64))
    bad.iep_ans_pos = ctypes.cast(pos, ctypes.POINTER(ctypes.c_int32))
    assert rc_of(bad) == -1 and "IEP position" in N.gm().gm_last_error().decode()
    bad2 = N.GmPlanDesc.from_buffer_copy(good)
    fat = (ctypes.c_int32 * 11)(*([0] * 11))
    bad2.father_prefix_id = ctypes.cast(fat, ctypes.POINTER(ctypes.c_int32))
    assert rc_of(bad2) == -1
    bad3 = N.GmPlanDesc.from_buffer_copy(good)
    bad3.mode = 7
    assert rc_of(bad3) == -1


def test_count_rejects_null_handles():
    raw = ctypes.c_uint64()
    assert N.gm().gm_count(None, None, None, 0, None, ctypes.byref(raw), None) == -1


@pytest.mark.parametrize("ctype,cname", [(N.GmStats, "gm_stats"), (N.GmTuning, "gm_tuning")])
def test_ctypes_mirrors_match_the_header_layout(tmp_path, ctype, cname):
    """The binding's structs have the header's size and field offsets (a C
    program built against include/graphindor_amd.h prints them)."""
    names = [f[0] for f in ctype._fields_]
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "graphindor_amd.h"', "int main(void) {",
             f'    printf("%zu\\n", sizeof({cname}));']
    lines += [f'    printf("%zu\\n", offsetof({cname}, {n}));' for n in names]
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    out = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert out[0] == ctypes.sizeof(ctype)
    assert out[1:] == [getattr(ctype, n).offset for n in names]
