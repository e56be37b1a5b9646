"""The C++ drop-in: a program compiled against include/cj/plan.h and linked
with libcalamarijoins_amd.so calls Contest::build_context / execute /
destroy_context the way the reference's callers do (tests/unit_tests.cpp:28-30,
tests/read_sql.cpp:1380), on all 21 reference known-answer plans and on the
100k x 1M join.  Checked: rows (KAT expectations, the oracle for the large
join), result pages structurally valid, pages unchanged until the next
execute() (execute.cpp:313), empty results with no pages and no mapping
(materialize.h:61-69), and TimingStats phases that never sum past the total.
The untruncated phases of the C ABI are checked on the same large join."""
import json
import os
import subprocess
import time

import numpy as np
import pytest

import cjtest
from calamarijoins_amd import abi, host, workloads

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(REPO, "tests", "dropin", "contest_dropin")
KATS = json.load(open(os.path.join(REPO, "tests", "golden", "unit_kats.json")))["cases"]


class TblResult:
    """A result read back from a .tbl file, shaped like an engine result."""

    def __init__(self, path):
        self.t = host.read_tbl(path)
        self.num_rows = self.t.num_rows
        self.types = [c.type for c in self.t.columns]
        self.page_lists = [[c.pages.ctypes.data + i * cjtest.PAGE for i in range(c.num_pages)] if c.num_pages else []
                           for c in self.t.columns]

    def close(self):
        self.t.close()


def spec_lines(name, plan, tmp):
    lines = [f"case {name.replace(' ', '_')}"]
    for i, t in enumerate(plan.inputs):
        path = os.path.join(tmp, f"{len(os.listdir(tmp))}_in{i}.tbl")
        host.write_tbl(path, t)
        lines.append(f"input {path}")
    for n in plan.nodes:
        attrs = " ".join(f"{a} {abi.TYPE_NAMES[t] if isinstance(t, str) else t}" for a, t in n["attrs"])
        k = len(n["attrs"])
        if n["kind"] == abi.CJ_NODE_SCAN:
            lines.append(f"scan {n['table']} {k} {attrs}")
        else:
            lines.append(f"join {int(n['build_left'])} {n['left']} {n['right']} {n['left_attr']} {n['right_attr']} "
                         f"{k} {attrs}")
    lines.append(f"root {plan.root}")
    return lines


def test_gpu_contest_api_dropin(tmp_path):
    assert os.path.exists(BIN), "tests/dropin/contest_dropin is built by __graft_entry__.build()"
    tmp = str(tmp_path)
    lines, expect = [], []
    for case in KATS:
        lines += spec_lines(case["name"], cjtest.kat_plan(case), tmp)
        out = os.path.join(tmp, f"out{len(expect)}.tbl")
        lines.append(f"run {out}")
        expect.append(("kat", case, out))
    b, p = workloads.c1_tables(seed=1001)
    c1 = workloads.c1_plan(b, p)
    for rep in range(2):  # the second run is warm (context memory pools filled)
        lines += spec_lines("c1_100k_x_1M", c1, tmp)
        out = os.path.join(tmp, f"out{len(expect)}.tbl")
        lines.append(f"run {out}")
        expect.append(("c1", None, out))
    print(f"running the contest-API program on {len(expect)} plans", flush=True)
    proc = subprocess.run([BIN], input="\n".join(lines) + "\n", capture_output=True, text=True, timeout=600)
    assert proc.returncode == 0, proc.stdout[-2000:] + proc.stderr[-2000:]
    res = [ln.split() for ln in proc.stdout.splitlines() if ln.startswith("result ")]
    assert len(res) == len(expect), proc.stdout[-2000:]
    oracle_c1 = cjtest.OracleRun(c1, threads=8)
    exp_c1 = cjtest.result_row_hashes(oracle_c1)
    phase_total = 0
    for words, (kind, case, out) in zip(res, expect):
        f = dict(zip(words[2::2], words[3::2]))
        rows = int(f["rows"])
        assert f["stable"] == "1", words  # pages unchanged until the next execute()
        assert f["empty_ok"] == "1", words
        ti = words.index("timing")
        build, probe, nl, mat, setup, total, inter = (int(x) for x in words[ti + 1:ti + 8])
        assert min(build, probe, nl, mat, setup, total, inter) >= 0
        assert build + probe + nl + mat + setup + inter <= total, words
        phase_total += build + probe + nl + mat + setup + inter
        r = TblResult(out)
        try:
            assert r.num_rows == rows
            if kind == "kat":
                cjtest.check_kat(case, r)
                if rows == 0:
                    assert f["pages"] == "-" or all(x == "0" for x in f["pages"].split(","))
            else:
                assert rows == oracle_c1.num_rows
                assert np.array_equal(cjtest.result_row_hashes(r), exp_c1)
        finally:
            r.close()
    assert phase_total > 0, "no phase of any plan took a millisecond"


def test_gpu_timing_phases_untruncated():
    """cj_execute's phase_ms (untruncated TimingStats phases) on a join whose
    inputs are not resident: every phase it runs is positive, the sum stays
    within the call's wall time, and the int64 fields are their floors."""
    b, p = workloads.c1_tables(seed=1001)
    plan = workloads.c1_plan(b, p)
    eng = cjtest.Engine()
    try:
        for _ in range(2):
            t0 = time.perf_counter()
            r = eng.execute(plan, timing=True)
            wall = (time.perf_counter() - t0) * 1e3
            tm = r.timing
            ph = list(tm.phase_ms)
            assert ph[abi.CJ_PHASE_SETUP] > 0  # the pages' upload
            assert ph[abi.CJ_PHASE_HASHTABLE_BUILD] > 0 and ph[abi.CJ_PHASE_HASH_JOIN_PROBE] > 0
            assert ph[abi.CJ_PHASE_MATERIALIZE] > 0
            assert ph[abi.CJ_PHASE_NESTED_LOOP_JOIN] == 0 and ph[abi.CJ_PHASE_INTERMEDIATE] == 0
            assert sum(ph) <= wall + 0.5, (ph, wall)
            whole = [tm.hashtable_build_ms, tm.hash_join_probe_ms, tm.nested_loop_join_ms, tm.materialize_ms,
                     tm.setup_ms, tm.intermediate_ms]
            assert whole == [int(x) for x in ph]
            r.release()
        # a multi-join plan of 3-row tables: nested-loop joins, no hash build
        case = next(c for c in KATS if c["name"] == "leftdeep 2-level join")
        r = eng.execute(cjtest.kat_plan(case), timing=True)
        assert r.timing.phase_ms[abi.CJ_PHASE_NESTED_LOOP_JOIN] > 0
        assert r.timing.phase_ms[abi.CJ_PHASE_HASHTABLE_BUILD] == 0
        r.release()
    finally:
        eng.close()
