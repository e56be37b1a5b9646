"""Serving loop (cj_serve_start / cj_submit / cj_wait): plans run concurrently
in the mix give the same row multisets as each plan run alone, and the
first plans are checked against the oracle."""
import numpy as np
import pytest

import cjtest
from calamarijoins_amd import abi, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c5_small():
    tables, meta = workloads.c5_tables(scale=0.01, keep_lo=0.5, keep_hi=0.95)
    plans = [workloads.c5_plan(tables, meta, 5000 + k) for k in range(12)]
    b, p = workloads.c2_tables(seed=91, build_rows=100_000, probe_rows=800_000)
    plans.append(workloads.c2_plan(b, p))
    return tables, plans, (b, p)


def hashes(res):
    return cjtest.result_row_hashes(res)


def test_gpu_served_results_equal_alone(engine, c5_small):
    tables, plans, _ = c5_small
    alone = []
    for pl in plans:
        r = engine.execute(pl)
        alone.append((r.num_rows, hashes(r)))
        r.release()
    assert sum(n for n, _ in alone) > 0
    for k in range(3):  # the oracle pins the first plans
        exp = cjtest.OracleRun(plans[k], threads=8)
        assert alone[k][0] == exp.num_rows
        assert np.array_equal(alone[k][1], cjtest.result_row_hashes(exp))
    engine.serve_start(4, 64 << 30)
    try:
        order = [k % len(plans) for k in range(2 * len(plans))]
        order = order[::2] + order[1::2]
        tickets = [(k, engine.submit(plans[k])) for k in order]
        for k, t in tickets:
            r = engine.wait(t)
            assert r.num_rows == alone[k][0], (k, r.num_rows, alone[k][0])
            assert np.array_equal(hashes(r), alone[k][1]), k
            assert r.timing.service_ms > 0 and r.timing.queue_ms >= 0
            r.release()
    finally:
        engine.serve_stop()


def test_gpu_served_tiny_budget_runs_one_at_a_time(engine, c5_small):
    """A budget below every plan's footprint still admits one plan when
    nothing runs, so the stream completes."""
    _, plans, _ = c5_small
    engine.serve_start(4, 1)
    try:
        ts = [(k, engine.submit(plans[k], abi.CJ_EXEC_OUTPUT_DEVICE)) for k in range(4)]
        for k, t in ts:
            r = engine.wait(t)
            r.release()
    finally:
        engine.serve_stop()


def test_gpu_served_error_and_timeout(engine, c5_small):
    tables, plans, _ = c5_small
    bad = cjtest.Plan()
    t = cjtest.table_from_rows(["INT32", "INT64"], [[1, 2]])
    bad.new_input(t)
    bad.new_scan_node(0, [(0, "INT32"), (1, "INT64")])
    bad.new_scan_node(0, [(0, "INT32"), (1, "INT64")])
    bad.root = bad.new_join_node(True, 0, 1, 1, 1, [(0, "INT32")])  # INT64 key
    engine.serve_start(2, 64 << 30)
    try:
        tb = engine.submit(bad)
        with pytest.raises(abi.CJError):
            engine.wait(tb)
        tg = engine.submit(plans[0])
        r = None
        while r is None:
            r = engine.wait(tg, timeout_ms=50)
        r.release()
    finally:
        engine.serve_stop()


def test_gpu_served_out_of_memory_runs_alone():
    """A device budget with room for the tables and one and a half plans:
    plans admitted together run out of memory, are queued again and run
    alone (nothing else admitted meanwhile), and every result equals the
    plan's result run alone.  (Results kept in device memory by the caller
    count against the budget too; these plans return host pages.)"""
    b, p = workloads.c2_tables(seed=97, build_rows=1_000_000, probe_rows=12_000_000)
    plan = workloads.c2_plan(b, p)
    probe = cjtest.Engine()
    try:
        for t in (b, p):
            probe.register(t)
        tables = probe.memory(reset_peak=True)["in_use"]
        r = probe.execute(plan, flags=abi.CJ_EXEC_OUTPUT_DEVICE)
        peak = probe.memory()["peak_in_use"]
        r.to_host()
        want = (r.num_rows, hashes(r))
        r.release()
    finally:
        probe.close()
    plan_bytes = peak - tables
    assert plan_bytes > (64 << 20)
    eng = cjtest.Engine(device_budget=tables + plan_bytes * 3 // 2 + (32 << 20), num_streams=4)
    try:
        for t in (b, p):
            eng.register(t)
        eng.serve_start(4, 1 << 40)  # admission lets all four in: the device budget is what binds
        try:
            # host output: a finished plan's pages leave the device, so results
            # not yet collected do not hold the budget
            tickets = [eng.submit(plan, 0) for _ in range(8)]
            alone = 0
            for t in tickets:
                r = eng.wait(t)
                assert (r.num_rows, None) == (want[0], None) and np.array_equal(hashes(r), want[1])
                alone += r.timing.served_alone
                r.release()
            st = eng.serve_stats()
        finally:
            eng.serve_stop()
        assert alone >= 1 and st["retried_alone"] == alone, (alone, st)
        assert st["peak_running"] >= 2
    finally:
        eng.close()


def test_gpu_served_budget_binds(engine, c5_small):
    """A budget of about two plan footprints: plans wait for admission, at
    most as many run at once as fit, and results are unchanged."""
    tables, plans, _ = c5_small
    foot = sorted(engine.footprint(pl, engine) for pl in plans)
    budget = 2 * foot[len(foot) // 2]
    alone = []
    for pl in plans:
        r = engine.execute(pl)
        alone.append((r.num_rows, hashes(r)))
        r.release()
    engine.serve_start(4, budget)
    try:
        ts = [(k % len(plans), engine.submit(plans[k % len(plans)])) for k in range(3 * len(plans))]
        for k, t in ts:
            r = engine.wait(t)
            assert r.num_rows == alone[k][0] and np.array_equal(hashes(r), alone[k][1]), k
            r.release()
        st = engine.serve_stats()
    finally:
        engine.serve_stop()
    assert st["budget_waits"] > 0 and st["peak_admitted_bytes"] <= max(budget, foot[-1]), (st, budget)


def test_gpu_learned_footprint():
    """A plan's footprint charge starts from the bound on its shape and, once
    the plan has run on the context, falls to its recorded peak plus a
    quarter: never below the device memory the run actually held."""
    tables, meta = workloads.c5_tables(scale=0.01, keep_lo=0.5, keep_hi=0.95)
    plans = [workloads.c5_plan(tables, meta, 5000 + k) for k in range(4)]
    eng = cjtest.Engine(cache_decoded=0)
    try:
        for tb in tables:
            eng.register(tb)
        for pl in plans:
            shape = eng.footprint(pl, eng)
            base = eng.memory(reset_peak=True)["in_use"]
            r = eng.execute(pl, flags=abi.CJ_EXEC_OUTPUT_DEVICE)
            held = eng.memory()["peak_in_use"] - base
            r.release()
            charged = eng.footprint(pl, eng)
            assert charged <= shape, (charged, shape)
            assert charged >= held, (charged, held)
            # a repeat keeps the charge (the recorded peak is the largest seen)
            eng.execute(pl, flags=abi.CJ_EXEC_OUTPUT_DEVICE).release()
            assert eng.footprint(pl, eng) >= charged
    finally:
        eng.close()
