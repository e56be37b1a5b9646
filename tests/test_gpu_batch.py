"""Batched execution (cj_execute_batch): every plan's result equals its
per-plan cj_execute result (row multiset), including plans that run one at a
time (non-resident inputs, nested-loop builds), a repeated batch that
replays the phase graph, and the autotuned configuration.  The first plans
are pinned to the oracle."""
import numpy as np
import pytest

import cjtest
from calamarijoins_amd import abi, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c4_small(engine):
    pairs = [workloads.c4_plan_tables(k, scale=0.01) for k in range(24)]
    ids = []
    for b, p in pairs:
        ids += [engine.register(b), engine.register(p)]
    plans = [workloads.c4_plan(b, p) for b, p in pairs]
    alone = []
    for pl in plans:
        r = engine.execute(pl)
        alone.append((r.num_rows, cjtest.result_row_hashes(r)))
        r.release()
    yield pairs, plans, alone
    for t in ids:
        engine.unregister(t)


def check(results, alone, idx):
    for r, k in zip(results, idx):
        assert r.num_rows == alone[k][0], (k, r.num_rows, alone[k][0])
        assert np.array_equal(cjtest.result_row_hashes(r), alone[k][1]), k
        r.release()


def test_gpu_batch_oracle_pins_per_plan(c4_small):
    _, plans, alone = c4_small
    assert sum(n for n, _ in alone) > 0
    for k in range(3):
        exp = cjtest.OracleRun(plans[k], threads=8)
        assert alone[k][0] == exp.num_rows
        assert np.array_equal(alone[k][1], cjtest.result_row_hashes(exp))


def test_gpu_batch_equals_per_plan(engine, c4_small):
    _, plans, alone = c4_small
    res, st = engine.execute_batch(plans)
    assert st.plans_batched == len(plans) and st.plans_single == 0
    assert st.output_rows == sum(n for n, _ in alone)
    check(res, alone, range(len(plans)))
    # same batch again: the phase graph replays, results unchanged
    res, st = engine.execute_batch(plans)
    assert st.graph_replayed == 1
    check(res, alone, range(len(plans)))


def test_gpu_batch_device_output_and_order(engine, c4_small):
    _, plans, alone = c4_small
    idx = list(range(len(plans)))[::-1][:10]
    res, st = engine.execute_batch([plans[k] for k in idx], abi.CJ_EXEC_OUTPUT_DEVICE)
    assert st.plans_batched == len(idx)
    for r in res:
        r.to_host()
    check(res, alone, idx)


def test_gpu_batch_mixed_with_single_plans(engine, c4_small):
    """A non-resident input and a nested-loop build run one at a time
    inside the same call."""
    _, plans, alone = c4_small
    b, p = workloads.c2_tables(seed=93, build_rows=5_000, probe_rows=40_000)
    c2 = workloads.c2_plan(b, p)
    nl_b = cjtest.table_from_rows(["INT32", "INT32"], [[i, 10 * i] for i in range(5)])
    nl_p = cjtest.table_from_rows(["INT32", "INT32"], [[i % 7, i] for i in range(100)])
    nl = workloads.c4_plan(nl_b, nl_p)
    extra = []
    for pl in (c2, nl):
        r = engine.execute(pl)
        extra.append((r.num_rows, cjtest.result_row_hashes(r)))
        r.release()
    assert extra[1][0] == sum(1 for i in range(100) if i % 7 < 5)
    all_alone = alone[:6] + extra
    mix = plans[:3] + [c2] + plans[3:6] + [nl]
    order = [0, 1, 2, 6, 3, 4, 5, 7]
    res, st = engine.execute_batch(mix)
    assert st.plans_batched == 6 and st.plans_single == 2
    check(res, all_alone, order)


def test_gpu_batch_autotune(engine, c4_small):
    _, plans, alone = c4_small
    sub = plans[4:20]
    res, st = engine.execute_batch(sub, abi.CJ_BATCH_AUTOTUNE)
    # the tuner runs every configuration (and fails the call if their result
    # sizes disagree); the chosen one's results equal the per-plan results
    assert st.tuned == 1 and st.load_factor in (2, 4) and st.grid > 0
    assert st.log_part in (11, 12) and st.chunk in (4096, 8192, 16384) and st.join_queue in (0, 1)
    check(res, alone, range(4, 20))
    cfg = (st.load_factor, st.log_part, st.chunk, st.join_queue)
    res, st = engine.execute_batch(sub, abi.CJ_BATCH_AUTOTUNE)
    assert st.tuned == 0 and (st.load_factor, st.log_part, st.chunk, st.join_queue) == cfg
    check(res, alone, range(4, 20))


@pytest.mark.parametrize("cfg", [(2, 8, 1000, 0, 0), (4, 11, 4096, 1, 0), (2, 12, 16384, 0, 300), (4, 10, 2047, 1, 7)],
                         ids=["lp8-tile1000", "lp11-queue", "lp12-grid300", "lp10-grid7-queue"])
def test_gpu_batch_pinned_configs(engine, c4_small, cfg):
    """Every configuration gives the per-plan results: radix bits 8..12,
    tiles that are not a multiple of the probe tile, strided or queued join
    units, resident or small explicit grids."""
    _, plans, alone = c4_small
    engine.batch_config(*cfg)
    try:
        res, st = engine.execute_batch(plans, abi.CJ_BATCH_AUTOTUNE)  # a pinned configuration is not tuned
        assert st.tuned == 0 and (st.load_factor, st.log_part, st.chunk, st.join_queue) == cfg[:4]
        assert st.grid == (cfg[4] or st.grid) and st.grid > 0
        check(res, alone, range(len(plans)))
    finally:
        engine.batch_config()


def test_gpu_batch_config_rejected(engine):
    for bad in [(3, 12, 8192, 0, 0), (2, 13, 8192, 0, 0), (2, 7, 8192, 0, 0), (2, 12, 0, 0, 0), (2, 12, 16385, 0, 0),
                (2, 12, 8192, 2, 0), (2, 12, 8192, 0, 65537)]:
        with pytest.raises(abi.CJError):
            engine.batch_config(*bad)


def test_gpu_batch_empty(engine):
    res, st = engine.execute_batch([])
    assert res == [] and st.plans_batched == 0 and st.plans_single == 0


@pytest.fixture(scope="module")
def wide_plans(engine):
    """Plans whose outputs carry 8-byte columns with NULLs on both sides, key
    columns of either side, a column twice, a build partition too large for
    LDS (two keys with 8000 rows each), and a plan with more output columns
    than the batch writes (it runs one at a time)."""
    b, p = workloads.c2_tables(seed=95, build_rows=60_000, probe_rows=500_000)
    c2 = workloads.c2_plan(b, p)
    # keys of both sides, 8-byte columns of both sides, a column twice, build_left=False
    mix = cjtest.Plan()
    mix.new_input(p)
    mix.new_input(b)
    mix.new_scan_node(0, [(3, "FP64"), (0, "INT32"), (2, "INT64")])
    mix.new_scan_node(1, [(2, "INT64"), (0, "INT32"), (1, "INT32")])
    mix.root = mix.new_join_node(False, 0, 1, 1, 1, [(4, "INT32"), (1, "INT32"), (0, "FP64"), (3, "INT64"),
                                                      (5, "INT32"), (2, "INT64"), (0, "FP64")])
    # duplicate-heavy build: key i % 2 over 16000 rows (partitions above the LDS capacity)
    dup_b = workloads.generate(16_000, 77, [workloads.col(workloads.GEN_UNIFORM, "INT32", 1, 0.0, lo=0, hi=2),
                                            workloads.col(workloads.GEN_I64_FULL, "INT64", 2, 0.1)])
    dup_p = workloads.generate(20_000, 78, [workloads.col(workloads.GEN_UNIFORM, "INT32", 3, 0.1, lo=0, hi=500),
                                        workloads.col(workloads.GEN_I32_FULL, "INT32", 4, 0.1)])
    dup = cjtest.Plan()
    dup.new_input(dup_b)
    dup.new_input(dup_p)
    dup.new_scan_node(0, [(0, "INT32"), (1, "INT64")])
    dup.new_scan_node(1, [(0, "INT32"), (1, "INT32")])
    dup.root = dup.new_join_node(True, 0, 1, 0, 0, [(1, "INT64"), (3, "INT32"), (2, "INT32")])
    # nine output columns: more than the batch writes (runs one at a time)
    nine = cjtest.Plan()
    nine.new_input(b)
    nine.new_input(p)
    nine.new_scan_node(0, [(0, "INT32")])
    nine.new_scan_node(1, [(0, "INT32"), (1, "INT32"), (2, "INT64"), (3, "FP64")])
    nine.root = nine.new_join_node(True, 0, 1, 0, 0, [(1, "INT32"), (2, "INT32"), (3, "INT64"), (4, "FP64"),
                                                      (2, "INT32"), (3, "INT64"), (4, "FP64"), (1, "INT32"),
                                                      (0, "INT32")])
    ids = [engine.register(t) for t in (b, p, dup_b, dup_p)]
    plans = [c2, mix, dup, nine]
    alone = []
    for pl in plans:
        r = engine.execute(pl)
        alone.append((r.num_rows, cjtest.result_row_hashes(r)))
        r.release()
    yield plans, alone
    for t in ids:
        engine.unregister(t)


def test_gpu_batch_wide_and_duplicate_plans(engine, wide_plans):
    plans, alone = wide_plans
    assert alone[2][0] > 100_000  # every probe key 0 or 1 meets about 8000 build rows
    for k in (1, 2):
        exp = cjtest.OracleRun(plans[k], threads=8)
        assert alone[k][0] == exp.num_rows
        assert np.array_equal(alone[k][1], cjtest.result_row_hashes(exp)), k
    res, st = engine.execute_batch(plans)
    assert st.plans_batched == 3 and st.plans_single == 1
    check(res, alone, range(len(plans)))
    res, st = engine.execute_batch(plans, abi.CJ_EXEC_OUTPUT_DEVICE)
    for r in res:
        r.to_host()
    check(res, alone, range(len(plans)))


def _carried_plan(seed, null_frac):
    """A one-join plan whose build side carries an INT64 column with
    `null_frac` NULLs: the same row counts give other page counts."""
    b = workloads.generate(30_000, seed, [workloads.col(workloads.GEN_UNIFORM, "INT32", 1, 0.0, lo=0, hi=20_000),
                                          workloads.col(workloads.GEN_I64_FULL, "INT64", 2, null_frac)])
    p = workloads.generate(90_000, seed + 1, [workloads.col(workloads.GEN_UNIFORM, "INT32", 3, 0.0, lo=0, hi=20_000),
                                              workloads.col(workloads.GEN_I32_FULL, "INT32", 4, 0.05)])
    pl = cjtest.Plan()
    pl.new_input(b)
    pl.new_input(p)
    pl.new_scan_node(0, [(0, "INT32"), (1, "INT64")])
    pl.new_scan_node(1, [(0, "INT32"), (1, "INT32")])
    pl.root = pl.new_join_node(True, 0, 1, 0, 0, [(1, "INT64"), (3, "INT32"), (0, "INT32")])
    return b, p, pl


def test_gpu_batch_template_carried_page_counts(engine):
    """Two batches with equal row counts whose carried INT64 column has a
    different NULL density (so a different page count) must not share a
    captured template: each result equals its per-plan result."""
    made = [_carried_plan(501, 0.0), _carried_plan(503, 0.6)]
    assert made[0][0].num_rows == made[1][0].num_rows and made[0][1].num_rows == made[1][1].num_rows
    assert made[0][0].columns[1].num_pages != made[1][0].columns[1].num_pages
    ids = []
    try:
        for b, p, _ in made:
            ids += [engine.register(b), engine.register(p)]
        for rep in range(2):
            for b, p, pl in (made if rep == 0 else made[::-1]):
                r = engine.execute(pl)
                alone = [(r.num_rows, cjtest.result_row_hashes(r))]
                r.release()
                res, st = engine.execute_batch([pl])
                assert st.plans_batched == 1
                check(res, alone, [0])
    finally:
        for t in ids:
            engine.unregister(t)


def test_gpu_batch_config_beyond_lds_fails_cleanly(engine):
    """2^8 slots a partition over a 4M-row build side (load factor 4) is 65536
    partitions: the scatter's per-partition LDS arrays cannot fit a
    workgroup, so the call fails before any launch; the tuned configurations
    then run the same batch."""
    b = workloads.generate(4_000_000, 61, [workloads.col(workloads.GEN_UNIFORM, "INT32", 1, 0.0, lo=0, hi=4_000_000),
                                           workloads.col(workloads.GEN_I32_FULL, "INT32", 2)])
    # the smaller side builds: the probe side must be the larger one
    p = workloads.generate(5_000_000, 62, [workloads.col(workloads.GEN_UNIFORM, "INT32", 3, 0.0, lo=0, hi=4_000_000),
                                        workloads.col(workloads.GEN_I32_FULL, "INT32", 4)])
    plan = workloads.c4_plan(b, p)
    ids = [engine.register(b), engine.register(p)]
    try:
        exp = engine.execute(plan)
        want = cjtest.result_fingerprint(exp)
        exp.release()
        engine.batch_config(4, 8, 8192, 0, 0)
        try:
            in_use = engine.memory()["in_use"]
            with pytest.raises(abi.CJError, match="LDS"):
                engine.execute_batch([plan])
            assert engine.memory()["in_use"] == in_use  # the failed call's workspace was released
        finally:
            engine.batch_config()
        res, st = engine.execute_batch([plan])
        assert st.plans_batched == 1 and cjtest.result_fingerprint(res[0]) == want
        res[0].release()
    finally:
        for t in ids:
            engine.unregister(t)
