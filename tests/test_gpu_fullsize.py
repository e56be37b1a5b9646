"""Parity at the full sizes of the benchmark workloads (the smaller tests
elsewhere cover every strategy and edge case; these show that nothing
changes at scale):

* the single large join, 4M-row build x 64M-row Zipf probe, on the
  page-aligned and the hash-table strategies, against the oracle;
* the 4-join star over the 36M-row fact table, against the oracle;
* 256 batched uneven joins at full size: every plan's batched result equals
  its per-plan result, and 8 plans (the largest among them) equal the oracle;
* the served stream over 12 resident tables with 1-40 % filters: every plan
  served in the mix equals the plan run alone, and 4 plans equal the oracle.

Every result is checked structurally page by page (cjtest.validate_pages)."""
import os

import numpy as np
import pytest

import cjtest
from calamarijoins_amd import abi, workloads

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 4)


def log(msg):
    print(msg, flush=True)


def hashes(r):
    return cjtest.result_row_hashes(r)


@pytest.fixture(scope="module")
def c2_full():
    b, p = workloads.c2_tables()
    plan = workloads.c2_plan(b, p)
    exp = cjtest.OracleRun(plan, threads=THREADS)
    exp_h = hashes(exp)
    n = exp.num_rows
    del exp
    pk = cjtest.decode_column_arrays(abi.CJ_INT32, p.columns[0].page_ptrs(), p.columns[0].num_pages)[0]
    non_miss = int(np.sum(pk < workloads.C2_BUILD))
    return b, p, plan, n, exp_h, non_miss


@pytest.mark.parametrize("flags", [0, abi.CJ_EXEC_NO_DIRECT], ids=["aligned", "fused-hash"])
def test_gpu_c2_full_size_bit_exact(engine, c2_full, flags):
    b, p, plan, n, exp_h, non_miss = c2_full
    assert n == non_miss  # unique build keys: one output row per non-miss probe row
    log(f"C2 4M x 64M flags={flags}: {n} rows expected")
    got = engine.execute(plan, flags=flags)
    assert got.num_rows == n
    assert np.array_equal(hashes(got), exp_h)
    got.release()


def test_gpu_c2_full_size_resident_device_output(engine, c2_full):
    """The benchmark's form: tables resident in HBM, result pages in HBM."""
    b, p, plan, n, exp_h, _ = c2_full
    ids = [engine.register(b), engine.register(p)]
    try:
        got = engine.execute(plan, flags=abi.CJ_EXEC_OUTPUT_DEVICE)
        got.to_host()
        assert got.num_rows == n and np.array_equal(hashes(got), exp_h)
        got.release()
    finally:
        for t in ids:
            engine.unregister(t)


def test_gpu_c3_full_size_bit_exact(engine):
    f, ds, _ = workloads.c3_tables()
    plan = workloads.c3_plan(f, ds)
    exp = cjtest.OracleRun(plan, threads=THREADS)
    log(f"C3 36M star: {exp.num_rows} rows expected")
    got = engine.execute(plan)
    assert got.num_rows == exp.num_rows > 0
    assert np.array_equal(hashes(got), hashes(exp))
    got.release()


def test_gpu_c4_full_size_batch(engine):
    pairs = []
    for k in range(256):
        pairs.append(workloads.c4_plan_tables(k))
        if k % 32 == 31:
            log(f"C4 tables: {k + 1} of 256 plans generated")
    plans = [workloads.c4_plan(b, p) for b, p in pairs]
    ids = []
    try:
        for b, p in pairs:
            ids += [engine.register(b), engine.register(p)]
        alone = []
        for k, pl in enumerate(plans):
            r = engine.execute(pl)
            alone.append(cjtest.result_fingerprint(r))
            r.release()
            if k % 32 == 31:
                log(f"C4 per plan: {k + 1} of 256 run")
        log(f"C4 256 plans: {sum(a[0] for a in alone)} rows alone")
        res, st = engine.execute_batch(plans, abi.CJ_EXEC_OUTPUT_DEVICE)
        assert st.plans_batched + st.plans_single == 256
        assert st.plans_batched >= 250
        for k, r in enumerate(res):
            r.to_host()
            assert cjtest.result_fingerprint(r) == alone[k], k
            r.release()
            if k % 32 == 31:
                log(f"C4 batched: {k + 1} of 256 checked")
        # the oracle pins 8 plans, the largest probe side among them
        sizes = [p.num_rows for _, p in pairs]
        pick = sorted(set([int(np.argmax(sizes))] + list(range(0, 256, 37))))[:8]
        for k in pick:
            exp = cjtest.OracleRun(plans[k], threads=THREADS)
            assert (exp.num_rows,) + cjtest.result_fingerprint(exp)[1:] == alone[k], k
    finally:
        for t in ids:
            engine.unregister(t)


def test_gpu_c5_full_size_served(engine):
    tables, meta = workloads.c5_tables()  # scale 1, 1-40 % of each table's rows kept
    plans = [workloads.c5_plan(tables, meta, 5000 + k) for k in range(48)]
    ids = [engine.register(t) for t in tables]
    try:
        alone = []
        for pl in plans:
            r = engine.execute(pl)
            alone.append((r.num_rows, hashes(r)))
            r.release()
        log(f"C5 48 plans: {sum(a[0] for a in alone)} rows alone")
        assert sum(a[0] for a in alone) > 0
        for k in range(4):
            exp = cjtest.OracleRun(plans[k], threads=THREADS)
            assert exp.num_rows == alone[k][0] and np.array_equal(hashes(exp), alone[k][1]), k
        engine.serve_start(4, 8 << 30)
        try:
            order = list(range(len(plans))) * 2
            tickets = [(k, engine.submit(plans[k])) for k in order]
            for k, t in tickets:
                r = engine.wait(t)
                assert r.num_rows == alone[k][0], k
                assert np.array_equal(hashes(r), alone[k][1]), k
                r.release()
        finally:
            engine.serve_stop()
    finally:
        for t in ids:
            engine.unregister(t)
