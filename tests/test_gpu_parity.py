"""GPU parity: the HIP engine through the C ABI against the oracle, bit-exact
(the same multiset of output rows, NULLs and raw bytes included)."""
import json
import os
import random

import numpy as np
import pytest

import bruteforce
import cjtest
import test_oracle as to
from calamarijoins_amd import abi, workloads

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KATS = json.load(open(os.path.join(GOLDEN, "unit_kats.json")))["cases"]


def same_multiset(got, exp):
    assert got.num_rows == exp.num_rows, (got.num_rows, exp.num_rows)
    assert got.types == exp.types
    assert np.array_equal(cjtest.result_row_hashes(got), cjtest.result_row_hashes(exp))


@pytest.mark.parametrize("flags", [0, 8, 4, 2], ids=["aligned", "fused", "fused-hash", "general"])
@pytest.mark.parametrize("case", KATS, ids=[c["name"] for c in KATS])
def test_gpu_unit_kat(engine, case, flags):
    res = engine.execute(cjtest.kat_plan(case), flags=flags)
    cjtest.check_kat(case, res)


PATHS = [0, abi.CJ_EXEC_NO_ALIGNED, abi.CJ_EXEC_NO_DIRECT, abi.CJ_EXEC_NO_FAST_PATH]
PATH_IDS = ["aligned", "fused", "fused-hash", "general"]


@pytest.mark.parametrize("flags", PATHS, ids=PATH_IDS)
@pytest.mark.parametrize("seed", range(40))
def test_gpu_random_plans_vs_bruteforce(engine, seed, flags):
    rng = random.Random(1000 + seed)
    njoins = 1 + seed % 3
    tables = []
    for _ in range(njoins + 1):
        nrows = 0 if rng.random() < 0.04 else rng.choice([1, 3, 5, 7, 8, 9, 20, 60, 150])
        tables.append(to.rand_table(rng, nrows, key_domain=rng.choice([2, 4, 8, 16])))
    plan = to.rand_plan(rng, njoins, tables)
    exp = bruteforce.evaluate(plan.nodes, tables, plan.root)
    got = engine.execute(plan, flags=flags)
    assert got.num_rows == len(exp)
    assert to.norm_rows(cjtest.result_rows(got)) == to.norm_rows(exp)


@pytest.mark.parametrize("flags", PATHS, ids=PATH_IDS)
@pytest.mark.parametrize("seed", range(8))
def test_gpu_random_plans_vs_oracle_larger(engine, seed, flags):
    rng = random.Random(5000 + seed)
    njoins = 1 + seed % 3
    tables = []
    for _ in range(njoins + 1):
        nrows = rng.choice([500, 3000, 20000])
        tables.append(to.rand_table(rng, nrows, key_domain=rng.choice([100, 2000, 50000])))
    plan = to.rand_plan(rng, njoins, tables)
    same_multiset(engine.execute(plan, flags=flags), cjtest.OracleRun(plan, threads=8))


@pytest.mark.parametrize("flags", PATHS, ids=PATH_IDS)
def test_gpu_c1_config_bit_exact(engine, flags):
    b, p = workloads.c1_tables()
    plan = workloads.c1_plan(b, p)
    same_multiset(engine.execute(plan, flags=flags), cjtest.OracleRun(plan, threads=8))


@pytest.mark.parametrize("flags", PATHS, ids=PATH_IDS)
def test_gpu_c2_scaled_bit_exact(engine, flags):
    b, p = workloads.c2_tables(build_rows=200_000, probe_rows=3_200_000)
    plan = workloads.c2_plan(b, p)
    got = engine.execute(plan, flags=flags)
    exp = cjtest.OracleRun(plan, threads=8)
    same_multiset(got, exp)
    # each output row's payloads come from rows whose keys matched: all 5
    # columns present, and the row count equals the non-miss probe rows
    pk = cjtest.decode_column_arrays(abi.CJ_INT32, p.columns[0].page_ptrs(), p.columns[0].num_pages)[0]
    assert got.num_rows == int(np.sum(pk < 200_000))


def test_gpu_repeated_resident_plans_reuse_key_statistics(engine):
    """A plan over registered tables run again: the direct maps' key
    statistics and the page-aligned join's match count are taken from the
    first run (no read-back); every run equals the oracle, on the aligned
    path (C2 shape) and through a multi-join plan (C3 shape, direct maps of
    row ids over the dimension keys)."""
    b, p = workloads.c2_tables(seed=77, build_rows=150_000, probe_rows=2_000_000)
    c2 = workloads.c2_plan(b, p)
    f, ds, _ = workloads.c3_tables(fact_rows=workloads.C3_FACT // 100, dims=tuple(d // 100 for d in workloads.C3_DIMS))
    c3 = workloads.c3_plan(f, ds)
    ids = [engine.register(t) for t in [b, p, f] + list(ds)]
    try:
        for plan in (c2, c3):
            exp = cjtest.OracleRun(plan, threads=8)
            for _ in range(3):
                same_multiset(engine.execute(plan), exp)
                r = engine.execute(plan, flags=abi.CJ_EXEC_OUTPUT_DEVICE)
                r.to_host()
                same_multiset(r, exp)
    finally:
        for t in ids:
            engine.unregister(t)


def test_gpu_registered_key_statistics_recount_for_a_wider_map(engine):
    """Key statistics recorded by a build whose direct map could not hold the
    key range (a root join: a map of records, at most ~2x the build rows)
    carry no duplicate count; a later build with a wider map (a non-root
    join: a map of row ids) over the same registered key column counts the
    keys itself, finds the duplicates and takes the hash table.  Both plans,
    run twice each, equal the oracle."""
    rng = random.Random(777)
    keys = rng.sample(range(0, 20000), 3800)
    keys += [rng.choice(keys) for _ in range(200)]  # duplicates
    keys += [0, 19999]
    rng.shuffle(keys)
    d_rows = [[k, rng.randrange(-500, 500)] for k in keys]
    p_rows = [[rng.randrange(0, 20000), i] for i in range(9000)]
    e_rows = [[i, rng.randrange(100)] for i in range(9000)]
    d, p, e = (cjtest.table_from_rows(["INT32"] * 2, r) for r in (d_rows, p_rows, e_rows))
    root = cjtest.Plan()
    for t in (p, d):
        root.new_input(t)
    root.new_scan_node(0, [(0, "INT32"), (1, "INT32")])
    root.new_scan_node(1, [(0, "INT32"), (1, "INT32")])
    root.root = root.new_join_node(False, 0, 1, 0, 0, [(1, "INT32"), (3, "INT32")])
    two = cjtest.Plan()
    for t in (p, d, e):
        two.new_input(t)
    two.new_scan_node(0, [(0, "INT32"), (1, "INT32")])
    two.new_scan_node(1, [(0, "INT32"), (1, "INT32")])
    two.new_scan_node(2, [(0, "INT32"), (1, "INT32")])
    j1 = two.new_join_node(False, 0, 1, 0, 0, [(1, "INT32"), (3, "INT32")])
    two.root = two.new_join_node(True, j1, 2, 0, 0, [(0, "INT32"), (1, "INT32"), (3, "INT32")])
    ids = [engine.register(t) for t in (d, p, e)]
    try:
        for plan in (root, two, root, two):
            exp = cjtest.OracleRun(plan, threads=8)
            assert exp.num_rows > 1000
            same_multiset(engine.execute(plan), exp)
    finally:
        for t in ids:
            engine.unregister(t)


@pytest.mark.parametrize("dim_rows", [1_000_000, 2_400_000], ids=["one-window", "two-windows"])
def test_gpu_count_pass_presence_bitmap_in_lds(engine, dim_rows):
    """A probe side of more than 4096 dense key pages against a direct map
    whose keys leave gaps in their range (30 % of a permuted domain kept):
    the counting pass takes the presence bitmap into LDS, in one window
    (1M-key range) or two (2.4M); a two-join plan whose first join builds
    that map equals the oracle."""
    fact = workloads.generate(8_600_000, 4242, [
        workloads.col(workloads.GEN_UNIFORM, "INT32", 20, 0.0, lo=0, hi=dim_rows),
        workloads.col(workloads.GEN_UNIFORM, "INT32", 21, 0.0, lo=0, hi=1000),
        workloads.col(workloads.GEN_UNIFORM, "INT32", 22, 0.0, lo=-50, hi=50)])
    dim = workloads.generate(dim_rows, 4243, [
        workloads.col(workloads.GEN_PERM, "INT32", 1, 0.0, domain=dim_rows, perm_stream=50),
        workloads.col(workloads.GEN_UNIFORM, "INT32", 2, 0.0, lo=0, hi=100)], keep_frac=0.3)
    small = workloads.generate(1000, 4244, [
        workloads.col(workloads.GEN_PERM, "INT32", 1, 0.0, domain=1000, perm_stream=51),
        workloads.col(workloads.GEN_UNIFORM, "INT32", 2, 0.0, lo=0, hi=7)], keep_frac=0.5)
    p = cjtest.Plan()
    for t in (fact, dim, small):
        p.new_input(t)
    f = p.new_scan_node(0, [(0, "INT32"), (1, "INT32"), (2, "INT32")])
    d = p.new_scan_node(1, [(0, "INT32"), (1, "INT32")])
    sm = p.new_scan_node(2, [(0, "INT32"), (1, "INT32")])
    j1 = p.new_join_node(False, f, d, 0, 0, [(1, "INT32"), (2, "INT32"), (4, "INT32")])
    p.root = p.new_join_node(False, j1, sm, 0, 0, [(1, "INT32"), (2, "INT32"), (4, "INT32")])
    exp = cjtest.OracleRun(p, threads=8)
    assert exp.num_rows > 10_000
    same_multiset(engine.execute(p), exp)


@pytest.mark.parametrize("scale", [100, 10])
def test_gpu_c3_star_pipeline_bit_exact(engine, scale):
    """Left-deep 4-join star with VARCHAR (incl. long strings) through
    late-materialized intermediates."""
    f, ds, _ = workloads.c3_tables(fact_rows=workloads.C3_FACT // scale,
                                   dims=tuple(d // scale for d in workloads.C3_DIMS))
    plan = workloads.c3_plan(f, ds)
    same_multiset(engine.execute(plan), cjtest.OracleRun(plan, threads=8))


def test_gpu_resident_tables_and_device_output(engine):
    b, p = workloads.c2_tables(seed=77, build_rows=50_000, probe_rows=400_000)
    plan = workloads.c2_plan(b, p)
    exp = cjtest.OracleRun(plan, threads=8)
    ib, ip = engine.register(b), engine.register(p)
    try:
        for _ in range(2):  # second run hits the decoded-column cache
            got = engine.execute(plan, flags=abi.CJ_EXEC_OUTPUT_DEVICE, timing=True)
            assert got.page_lists == [[]] * 5
            got.to_host()
            same_multiset(got, exp)
            assert got.timing.bytes_alg > 0 and got.timing.device_ms > 0
    finally:
        engine.unregister(ib)
        engine.unregister(ip)


def test_gpu_varchar_long_strings(engine):
    rng = random.Random(3)
    n = 4000
    keys = [rng.randrange(500) if rng.random() > 0.05 else None for _ in range(n)]
    strs = []
    for _ in range(n):
        r = rng.random()
        strs.append(None if r < 0.1 else ("Q" * rng.choice([8185, 8186, 9000, 20000]) if r < 0.11 else
                                          "".join(chr(rng.randrange(33, 127)) for _ in range(rng.randrange(0, 70)))))
    t1 = cjtest.table_from_rows(["INT32", "VARCHAR"], [[k, s] for k, s in zip(keys, strs)])
    t2 = cjtest.table_from_rows(["INT32", "VARCHAR"], [[k, s] for k, s in zip(keys[:700], strs[:700])])
    plan = cjtest.Plan()
    plan.new_input(t1)
    plan.new_input(t2)
    plan.new_scan_node(0, [(0, "INT32"), (1, "VARCHAR")])
    plan.new_scan_node(1, [(1, "VARCHAR"), (0, "INT32")])
    plan.root = plan.new_join_node(False, 0, 1, 0, 1, [(1, "VARCHAR"), (2, "VARCHAR"), (0, "INT32")])
    same_multiset(engine.execute(plan), cjtest.OracleRun(plan, threads=8))


def test_gpu_int32_min_quirks(engine):
    """INT32_MIN payloads come out NULL; through an intermediate, NULL and
    INT32_MIN keys behave as the reference's value_t sentinel."""
    I = bruteforce.I32_MIN
    rows_a = [[1, I], [I, 5], [None, 6], [2, 7]] + [[k, k] for k in range(10, 40)]
    rows_b = [[1, 1], [I, 2], [2, I], [None, 3]] + [[k, -k] for k in range(10, 60)]
    rows_c = [[I, 100], [1, 101], [5, 102]] + [[k, k] for k in range(10, 70)]
    tabs = [(["INT32", "INT32"], r) for r in (rows_a, rows_b, rows_c)]
    plan = cjtest.Plan()
    for t in tabs:
        plan.new_input(cjtest.table_from_rows(*t))
    for i in range(3):
        plan.new_scan_node(i, [(0, "INT32"), (1, "INT32")])
    j = plan.new_join_node(True, 0, 1, 0, 0, [(1, "INT32"), (2, "INT32"), (3, "INT32")])
    plan.root = plan.new_join_node(True, j, 2, 1, 0, [(0, "INT32"), (1, "INT32"), (2, "INT32"), (4, "INT32")])
    exp = bruteforce.evaluate(plan.nodes, tabs, plan.root)
    got = engine.execute(plan)
    assert to.norm_rows(cjtest.result_rows(got)) == to.norm_rows(exp)


def test_gpu_empty_inputs(engine):
    t0 = cjtest.table_from_rows(["INT32"], [])
    t1 = cjtest.table_from_rows(["INT32"], [[1], [2]])
    for a, bt in ((t0, t1), (t1, t0), (t0, t0)):
        plan = cjtest.Plan()
        plan.new_input(a)
        plan.new_input(bt)
        plan.new_scan_node(0, [(0, "INT32")])
        plan.new_scan_node(1, [(0, "INT32")])
        plan.root = plan.new_join_node(True, 0, 1, 0, 0, [(0, "INT32"), (1, "INT32")])
        got = engine.execute(plan)
        assert got.num_rows == 0 and got.types == [0, 0] and got.num_pages() == [0, 0]


def test_gpu_rejects_bad_plans(engine):
    t = cjtest.table_from_rows(["INT32", "INT64"], [[1, 2]])
    plan = cjtest.Plan()
    plan.new_input(t)
    plan.new_scan_node(0, [(0, "INT32"), (1, "INT64")])
    plan.new_scan_node(0, [(0, "INT32"), (1, "INT64")])
    plan.root = plan.new_join_node(True, 0, 1, 1, 1, [(0, "INT32")])  # INT64 key
    with pytest.raises(abi.CJError):
        engine.execute(plan)
    plan.nodes[-1]["left_attr"] = 0
    plan.nodes[-1]["right_attr"] = 0
    plan.nodes[-1]["attrs"] = [(1, "INT32")]  # declared type != source type
    with pytest.raises(abi.CJError):
        engine.execute(plan)


def test_gpu_rejects_page_bitmap_value_count_mismatch(engine):
    """A page with NULLs lists exactly its valid rows' values: a bitmap that
    marks one valid row fewer than the header's value count is rejected at
    upload (kernels that count a page's keys from its value array and ones
    that place them by the bitmap must see the same keys)."""
    rows = [[i if i % 3 else None, i] for i in range(300)]
    good = cjtest.table_from_rows(["INT32", "INT32"], rows)
    bad = cjtest.table_from_rows(["INT32", "INT32"], rows)
    page = bad.columns[0].pages[0]
    n = int(page[0]) | int(page[1]) << 8
    nv = int(page[2]) | int(page[3]) << 8
    assert nv < n
    bm0 = 8192 - (n + 7) // 8
    for r in range(n):  # clear the first valid row's bit
        if page[bm0 + r // 8] >> (r % 8) & 1:
            page[bm0 + r // 8] &= ~(1 << (r % 8)) & 0xFF
            break
    for t, ok in ((good, True), (bad, False)):
        plan = cjtest.Plan()
        plan.new_input(t)
        plan.new_input(t)
        plan.new_scan_node(0, [(0, "INT32"), (1, "INT32")])
        plan.new_scan_node(1, [(0, "INT32"), (1, "INT32")])
        plan.root = plan.new_join_node(True, 0, 1, 0, 0, [(1, "INT32"), (3, "INT32")])
        if ok:
            same_multiset(engine.execute(plan), cjtest.OracleRun(plan, threads=1))
        else:
            with pytest.raises(abi.CJError):
                engine.execute(plan)


@pytest.mark.parametrize("flags", PATHS, ids=PATH_IDS)
@pytest.mark.parametrize("key_null,pay_null", [(0.3, 0.95), (0.6, 0.97), (0.0, 0.5)],
                         ids=["sparse-payload", "sparse-keys", "half-null"])
def test_gpu_sparse_pages(engine, key_null, pay_null, flags):
    """Pages far beyond the dense row counts: 95-97 % NULL payloads give
    pages of 20k-40k rows, 60 % NULL keys give key pages of several 2048-row
    tiles; unique dense build keys take the direct-mapped build."""
    rng = random.Random(int(key_null * 100) * 7 + int(pay_null * 100))
    nb, npr = 3000, 60000
    perm = list(range(nb))
    rng.shuffle(perm)

    def val(p, lo, hi):
        return None if rng.random() < p else rng.randrange(lo, hi)

    brows = [[perm[i], val(pay_null, -2**31, 2**31), val(pay_null, -2**63, 2**63)] for i in range(nb)]
    prows = []
    for _ in range(npr):
        k = None if rng.random() < key_null else (rng.randrange(nb) if rng.random() < 0.85 else rng.randrange(nb, 10**6))
        prows.append([k, val(pay_null, -2**31, 2**31), val(0.0, -2**63, 2**63),
                      None if rng.random() < pay_null else rng.random() * 1e6])
    plan = cjtest.Plan()
    plan.new_input(cjtest.table_from_rows(["INT32", "INT32", "INT64"], brows))
    plan.new_input(cjtest.table_from_rows(["INT32", "INT32", "INT64", "FP64"], prows))
    plan.new_scan_node(0, [(0, "INT32"), (1, "INT32"), (2, "INT64")])
    plan.new_scan_node(1, [(0, "INT32"), (1, "INT32"), (2, "INT64"), (3, "FP64")])
    plan.root = plan.new_join_node(True, 0, 1, 0, 0, [(1, "INT32"), (2, "INT64"), (4, "INT32"), (5, "INT64"),
                                                     (6, "FP64"), (3, "INT32")])
    same_multiset(engine.execute(plan, flags=flags), cjtest.OracleRun(plan, threads=8))


@pytest.mark.parametrize("flags", [0, 8, 4, 2], ids=["aligned", "fused", "fused-hash", "general"])
@pytest.mark.parametrize("seed", range(2))
def test_gpu_intermediate_probes_direct_maps(engine, seed, flags):
    """Intermediates probing base tables with unique, sparse keys (direct maps
    of row ids, probed through the intermediate's match bits): one key column
    with NULLs and INT32_MIN (decoded, NULL semantics through the
    intermediate), one without NULLs (read in place from its pages), a
    non-root and a root join."""
    rng = random.Random(4100 + seed)
    I = bruteforce.I32_MIN
    n = 6000 + 1500 * seed
    p_rows = []
    for i in range(n):
        r = rng.random()
        k1 = None if r < 0.05 else (I if r < 0.08 else rng.randrange(-5, 9000))
        p_rows.append([None if rng.random() < 0.03 else rng.randrange(0, 3000), k1, rng.randrange(0, 4000), i])
    d_keys = list(range(3000))
    rng.shuffle(d_keys)
    e_keys = rng.sample(range(0, 9000), 2200)  # sparse: a map of row ids over 9000 values
    f_keys = rng.sample(range(0, 4000), 1700)
    tabs = [
        cjtest.table_from_rows(["INT32"] * 4, p_rows),
        cjtest.table_from_rows(["INT32", "INT32"], [[k, rng.randrange(-50, 50)] for k in d_keys]),
        cjtest.table_from_rows(["INT32", "INT32"], [[k, rng.randrange(-50, 50)] for k in e_keys]),
        cjtest.table_from_rows(["INT32", "INT32"], [[k, None if rng.random() < 0.1 else k] for k in f_keys]),
    ]
    plan = cjtest.Plan()
    for t in tabs:
        plan.new_input(t)
    plan.new_scan_node(0, [(0, "INT32"), (1, "INT32"), (2, "INT32"), (3, "INT32")])
    for i in range(1, 4):
        plan.new_scan_node(i, [(0, "INT32"), (1, "INT32")])
    # P.k0 = D.key -> (P.k1, P.k2, P.row, D.pay)
    j1 = plan.new_join_node(False, 0, 1, 0, 0, [(1, "INT32"), (2, "INT32"), (3, "INT32"), (5, "INT32")])
    # k1 = E.key -> (k2, row, D.pay, E.pay)
    j2 = plan.new_join_node(False, j1, 2, 0, 0, [(1, "INT32"), (2, "INT32"), (3, "INT32"), (5, "INT32")])
    # k2 = F.key -> (row, D.pay, E.pay, F.pay)
    plan.root = plan.new_join_node(False, j2, 3, 0, 0, [(1, "INT32"), (2, "INT32"), (3, "INT32"), (5, "INT32")])
    exp = cjtest.OracleRun(plan, threads=8)
    assert exp.num_rows > 500
    same_multiset(engine.execute(plan, flags=flags), exp)


@pytest.mark.parametrize("flags", [0, 2], ids=["aligned", "general"])
def test_gpu_keys_through_direct_map_rows(engine, flags):
    """A non-root join over a map of row ids keeps map slots for its build
    side; a later join keyed on a column of that side (probing, and then as
    the build side of the root) reads its row ids through the map."""
    rng = random.Random(4242)
    p_rows = [[rng.randrange(0, 5000), i] for i in range(9000)]
    d_keys = rng.sample(range(0, 20000), 4000)  # sparse unique keys: a map of row ids
    d_rows = [[k, None if rng.random() < 0.05 else rng.randrange(0, 700), rng.randrange(-9, 9)] for k in d_keys]
    e_keys = list(range(700))
    rng.shuffle(e_keys)
    e_rows = [[k, k * 3] for k in e_keys]
    f_rows = [[rng.randrange(0, 9000), rng.randrange(100)] for _ in range(30000)]
    tabs = [cjtest.table_from_rows(["INT32"] * 2, p_rows), cjtest.table_from_rows(["INT32"] * 3, d_rows),
            cjtest.table_from_rows(["INT32"] * 2, e_rows), cjtest.table_from_rows(["INT32"] * 2, f_rows)]
    plan = cjtest.Plan()
    for t in tabs:
        plan.new_input(t)
    plan.new_scan_node(0, [(0, "INT32"), (1, "INT32")])
    plan.new_scan_node(1, [(0, "INT32"), (1, "INT32"), (2, "INT32")])
    plan.new_scan_node(2, [(0, "INT32"), (1, "INT32")])
    plan.new_scan_node(3, [(0, "INT32"), (1, "INT32")])
    # P.k = D.key -> (P.row, D.fk, D.pay)
    j1 = plan.new_join_node(False, 0, 1, 0, 0, [(1, "INT32"), (3, "INT32"), (4, "INT32")])
    # D.fk = E.key -> (P.row, D.pay, E.val)
    j2 = plan.new_join_node(False, j1, 2, 1, 0, [(0, "INT32"), (2, "INT32"), (4, "INT32")])
    # F.rowref = P.row, the intermediate as the build side -> (P.row, D.pay, E.val, F.x)
    plan.root = plan.new_join_node(True, j2, 3, 0, 0, [(0, "INT32"), (1, "INT32"), (2, "INT32"), (4, "INT32")])
    exp = cjtest.OracleRun(plan, threads=8)
    assert exp.num_rows > 1000
    same_multiset(engine.execute(plan, flags=flags), exp)


def test_gpu_aligned_pages_without_matches(engine):
    """The page-aligned writers put payload page q at output page q and close
    up the pages whose rows all miss: a probe table whose first quarter of
    rows holds keys the build side lacks (whole payload pages without a
    match) while three quarters match."""
    nb, npr = 100_000, 400_000
    rng = np.random.default_rng(11)
    bkeys = rng.permutation(nb).astype(np.int32)
    pkeys = np.concatenate([200_000 + np.arange(npr // 4), rng.integers(0, nb, npr - npr // 4)]).astype(np.int32)
    ones = lambda n: np.ones(n, np.uint8)
    b = cjtest.Table(nb, [cjtest.Column(abi.CJ_INT32, cjtest.encode_fixed(abi.CJ_INT32, bkeys, ones(nb))),
                          cjtest.Column(abi.CJ_INT64, cjtest.encode_fixed(abi.CJ_INT64, bkeys.astype(np.int64) * 3,
                                                                          (rng.random(nb) > 0.05).astype(np.uint8)))])
    pv = rng.integers(-2**31, 2**31, npr).astype(np.int32)
    p = cjtest.Table(npr, [cjtest.Column(abi.CJ_INT32, cjtest.encode_fixed(abi.CJ_INT32, pkeys, ones(npr))),
                           cjtest.Column(abi.CJ_INT32, cjtest.encode_fixed(abi.CJ_INT32, pv,
                                                                           (rng.random(npr) > 0.05).astype(np.uint8))),
                           cjtest.Column(abi.CJ_INT64, cjtest.encode_fixed(abi.CJ_INT64, pv.astype(np.int64) << 20,
                                                                           ones(npr)))])
    plan = cjtest.Plan()
    plan.new_input(b)
    plan.new_input(p)
    plan.new_scan_node(0, [(0, "INT32"), (1, "INT64")])
    plan.new_scan_node(1, [(0, "INT32"), (1, "INT32"), (2, "INT64")])
    plan.root = plan.new_join_node(True, 0, 1, 0, 0, [(1, "INT64"), (3, "INT32"), (4, "INT64"), (2, "INT32")])
    got = engine.execute(plan, timing=True)
    exp = cjtest.OracleRun(plan, threads=8)
    assert got.num_rows == npr - npr // 4
    same_multiset(got, exp)
    # fewer probe-side pages than payload pages: the empty ones were closed up
    assert got.num_pages()[1] < p.columns[1].num_pages


@pytest.mark.parametrize("build_rows,log_domain", [(17_000_000, 28), (70_000_000, 30)], ids=["above_2_24", "above_2_26"])
def test_gpu_large_hash_builds(engine, build_rows, log_domain):
    """Hash builds of more than 2^24 and more than 2^26 rows (the second
    counts each radix partition in two LDS passes): uniform keys over a
    domain too wide for a direct map, probed by a slightly larger table (the
    smaller side builds), against the oracle."""
    b = workloads.generate(build_rows, 1701, [workloads.col(workloads.GEN_UNIFORM, "INT32", 1, 0.01, lo=0,
                                                            hi=1 << log_domain),
                                              workloads.col(workloads.GEN_I32_FULL, "INT32", 2, 0.02)])
    p = workloads.generate(build_rows + 1_000_000, 1702, [workloads.col(workloads.GEN_UNIFORM, "INT32", 3, 0.01, lo=0,
                                                           hi=1 << log_domain),
                                             workloads.col(workloads.GEN_I32_FULL, "INT32", 4, 0.02)])
    plan = cjtest.Plan()
    plan.new_input(b)
    plan.new_input(p)
    plan.new_scan_node(0, [(0, "INT32"), (1, "INT32")])
    plan.new_scan_node(1, [(0, "INT32"), (1, "INT32")])
    plan.root = plan.new_join_node(True, 0, 1, 0, 0, [(2, "INT32"), (1, "INT32"), (3, "INT32")])
    exp = cjtest.OracleRun(plan, threads=16)
    assert exp.num_rows > 500_000
    print(f"hash build of {b.num_rows} rows: {exp.num_rows} result rows", flush=True)
    for flags in (0, abi.CJ_EXEC_NO_FAST_PATH):
        got = engine.execute(plan, flags=flags)
        same_multiset(got, exp)
        got.release()
