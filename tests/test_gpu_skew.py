"""GPU parity for skewed hash joins: build sides whose keys carry thousands,
hundreds, a handful and a few duplicates, probed from base tables and from
intermediates, against the oracle (the same multiset of output rows).  The
probes read the first two candidates of a key per lane, walk up to 8 more
per lane and hand longer candidate ranges to the whole wave; these cases
cover each of those with many long ranges in one wave."""
import random

import numpy as np
import pytest

import cjtest
from calamarijoins_amd import abi
from calamarijoins_amd.host import Plan

pytestmark = pytest.mark.gpu

PATHS = [0, abi.CJ_EXEC_NO_ALIGNED, abi.CJ_EXEC_NO_DIRECT, abi.CJ_EXEC_NO_FAST_PATH]
PATH_IDS = ["aligned", "fused", "fused-hash", "general"]

# key -> duplicates on the build side
HEAVY = {7: 2000, 9: 300, 11: 9, 13: 3, 15: 2, -4: 64, 2 ** 31 - 1: 40}


def skewed_build(rng, extra):
    rows = []
    for k, n in HEAVY.items():
        rows += [[k, rng.randrange(-1000, 1000)] for _ in range(n)]
    for _ in range(extra):
        rows.append([None if rng.random() < 0.05 else rng.randrange(100, 5000), rng.randrange(-1000, 1000)])
    rng.shuffle(rows)
    return rows


def skewed_probe(rng, n):
    keys = list(HEAVY) + [None]
    rows = []
    for i in range(n):
        r = rng.random()
        if r < 0.15:
            k = rng.choice(keys)  # runs of hot keys land in the same waves
        elif r < 0.2:
            k = 7
        else:
            k = rng.randrange(0, 6000)
        rows.append([k, i])
    return rows


def same_multiset(got, exp):
    assert got.num_rows == exp.num_rows, (got.num_rows, exp.num_rows)
    assert np.array_equal(cjtest.result_row_hashes(got), cjtest.result_row_hashes(exp))


@pytest.mark.parametrize("flags", PATHS, ids=PATH_IDS)
@pytest.mark.parametrize("seed", range(3))
def test_gpu_skewed_build_base_probe(engine, seed, flags):
    rng = random.Random(900 + seed)
    b = cjtest.table_from_rows(["INT32", "INT32"], skewed_build(rng, 3000))
    p = cjtest.table_from_rows(["INT32", "INT32"], skewed_probe(rng, 12000 + 4000 * seed))
    plan = Plan()
    plan.new_input(b)
    plan.new_input(p)
    plan.new_scan_node(0, [(0, "INT32"), (1, "INT32")])
    plan.new_scan_node(1, [(0, "INT32"), (1, "INT32")])
    plan.root = plan.new_join_node(True, 0, 1, 0, 0, [(0, "INT32"), (1, "INT32"), (3, "INT32")])
    same_multiset(engine.execute(plan, flags=flags), cjtest.OracleRun(plan, threads=8))


@pytest.mark.parametrize("flags", PATHS, ids=PATH_IDS)
def test_gpu_skewed_build_intermediate_probe(engine, flags):
    """(P x D) x S: the intermediate (P x D) is larger than the skewed S, so
    S is the hash build side and the intermediate's keys probe it."""
    rng = random.Random(77)
    p = cjtest.table_from_rows(["INT32", "INT32"], skewed_probe(rng, 15000))
    d = cjtest.table_from_rows(["INT32", "INT32"], [[i, rng.randrange(-9, 9)] for i in range(6000)])
    s = cjtest.table_from_rows(["INT32", "INT32"], skewed_build(rng, 2000))
    plan = Plan()
    for t in (p, d, s):
        plan.new_input(t)
    plan.new_scan_node(0, [(0, "INT32"), (1, "INT32")])
    plan.new_scan_node(1, [(0, "INT32"), (1, "INT32")])
    plan.new_scan_node(2, [(0, "INT32"), (1, "INT32")])
    # P.key = D.key: keys 0..5999 of P survive (the hot keys 7, 9, 11, 13, 15
    # among them); output P.key, P.row, D.pay
    j1 = plan.new_join_node(False, 0, 1, 0, 0, [(0, "INT32"), (1, "INT32"), (3, "INT32")])
    plan.root = plan.new_join_node(False, j1, 2, 0, 0, [(0, "INT32"), (1, "INT32"), (2, "INT32"), (4, "INT32")])
    same_multiset(engine.execute(plan, flags=flags), cjtest.OracleRun(plan, threads=8))
