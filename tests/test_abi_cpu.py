"""CPU checks of the product library: it loads, exports every function the
public headers declare, its host page writer matches the oracle's canonical
encoder byte for byte, the synthetic generator is deterministic, and the C++
API header compiles with the reference's layout."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import cjtest
from calamarijoins_amd import abi, workloads

REPO = cjtest.REPO
HEADERS = ["include/cj_api.h", "include/cj_workload.h", "include/cj_membench.h"]


def declared_functions():
    names = set()
    for h in HEADERS:
        src = open(os.path.join(REPO, h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(cj_[a-z_0-9]+)\s*\(", src, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_function():
    lib = abi.load_library()
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(abi.EXPORTS) <= set(names)


def test_context_creation_fails_cleanly_without_gpu():
    import glob
    if glob.glob("/dev/kfd"):
        pytest.skip("a GPU is present")
    lib = abi.load_library()
    ctx = C.c_void_p()
    rc = lib.cj_ctx_create(None, C.byref(ctx))
    assert rc != 0 and not ctx
    assert lib.cj_last_error()


@pytest.mark.parametrize("seed", [0, 1])
def test_product_page_writer_matches_oracle_encoder(seed):
    rng = np.random.default_rng(seed)
    n = 20000
    for t, dt in ((abi.CJ_INT32, np.int32), (abi.CJ_INT64, np.int64), (abi.CJ_FP64, np.float64)):
        vals = rng.integers(-2**31, 2**31, n).astype(dt)
        valid = (rng.random(n) > rng.random()).astype(np.uint8)
        a = workloads.encode_fixed(t, vals, valid)
        b = cjtest.encode_fixed(t, vals, valid)
        assert a.shape == b.shape and np.array_equal(a, b)


def test_generator_is_deterministic_and_shaped():
    b1, p1 = workloads.c2_tables(build_rows=20_000, probe_rows=100_000, threads=2)
    b2, p2 = workloads.c2_tables(build_rows=20_000, probe_rows=100_000, threads=5)
    for x, y in zip(b1.columns + p1.columns, b2.columns + p2.columns):
        assert np.array_equal(x.pages, y.pages)
    keys, ok = cjtest.decode_column_arrays(abi.CJ_INT32, b1.columns[0].page_ptrs(), b1.columns[0].num_pages)
    assert ok.all() and np.array_equal(np.sort(keys), np.arange(20_000))
    pk, pok = cjtest.decode_column_arrays(abi.CJ_INT32, p1.columns[0].page_ptrs(), p1.columns[0].num_pages)
    miss = np.mean(pk >= 20_000)
    assert pok.all() and 0.28 < miss < 0.32
    # Zipf(1): the most frequent key takes about 1 / H(20000) ~ 9 % of the hits
    hits = pk[pk < 20_000]
    top = np.bincount(hits).max() / len(hits)
    assert 0.07 < top < 0.12
    pay, pv = cjtest.decode_column_arrays(abi.CJ_INT64, p1.columns[2].page_ptrs(), p1.columns[2].num_pages)
    assert 0.015 < 1 - pv.mean() < 0.025


CPP_LAYOUT = r"""
#include "cj/plan.h"
#include <type_traits>
static_assert(sizeof(Page) == 8192 && alignof(Page) == 8, "page layout");
static_assert(std::is_same_v<decltype(Column::pages), std::vector<Page *>>, "column pages");
static_assert(sizeof(ScanNode) == sizeof(size_t), "scan node");
static_assert(sizeof(JoinNode) == 5 * sizeof(size_t), "join node");
static_assert(sizeof(MappedMemory) == 3 * sizeof(size_t), "mapped memory");
static_assert(sizeof(Contest::TimingStats) == 7 * sizeof(int64_t), "timing stats");
int main() {
    Plan plan;
    ColumnarTable t;
    t.columns.emplace_back(DataType::INT32);
    ColumnInserter<int32_t> ins(t.columns.back());
    for (int i = 0; i < 5000; ++i) ins.insert(i);
    ins.finalize();
    t.num_rows = 5000;
    plan.new_input(std::move(t));
    plan.new_scan_node(0, {{0, DataType::INT32}});
    plan.new_scan_node(0, {{0, DataType::INT32}});
    plan.root = plan.new_join_node(true, 0, 1, 0, 0, {{0, DataType::INT32}, {1, DataType::INT32}});
    return plan.inputs[0].columns[0].pages.size() == 3 ? 0 : 1;
}
"""


def test_cpp_api_header_compiles_with_reference_layout():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "t.cpp")
        open(src, "w").write(CPP_LAYOUT)
        exe = os.path.join(d, "t")
        subprocess.run(["g++", "-std=c++17", "-I", os.path.join(REPO, "include"), src, "-o", exe], check=True)
        assert subprocess.run([exe]).returncode == 0


def test_plan_footprint_grows_with_tables():
    """The serving loop's admission estimate (host only)."""
    from calamarijoins_amd import workloads
    from calamarijoins_amd.host import Engine
    small, big = workloads.c2_tables(build_rows=1000, probe_rows=10_000), workloads.c2_tables(build_rows=10_000,
                                                                                               probe_rows=100_000)
    fs = Engine.footprint(workloads.c2_plan(*small))
    fb = Engine.footprint(workloads.c2_plan(*big))
    assert 0 < fs < fb


def test_serve_calls_fail_cleanly_without_context():
    lib = abi.load_library()
    t = C.c_uint64()
    assert lib.cj_serve_stop(None) == abi.CJ_ERR_INVALID
    assert lib.cj_submit(None, None, 0, C.byref(t)) == abi.CJ_ERR_INVALID


def test_batch_call_fails_cleanly_without_context():
    lib = abi.load_library()
    st = abi.BatchStats()
    assert lib.cj_execute_batch(None, None, 0, 0, None, C.byref(st)) == abi.CJ_ERR_INVALID
    assert "null" in lib.cj_last_error().decode()


def test_c4_generator_follows_the_plan_seeds():
    """Plan k of the batch workload uses seed 2000 + k: sizes inside the
    log-uniform ranges (times the scale), keys inside [0, build rows)."""
    from calamarijoins_amd import workloads
    import cjtest
    for k in (0, 1, 7):
        b, p = workloads.c4_plan_tables(k, scale=0.01)
        b2, _ = workloads.c4_plan_tables(k, scale=0.01)
        assert b.num_rows == b2.num_rows
        assert 8 <= b.num_rows <= 20_000 and 100 <= p.num_rows <= 200_000
        kb, vb = cjtest.decode_column_arrays(abi.CJ_INT32, b.columns[0].page_ptrs(), b.columns[0].num_pages)
        kp, vp = cjtest.decode_column_arrays(abi.CJ_INT32, p.columns[0].page_ptrs(), p.columns[0].num_pages)
        assert len(kb) == b.num_rows and len(kp) == p.num_rows
        assert kb[vb].min() >= 0 and kb[vb].max() < b.num_rows
        assert kp[vp].min() >= 0 and kp[vp].max() < b.num_rows
