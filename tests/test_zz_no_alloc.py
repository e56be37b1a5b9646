"""cj_execute_noalloc: the entry point that allocates nothing (temporaries in
the GPU's fixed scratch array, result pages in caller buffers) gives the same
row multisets as cj_execute with device output, on two calls in a row (the
second with a smaller input, over the scratch the first one left), on the
page-aligned, block-writer and general strategies; and it rejects what does
not fit with a clear error."""
import ctypes as C

import numpy as np
import pytest

import cjtest
from calamarijoins_amd import abi, workloads

pytestmark = pytest.mark.gpu

PAGE = abi.PAGE_SIZE


class Hip:
    """hipMalloc / hipMemcpy for the caller's output buffers."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]

    def malloc(self, nbytes):
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), nbytes) == 0
        assert self.lib.hipMemset(p, 0xA5, nbytes) == 0  # stale bytes: nothing may rely on zeroed buffers
        return p

    def to_host(self, p, nbytes):
        out = np.empty(max(nbytes, 1), dtype=np.uint8)
        assert self.lib.hipMemcpy(out.ctypes.data, p, nbytes, 2) == 0
        return out[:nbytes]

    def free(self, p):
        self.lib.hipFree(p)


class PagesResult:
    """Host copies of the pages a no-allocation call wrote, shaped like an
    engine result (num_rows, types, page_lists)."""

    def __init__(self, rows, types, arrays):
        self.num_rows = rows
        self.types = types
        self.arrays = arrays
        self.page_lists = [[a.ctypes.data + i * PAGE for i in range(len(a) // PAGE)] for a in arrays]


def run_noalloc(eng, hip, plan, caps, flags=0):
    """cj_execute_noalloc with len(caps) output buffers of caps[i] pages."""
    lib = abi.load_library()
    bufs = [hip.malloc(max(c, 1) * PAGE) for c in caps]
    try:
        ptrs = (C.c_void_p * len(caps))(*[b.value for b in bufs])
        cap = (C.c_uint32 * len(caps))(*caps)
        npages = (C.c_uint32 * len(caps))()
        rows = C.c_uint64()
        rc = lib.cj_execute_noalloc(eng.ctx, C.byref(plan.desc()), flags, ptrs, cap, len(caps), C.byref(rows),
                                    npages)
        if rc != abi.CJ_OK:
            return rc, lib.cj_last_error().decode(), None
        arrays = [hip.to_host(bufs[i], npages[i] * PAGE).copy() for i in range(len(caps))]
        return rc, "", (rows.value, list(npages), arrays)
    finally:
        for b in bufs:
            hip.free(b)


def expected(eng, plan, flags):
    r = eng.execute(plan, flags=flags | abi.CJ_EXEC_OUTPUT_DEVICE)
    r.to_host()
    out = (r.num_rows, r.types, cjtest.result_row_hashes(r), r.num_pages())
    r.release()
    return out


@pytest.fixture(scope="module")
def setup():
    eng = cjtest.Engine()
    hip = Hip()
    big = workloads.c2_tables(seed=1402, build_rows=400_000, probe_rows=6_400_000)
    small = workloads.c2_tables(seed=1403, build_rows=30_000, probe_rows=200_000)
    ids = [eng.register(t) for t in big + small]
    yield eng, hip, big, small
    for t in ids:
        eng.unregister(t)
    eng.close()


@pytest.mark.parametrize("flags", [0, abi.CJ_EXEC_NO_ALIGNED, abi.CJ_EXEC_NO_DIRECT, abi.CJ_EXEC_NO_FAST_PATH],
                         ids=["aligned", "fused", "fused-hash", "general"])
def test_gpu_noalloc_equals_execute_twice(setup, flags):
    eng, hip, big, small = setup
    for tabs in (big, small):  # the second call is smaller and reuses the scratch
        plan = workloads.c2_plan(*tabs)
        rows, types, want, pages = expected(eng, plan, flags)
        caps = [p + 4 + p // 64 for p in pages]  # the strategies' page counts differ slightly
        rc, msg, got = run_noalloc(eng, hip, plan, caps, flags)
        assert rc == abi.CJ_OK, msg
        n, npages, arrays = got
        assert n == rows
        res = PagesResult(n, types, arrays)
        assert np.array_equal(cjtest.result_row_hashes(res), want)


def test_gpu_noalloc_rejects_what_does_not_fit(setup):
    eng, hip, big, small = setup
    plan = workloads.c2_plan(*small)
    rows, types, want, pages = expected(eng, plan, abi.CJ_EXEC_NO_ALIGNED)  # the block writers, as no-alloc uses
    # an output buffer one page too small
    caps = [p + 2 for p in pages]
    caps[1] = max(pages[1] - 1, 0)
    rc, msg, _ = run_noalloc(eng, hip, plan, caps)
    assert rc == abi.CJ_ERR_OOM and "needs room for" in msg, msg
    # fewer buffers than result columns
    rc, msg, _ = run_noalloc(eng, hip, plan, [p + 2 for p in pages][:3])
    assert rc == abi.CJ_ERR_INVALID and "fewer output buffers" in msg, msg
    # an input that is not resident
    b, p = workloads.c2_tables(seed=1404, build_rows=2_000, probe_rows=10_000)
    rc, msg, _ = run_noalloc(eng, hip, workloads.c2_plan(b, p), [64] * 5)
    assert rc == abi.CJ_ERR_UNSUPPORTED and "resident" in msg, msg
    # temporaries beyond the scratch: a 12M-row hash build (directory and
    # partitioned tuples alone exceed 256 MiB)
    hb = workloads.generate(12_000_000, 1405, [workloads.col(workloads.GEN_UNIFORM, "INT32", 1, 0.0, lo=0, hi=1 << 30),
                                               workloads.col(workloads.GEN_I32_FULL, "INT32", 2)])
    hp = workloads.generate(13_000_000, 1406, [workloads.col(workloads.GEN_UNIFORM, "INT32", 3, 0.0, lo=0, hi=1 << 30),
                                               workloads.col(workloads.GEN_I32_FULL, "INT32", 4)])
    ids = [eng.register(hb), eng.register(hp)]
    try:
        big_plan = workloads.c4_plan(hb, hp)
        rc, msg, _ = run_noalloc(eng, hip, big_plan, [4096] * 3, abi.CJ_EXEC_NO_DIRECT)
        assert rc == abi.CJ_ERR_OOM and "scratch exhausted" in msg, msg
    finally:
        for t in ids:
            eng.unregister(t)
    # the scratch still serves a plan that fits after the failed call
    caps = [p + 2 for p in pages]
    rc, msg, got = run_noalloc(eng, hip, plan, caps)
    assert rc == abi.CJ_OK, msg
    assert np.array_equal(cjtest.result_row_hashes(PagesResult(got[0], types, got[2])), want)
