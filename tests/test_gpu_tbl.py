"""cj_tbl_load: a .tbl file streamed into HBM is a resident table; plans
over it give the results of plans over the same table registered from host
memory and of the non-resident path, and its pages in HBM are the file's."""
import numpy as np
import pytest

import cjtest
from calamarijoins_amd import abi, host, workloads

pytestmark = pytest.mark.gpu


def test_gpu_tbl_load_equals_host_tables(engine, tmp_path):
    b, p = workloads.c2_tables(seed=97, build_rows=80_000, probe_rows=1_280_000)
    host.write_tbl(str(tmp_path / "b.tbl"), b)
    host.write_tbl(str(tmp_path / "p.tbl"), p)
    plan = workloads.c2_plan(b, p)
    r = engine.execute(plan)
    exp = (r.num_rows, cjtest.result_row_hashes(r))
    r.release()
    lb = engine.load_tbl(str(tmp_path / "b.tbl"))
    lp = engine.load_tbl(str(tmp_path / "p.tbl"))
    try:
        assert lb.num_rows == b.num_rows and lp.num_rows == p.num_rows
        for x, y in zip(lp.columns, p.columns):
            assert np.array_equal(x.pages, y.pages)
        plan2 = workloads.c2_plan(lb, lp)
        for flags in (0, abi.CJ_EXEC_NO_ALIGNED, abi.CJ_EXEC_NO_FAST_PATH):
            r = engine.execute(plan2, flags)
            assert r.num_rows == exp[0]
            assert np.array_equal(cjtest.result_row_hashes(r), exp[1]), flags
            r.release()
        # the loaded tables are resident: the batch takes the plan
        res, st = engine.execute_batch([plan2])
        assert st.plans_batched == 1 and res[0].num_rows == exp[0]
        assert np.array_equal(cjtest.result_row_hashes(res[0]), exp[1])
        res[0].release()
    finally:
        lb.close()
        lp.close()


def test_gpu_tbl_load_rejects_malformed_file(engine, tmp_path):
    path = tmp_path / "bad.tbl"
    path.write_bytes(b"\0" * 100)
    with pytest.raises(abi.CJError):
        engine.load_tbl(str(path))
    # a header that promises more pages than the file has
    t = cjtest.table_from_rows(["INT32"], [[i] for i in range(5000)])
    host.write_tbl(str(path), t)
    data = path.read_bytes()
    path.write_bytes(data[:-8192])
    with pytest.raises(abi.CJError):
        engine.load_tbl(str(path))
