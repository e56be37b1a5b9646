"""The 113 JOB plans loaded from their EXPLAIN trees (calamarijoins_amd.job)
run on the engine with the brute-force evaluator's results, on every root
strategy (small synthetic tables of the IMDB schema)."""
import pytest

import bruteforce
import cjtest
import test_job as tj
import test_oracle as to
from calamarijoins_amd import abi

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", tj.NAMES)
def test_gpu_job_plans_vs_bruteforce(engine, name):
    parsed = tj.job.ParsedSQL(tj.JOB[name]["sql"])
    tabs = tj.tiny_tables(parsed, 11, rows=16)
    order = []

    def table_for(ent, filt):
        order.append(tabs[ent])
        return cjtest.table_from_rows(*tabs[ent])

    plan, _ = tj.job.load_plan(tj.JOB[name]["plan"], parsed, table_for)
    cplan = cjtest.Plan()
    cplan.nodes, cplan.inputs, cplan.root = plan.nodes, plan.inputs, plan.root
    exp = to.norm_rows(bruteforce.evaluate(cplan.nodes, order, cplan.root))
    for flags in (0, abi.CJ_EXEC_NO_ALIGNED, abi.CJ_EXEC_NO_DIRECT, abi.CJ_EXEC_NO_FAST_PATH):
        got = engine.execute(cplan, flags=flags)
        assert got.num_rows == len(exp), (name, flags)
        assert to.norm_rows(cjtest.result_rows(got)) == exp, (name, flags)
        got.release()
