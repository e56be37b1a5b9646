"""JOB front end (calamarijoins_amd.job): the 113 JOB queries of the
reference harness (tests/golden/job_plans.json: their SQL and PostgreSQL
EXPLAIN trees from the reference's plans.json) parse, load into Plans whose
structure the EXPLAIN tree and the SQL fix, filters print in the reference's
format, and the loaded plans run on the CPU oracle with the results of the
brute-force evaluator.  Host only."""
import json
import os
import random
import re

import pytest

import bruteforce
import cjtest
import test_oracle as to
from calamarijoins_amd import abi, host, job

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "job_plans.json")
JOB = json.load(open(GOLDEN))
NAMES = sorted(JOB)


class Stub:
    """Stands in for a table where the test only looks at plan structure."""

    def __init__(self, ent, filt):
        self.ent, self.filt = ent, filt


def walk(node):
    yield node
    for c in node.get("Plans", []):
        yield from walk(c)


def test_all_113_queries_parse_and_load():
    assert len(NAMES) == 113
    for name in NAMES:
        parsed = job.ParsedSQL(JOB[name]["sql"])
        plan, cols = job.load_plan(JOB[name]["plan"], parsed, lambda e, f: Stub(e, f))
        assert [(e, c) for e, c, _ in cols] == parsed.output_attrs, name


@pytest.mark.parametrize("name", NAMES)
def test_plan_structure_follows_explain_and_sql(name):
    parsed = job.ParsedSQL(JOB[name]["sql"])
    explain = JOB[name]["plan"]
    plan, cols = job.load_plan(explain, parsed, lambda e, f: Stub(e, f))
    hash_joins = [n for n in walk(explain) if n["Node Type"] == "Hash Join"]
    joins = [n for n in plan.nodes if n["kind"] == abi.CJ_NODE_JOIN]
    scans = [n for n in plan.nodes if n["kind"] == abi.CJ_NODE_SCAN]
    assert len(joins) == len(hash_joins)
    assert len(scans) == len(plan.inputs) == len(parsed.entities)
    assert sorted(t.ent for t in plan.inputs) == sorted(parsed.entities)
    # every input carries its entity's filter
    for t in plan.inputs:
        assert t.filt is parsed.filters.get(t.ent)

    # the columns each node outputs, to check every join's condition
    def out_cols(i):
        n = plan.nodes[i]
        if n["kind"] == abi.CJ_NODE_SCAN:
            ent = plan.inputs[n["table"]].ent
            return [(ent, job.SCHEMA[ent[0]][c][0]) for c, _ in n["attrs"]]
        both = out_cols(n["left"]) + out_cols(n["right"])
        return [both[k] for k, _ in n["attrs"]]

    def ents(i):
        n = plan.nodes[i]
        if n["kind"] == abi.CJ_NODE_SCAN:
            return {plan.inputs[n["table"]].ent}
        return ents(n["left"]) | ents(n["right"])

    for n in joins:
        le, lc = out_cols(n["left"])[n["left_attr"]]
        re_, rc = out_cols(n["right"])[n["right_attr"]]
        assert le in ents(n["left"]) and re_ in ents(n["right"])
        assert parsed.join_class(le, lc) == parsed.join_class(re_, rc)
        # exactly one equivalence class connects the two sides, so the
        # choice among connecting column pairs cannot change the result
        classes = {parsed.join_class(a, c) for a in ents(n["left"]) for c, _ in job.SCHEMA[a[0]]
                   for b in ents(n["right"]) for d, _ in job.SCHEMA[b[0]]
                   if parsed.join_class(a, c) == parsed.join_class(b, d)}
        assert len(classes) == 1, (name, classes)
    # PostgreSQL's hash condition of each join names a pair of the same class
    for hj in hash_joins:
        m = re.match(r"\(+(\w+)\.(\w+) = (\w+)\.(\w+)\)", hj["Hash Cond"])
        a, ac, b, bc = m.groups()
        ea, eb = parsed.alias_map[a], parsed.alias_map[b]
        assert parsed.join_class(ea, ac) == parsed.join_class(eb, bc)
    # root columns: the select list, typed by the schema
    assert [t for _, _, t in cols] == [job.SCHEMA[e[0]][job.column_index(e[0], c)][1]
                                       for e, c in parsed.output_attrs]


def test_filter_text_follows_the_reference_printer():
    p = job.ParsedSQL(JOB["1a"]["sql"])
    mc, ct, it = p.alias_map["mc"], p.alias_map["ct"], p.alias_map["it"]
    assert p.filters[ct].pretty() == "1 = 'production companies'"
    assert p.filters[it].pretty() == "1 = 'top 250 rank'"
    assert p.filters[mc].pretty() == ("[AND]\n"
                                      "  4 NOT LIKE '%(as Metro-Goldwyn-Mayer Pictures)%'\n"
                                      "  [OR]\n"
                                      "    4 LIKE '%(co-production)%'\n"
                                      "    4 LIKE '%(presents)%'")
    q = job.ParsedSQL(JOB["33c"]["sql"])
    assert q.filters[q.alias_map["kt1"]].pretty() == "[OR]\n  1 = 'tv series'\n  1 = 'episode'"
    assert q.filters[q.alias_map["t2"]].pretty() == "[AND]\n  4 >= 2000\n  4 <= 2010"
    assert q.filters[q.alias_map["lt"]].pretty() == ("[OR]\n  [OR]\n    1 = 'sequel'\n    1 = 'follows'\n"
                                                     "  1 = 'followed by'")
    # IS NOT NULL is NOT (IS NULL); the comparison prints a trailing blank
    s = job.ParsedSQL("SELECT MIN(t.title) FROM title AS t WHERE t.imdb_index IS NOT NULL AND t.id = t.id")
    assert s.filters[("title", 0)].pretty() == "[NOT]\n  2 IS NULL "


def test_cache_names_of_filtered_scans():
    p = job.ParsedSQL(JOB["1a"]["sql"])
    ct = p.alias_map["ct"]
    name = host.tbl_cache_name("company_type", p.filters[ct].pretty())
    assert re.fullmatch(r"company_type_\d+\.tbl", name)
    assert host.tbl_cache_name("title") != name


def test_filter_evaluation_on_records():
    p = job.ParsedSQL(JOB["1a"]["sql"])
    f = p.filters[p.alias_map["mc"]]
    row = [1, 2, 3, 4, None]
    assert not f.eval(row)  # NULL note
    row[4] = "(co-production)"
    assert f.eval(row)
    row[4] = "x (presents) (as Metro-Goldwyn-Mayer Pictures)"
    assert not f.eval(row)
    assert job.like_match("a.b", "a_b") and not job.like_match("a.bc", "a_b")
    assert job.like_match("", "%") and job.like_match("100%", "100%")


@pytest.mark.parametrize("sql,err", [
    ("SELECT MIN(t.title) FROM title AS t, movie_companies AS mc WHERE t.id = mc.movie_id OR t.kind_id = 1",
     "join condition"),
    ("SELECT MIN(t.title) FROM title AS t, movie_companies AS mc WHERE t.id < mc.movie_id", "Non-EuqalJoins"),
    ("SELECT MIN(t.title) FROM title AS t, movie_companies AS mc WHERE (t.kind_id = 1 OR mc.note = 'x') "
     "AND t.id = mc.movie_id", "pushed down"),
    ("SELECT MIN(t.title) FROM nosuchtable AS t", "schema"),
    ("SELECT MIN(t.nosuchcolumn) FROM title AS t", "column"),
])
def test_unsupported_sql_is_rejected(sql, err):
    with pytest.raises(job.JobError, match=err):
        job.ParsedSQL(sql)


def test_non_hash_joins_are_rejected():
    p = job.ParsedSQL(JOB["1a"]["sql"])
    explain = json.loads(json.dumps(JOB["1a"]["plan"]))
    for n in walk(explain):
        if n["Node Type"] == "Hash Join":
            n["Node Type"] = "Nested Loop"
            break
    with pytest.raises(job.JobError, match="Not Hash Join"):
        job.load_plan(explain, p, lambda e, f: Stub(e, f))


# ------------------------------------------------------------ execution
def tiny_tables(parsed, seed, rows=10):
    """Per entity a small table of its schema: ids 0..rows-1, other INT32
    columns in [0, rows) with a few NULLs, short strings."""
    rng = random.Random(seed)
    out = {}
    for ent in parsed.entities:
        types = [abi.TYPE_IDS[t] for _, t in job.SCHEMA[ent[0]]]
        data = []
        for i in range(rows):
            r = []
            for (cname, t) in job.SCHEMA[ent[0]]:
                if t == job.INT32:
                    r.append(i if cname == "id" else (None if rng.random() < 0.08 else rng.randrange(rows)))
                else:
                    r.append(None if rng.random() < 0.1 else "".join(rng.choice("abc") for _ in range(rng.randrange(4))))
            data.append(r)
        out[ent] = (types, data)
    return out


def load_tiny(name, seed):
    parsed = job.ParsedSQL(JOB[name]["sql"])
    tabs = tiny_tables(parsed, seed)
    order = []

    def table_for(ent, filt):  # filters are left out so that the joins see rows
        order.append(tabs[ent])
        return cjtest.table_from_rows(*tabs[ent])

    plan, _ = job.load_plan(JOB[name]["plan"], parsed, table_for)
    cplan = cjtest.Plan()
    cplan.nodes, cplan.inputs, cplan.root = plan.nodes, plan.inputs, plan.root
    return cplan, order


@pytest.mark.parametrize("name", NAMES[::4])
def test_loaded_plans_run_on_the_oracle(name):
    plan, tables = load_tiny(name, 7)
    exp = bruteforce.evaluate(plan.nodes, tables, plan.root)
    got = cjtest.OracleRun(plan, threads=2)
    assert got.num_rows == len(exp)
    assert to.norm_rows(cjtest.result_rows(got)) == to.norm_rows(exp)
