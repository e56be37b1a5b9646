"""Writes tests/golden/job_plans.json from the reference's JOB inputs:
plans.json (PostgreSQL EXPLAIN (FORMAT JSON) of the 113 JOB queries, read by
tests/read_sql.cpp) and job/<name>.sql.  Each plan tree keeps only the keys
the plan loader reads ("Node Type", "Alias", "Relation Name", "Hash Cond",
"Plans").  Run once in a checkout that has the reference next to it:

    python tests/golden/make_job_fixture.py /root/reference
"""
import json
import os
import sys

KEEP = ("Node Type", "Alias", "Relation Name", "Hash Cond")


def reduce(node):
    out = {k: node[k] for k in KEEP if k in node}
    if "Plans" in node:
        out["Plans"] = [reduce(c) for c in node["Plans"]]
    return out


def main(ref):
    d = json.load(open(os.path.join(ref, "plans.json")))
    sql_dir = os.path.join(ref, d["sql_directory"])
    out = {}
    for name, plan in zip(d["names"], d["plans"]):
        with open(os.path.join(sql_dir, name + ".sql")) as f:
            sql = f.read().strip()
        out[name] = {"sql": sql, "plan": reduce(plan["Plan"])}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "job_plans.json")
    with open(path, "w") as f:
        json.dump(out, f, separators=(",", ":"), sort_keys=True)
    print(path, os.path.getsize(path), "bytes,", len(out), "queries")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
