"""Writes tests/golden/job_queries.json: the JOB query set as the reference
holds it (plans.json: PostgreSQL EXPLAIN plans of the 113 queries; job/*.sql;
schema.json), trimmed to what the plan adapter reads: per plan node its type,
children, alias / relation, row estimate, parallel flags and whether it has a
filter. Run from the repository root where the reference is checked out:

    python tests/golden/make_job_fixture.py /root/reference
"""
import json
import os
import sys

KEEP = ("Node Type", "Alias", "Relation Name", "Plan Rows", "Parallel Aware", "Workers Planned")


def trim(node):
    out = {k: node[k] for k in KEEP if k in node}
    if "Filter" in node:
        out["Filter"] = True
    if "Plans" in node:
        out["Plans"] = [trim(c) for c in node["Plans"]]
    return out


def main(ref):
    plans = json.load(open(os.path.join(ref, "plans.json")))
    schema = json.load(open(os.path.join(ref, "schema.json")))
    sql_dir = os.path.join(ref, plans["sql_directory"])
    queries = []
    for name, plan in zip(plans["names"], plans["plans"]):
        with open(os.path.join(sql_dir, name + ".sql")) as f:
            sql = " ".join(f.read().split())
        queries.append({"name": name, "sql": sql, "plan": trim(plan["Plan"])})
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "job_queries.json")
    with open(out, "w") as f:
        json.dump({"source": "reference plans.json, job/*.sql, schema.json", "schema": schema,
                   "queries": queries}, f, separators=(",", ":"))
    print(f"{len(queries)} queries -> {out}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
