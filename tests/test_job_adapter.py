"""The JOB adapter (CPU): the committed fixture (reference plans.json, job/*.sql
and schema.json, trimmed by tests/golden/make_job_fixture.py) turns into 113
contest Plans like the reference harness builds them
(tests/read_sql.cpp:732-1142), and the CPU oracle executes each of them on
small synthetic inputs."""
import json
import os

import pytest

import gpujoin as gj
import oracle
from gpujoin import job
from gpujoin.pages import INT32, VARCHAR

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "job_queries.json")
REF = "/root/reference"


@pytest.fixture(scope="module")
def js():
    return job.load_set(FIXTURE)


def test_all_queries_load(js):
    assert len(js.plans) == 113
    assert js.names[:3] == ["1a", "1b", "1c"]
    for ps, jp in zip(js.parsed, js.plans):
        # one input per FROM entity, each scanned once
        assert sorted(s.entity for s in jp.scans) == sorted(ps.entities)
        assert jp.types == [job.SQL_TYPES[js.schema[t][c]["type"]] for (t, _), c in ps.outputs]
        assert len(jp.plan.nodes) == 2 * len(jp.scans) - 1


def test_query_1a_structure(js):
    """1a (job/1a.sql): five tables; MIN over mc.note, t.title,
    t.production_year; three join classes; PostgreSQL hashes mc's subtree
    against a seq scan of title (the Hash child is on the right)."""
    ps = js.parsed[0]
    assert ps.entities == [("company_type", 0), ("info_type", 0), ("movie_companies", 0),
                           ("movie_info_idx", 0), ("title", 0)]
    assert ps.outputs == [(("movie_companies", 0), "note"), (("title", 0), "title"),
                          (("title", 0), "production_year")]
    classes = sorted(sorted((e[0], c) for e, c in g) for g in ps.classes)
    assert classes == [
        [("company_type", "id"), ("movie_companies", "company_type_id")],
        [("info_type", "id"), ("movie_info_idx", "info_type_id")],
        [("movie_companies", "movie_id"), ("movie_info_idx", "movie_id"), ("title", "id")],
    ]
    assert ps.filtered == {("company_type", 0), ("info_type", 0), ("movie_companies", 0)}
    jp = js.plans[0]
    root = jp.plan.nodes[jp.plan.root]
    assert root.kind == gj.NODE_JOIN and not root.build_left
    left = jp.plan.nodes[root.left]
    assert left.kind == gj.NODE_SCAN and jp.scans[left.base_table_id].entity == ("title", 0)
    assert [t for _, t in root.output_attrs] == [VARCHAR, VARCHAR, INT32]
    # PostgreSQL estimates per worker for the parallel title scan; the
    # divisor for two workers is 2.4
    assert jp.scans[left.base_table_id].rows_estimate == pytest.approx(1053435 * 2.4)


def test_sql_parser_details():
    schema = {"a": {"id": {"type": "integer"}, "x": {"type": "integer"}, "s": {"type": "text"}},
              "b": {"id": {"type": "integer"}, "a_id": {"type": "integer"}}}
    ps = job.parse_sql("SELECT MIN(a1.s) AS m, MIN(a2.x) FROM a AS a1, a AS a2, b "
                       "WHERE a1.x BETWEEN 3 AND 7 AND (a1.s LIKE '%AND%' OR a1.s = 'x') "
                       "AND a1.id = b.a_id AND b.a_id = a2.id AND a2.s IS NOT NULL;", schema)
    assert ps.entities == [("a", 0), ("a", 1), ("b", 0)]
    assert ps.outputs == [(("a", 0), "s"), (("a", 1), "x")]
    assert ps.filtered == {("a", 0), ("a", 1)}
    assert [sorted(g) for g in ps.classes] == [[(("a", 0), "id"), (("a", 1), "id"), (("b", 0), "a_id")]]
    assert ps.join_graph[("a", 0)][("a", 1)] == ("id", "id")
    with pytest.raises(ValueError):
        job.parse_sql("SELECT MIN(a.x) FROM a, b WHERE a.x < b.id", schema)


@pytest.mark.parametrize("part", range(4))
def test_oracle_runs_every_plan(js, part):
    """Every plan is well formed: the oracle executes it on tiny inputs, and
    the output has the query's column types."""
    for qi in range(part, len(js.plans), 4):
        jp = js.plans[qi]
        inputs = job.host_inputs(js, qi, 2e-4)
        res = oracle.execute(job.with_inputs(jp, inputs))
        assert [c.type for c in res.columns] == jp.types, jp.name


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "job")), reason="reference checkout not present")
def test_fixture_matches_reference_files(js):
    """The trimmed fixture gives the same plans as the reference's own files."""
    d = json.load(open(os.path.join(REF, "plans.json")))
    schema = json.load(open(os.path.join(REF, "schema.json")))
    sqls = [open(os.path.join(REF, "job", n + ".sql")).read() for n in d["names"]]
    full = job.build_set(d["names"], sqls, d["plans"], schema)
    assert full.names == js.names
    assert full.domains == js.domains and full.references == js.references
    for a, b in zip(full.plans, js.plans):
        assert [(s.entity, s.columns, s.rows_estimate, s.filtered) for s in a.scans] == \
               [(s.entity, s.columns, s.rows_estimate, s.filtered) for s in b.scans]
        assert [vars(n) for n in a.plan.nodes] == [vars(n) for n in b.plan.nodes]
