"""JOB (Join Order Benchmark) workload adapter: a query's SQL text and its
PostgreSQL plan (EXPLAIN FORMAT JSON, hash joins only) become a contest Plan,
as the reference's harness builds them (tests/read_sql.cpp:732-860 parses the
SQL, :862-1142 walks the plan), over synthetic inputs shaped like the
filtered IMDB scans.

* parse_sql: FROM entities (a table and its occurrence), the columns inside
  the select list's functions (the query's output), and the equivalence
  classes of equated columns (column = column predicates of the top-level
  conjunction; every other predicate is a filter of one entity).
* load_join_pipeline: Aggregate and Gather nodes pass through; a Hash Join
  takes its Hash child as the build side (build_left when the Hash is on the
  left) and joins on an edge of the join graph between the two sides' entity
  sets; every scan is a new input whose output attributes are the columns
  required above it (the query's outputs and the join columns).
* synthesize: the harness loads each scan's table from IMDB CSV files with
  the scan's filter applied (read_sql.cpp:1086-1089). The CSV files are not
  here, so each scan's input is generated instead: PostgreSQL's row estimate
  for the scan (per worker for a parallel-aware scan, so scaled by the
  parallel divisor) times `scale`; an `id` column is a random subset of its
  table's key domain, a column joined to some table's `id` in any query of
  the set is a foreign key into that table's domain, other columns are
  uniform integers or ASCII strings. The domain of a table is its largest
  unfiltered scan estimate over the set. Results are compared with the CPU
  oracle on the same generated inputs, never with IMDB answers.
"""
from __future__ import annotations

import json
import re
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

from . import (GEN_ASCII, GEN_FK_ZIPF, GEN_PERM_KEY, GEN_UNIFORM, INT32, VARCHAR, GenColumn, Plan)

Entity = Tuple[str, int]  # (table, occurrence in the FROM list)
Attr = Tuple[Entity, str]  # (entity, column)

SQL_TYPES = {"integer": INT32, "text": VARCHAR, "character varying": VARCHAR}
SCAN_NODES = ("Seq Scan", "Index Only Scan")
PASS_NODES = ("Aggregate", "Gather")
JOIN_NODES = ("Nested Loop", "Hash Join", "Merge Join")


# ---------------------------------------------------------------------------
# SQL
# ---------------------------------------------------------------------------
_TOKEN = re.compile(r"""\s*(?:
    (?P<str>'(?:[^']|'')*')
  | (?P<num>\d+(?:\.\d+)?)
  | (?P<id>[A-Za-z_][A-Za-z0-9_]*(?:\.[A-Za-z_][A-Za-z0-9_]*)?)
  | (?P<op><>|!=|<=|>=|\|\||::|[=<>(),;*+\-/%.])
)""", re.X)


def tokenize(sql: str) -> List[Tuple[str, str]]:
    """(kind, text) tokens: str, num, id (possibly `alias.column`), op."""
    out, pos = [], 0
    sql = sql.rstrip()
    while pos < len(sql):
        m = _TOKEN.match(sql, pos)
        if not m or m.end() == pos:
            if sql[pos:].strip() == "":
                break
            raise ValueError(f"cannot tokenize SQL at {sql[pos:pos + 30]!r}")
        kind = m.lastgroup
        out.append((kind, m.group(kind)))
        pos = m.end()
    return out


def _split(tokens, sep_test):
    """Splits a token list at depth-0 tokens for which sep_test holds."""
    parts, cur, depth = [], [], 0
    for t in tokens:
        if t == ("op", "("):
            depth += 1
        elif t == ("op", ")"):
            depth -= 1
        if depth == 0 and sep_test(t):
            parts.append(cur)
            cur = []
        else:
            cur.append(t)
    parts.append(cur)
    return parts


def _kw(t, word):
    return t[0] == "id" and t[1].lower() == word


def _strip_parens(tokens):
    while len(tokens) >= 2 and tokens[0] == ("op", "(") and tokens[-1] == ("op", ")"):
        depth = 0
        for i, t in enumerate(tokens):
            depth += t == ("op", "(")
            depth -= t == ("op", ")")
            if depth == 0 and i < len(tokens) - 1:
                return tokens  # the first paren closes before the end
        tokens = tokens[1:-1]
    return tokens


@dataclass
class ParsedSQL:
    """What the harness takes from a JOB query (read_sql.cpp:680-860)."""
    entities: List[Entity] = field(default_factory=list)         # FROM order
    alias: Dict[str, Entity] = field(default_factory=dict)
    outputs: List[Attr] = field(default_factory=list)            # select-list columns
    classes: List[List[Attr]] = field(default_factory=list)      # equated columns, 2+ each
    # join_graph[a][b] = (column of a, column of b) for entities with an equated column pair
    join_graph: Dict[Entity, Dict[Entity, Tuple[str, str]]] = field(default_factory=dict)
    filtered: set = field(default_factory=set)                   # entities with a filter


def parse_sql(sql: str, schema: Dict[str, Dict[str, dict]]) -> ParsedSQL:
    toks = tokenize(sql)
    if toks and toks[-1] == ("op", ";"):
        toks = toks[:-1]
    idx = {w: next((i for i, t in enumerate(toks) if _kw(t, w)), None) for w in ("select", "from", "where")}
    if idx["select"] is None or idx["from"] is None:
        raise ValueError("not a SELECT ... FROM query")
    ps = ParsedSQL()
    counts: Dict[str, int] = {}
    where_at = idx["where"] if idx["where"] is not None else len(toks)
    for item in _split(toks[idx["from"] + 1:where_at], lambda t: t == ("op", ",")):
        words = [t[1] for t in item if not _kw(t, "as")]
        if not words or words[0] not in schema:
            raise ValueError(f"unknown table in FROM: {words}")
        table = words[0]
        ent = (table, counts.get(table, 0))
        counts[table] = ent[1] + 1
        ps.entities.append(ent)
        if len(words) > 1:
            ps.alias[words[1]] = ent

    def resolve(ref: str) -> Attr:
        if "." in ref:
            a, c = ref.split(".", 1)
            if a in ps.alias:
                ent = ps.alias[a]
            elif counts.get(a) == 1:
                ent = (a, 0)
            else:
                raise ValueError(f"unknown alias {a}")
        else:
            c = ref
            owners = [e for e in ps.entities if c in schema[e[0]]]
            if len(owners) != 1:
                raise ValueError(f"column {c} is ambiguous or unknown")
            ent = owners[0]
        if c not in schema[ent[0]]:
            raise ValueError(f"no column {c} in {ent[0]}")
        return ent, c

    for item in _split(toks[idx["select"] + 1:idx["from"]], lambda t: t == ("op", ",")):
        # MIN(x.y) AS name | x.y [AS name]: every column reference of the item
        body = item
        for i, t in enumerate(item):
            if _kw(t, "as"):
                body = item[:i]
                break
        # identifiers that are not function names
        refs = [t[1] for i, t in enumerate(body)
                if t[0] == "id" and not (i + 1 < len(body) and body[i + 1] == ("op", "("))]
        if not refs:
            raise ValueError("select item without a column")
        ps.outputs.extend(resolve(r) for r in refs)

    parent: Dict[Attr, Attr] = {}

    def find(a):
        parent.setdefault(a, a)
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    if idx["where"] is not None:
        # the top-level conjunction; the AND of `x BETWEEN a AND b` stays in its term
        in_between = [False]

        def conj(t):
            if _kw(t, "between"):
                in_between[0] = True
            elif _kw(t, "and"):
                if in_between[0]:
                    in_between[0] = False
                    return False
                return True
            return False

        for term in _split(toks[idx["where"] + 1:], conj):
            term = _strip_parens(term)
            if (len(term) == 3 and term[1] == ("op", "=") and term[0][0] == "id" and term[2][0] == "id"
                    and not _kw(term[0], "null") and not _kw(term[2], "null")):
                a, b = resolve(term[0][1]), resolve(term[2][1])
                parent[find(a)] = find(b)
            else:
                ents = {resolve(t[1])[0] for t in term if t[0] == "id" and "." in t[1]}
                if len(ents) != 1:
                    raise ValueError("a filter that spans tables cannot be pushed down")
                ps.filtered |= ents
    groups: Dict[Attr, List[Attr]] = {}
    order = {e: i for i, e in enumerate(ps.entities)}
    for a in parent:
        groups.setdefault(find(a), []).append(a)
    for g in groups.values():
        if len(g) < 2:
            continue
        g.sort(key=lambda a: (order[a[0]], a[1]))
        ps.classes.append(g)
        for i in range(len(g)):
            for j in range(i + 1, len(g)):
                (ea, ca), (eb, cb) = g[i], g[j]
                if eb in ps.join_graph.get(ea, {}):
                    raise ValueError("two conditions between the same pair of tables")
                ps.join_graph.setdefault(ea, {})[eb] = (ca, cb)
                ps.join_graph.setdefault(eb, {})[ea] = (cb, ca)
    ps.classes.sort(key=lambda g: (order[g[0][0]], g[0][1]))
    return ps


# ---------------------------------------------------------------------------
# plan
# ---------------------------------------------------------------------------
@dataclass
class JobScan:
    entity: Entity
    columns: List[str]   # the input's columns, in input order (what the scan needs)
    rows_estimate: float  # PostgreSQL's estimate of the filtered scan (whole scan, not per worker)
    filtered: bool


@dataclass
class JobPlan:
    name: str
    plan: Plan           # inputs not attached; input i belongs to scans[i]
    scans: List[JobScan]
    types: List[int]     # output column types


def _entities(node, ps: ParsedSQL) -> List[Entity]:
    kind = node["Node Type"]
    if kind in PASS_NODES:
        return _entities(node["Plans"][0], ps)
    if kind in JOIN_NODES:
        left, right = _sides(node)
        return _entities(left, ps) + _entities(right, ps)
    if kind in SCAN_NODES:
        return [_scan_entity(node, ps)]
    raise ValueError(f"unsupported plan node {kind}")


def _sides(node):
    if node["Node Type"] != "Hash Join":
        raise ValueError("only hash joins are supported")
    lk, rk = node["Plans"][0]["Node Type"], node["Plans"][1]["Node Type"]
    if lk == "Hash" and rk != "Hash":
        return node["Plans"][0]["Plans"][0], node["Plans"][1]
    if lk != "Hash" and rk == "Hash":
        return node["Plans"][0], node["Plans"][1]["Plans"][0]
    raise ValueError("a hash join needs exactly one Hash child")


def _scan_entity(node, ps: ParsedSQL) -> Entity:
    if "Alias" in node:
        if node["Alias"] not in ps.alias:
            if (node["Alias"], 0) in ps.entities:
                return node["Alias"], 0
            raise ValueError(f"unknown alias {node['Alias']}")
        return ps.alias[node["Alias"]]
    name = node.get("Relation Name")
    if name is None:
        raise ValueError("scan node without alias or relation")
    ents = [e for e in ps.entities if e[0] == name]
    if len(ents) != 1:
        raise ValueError(f"table {name} is used {len(ents)} times")
    return ents[0]


def load_join_pipeline(name: str, root: dict, ps: ParsedSQL, schema: Dict[str, Dict[str, dict]]) -> JobPlan:
    """The contest Plan of a PostgreSQL plan tree (read_sql.cpp:862-1142). Of
    several equated column pairs between a join's two sides, the first in
    FROM order (left side's entity, then right side's) is the join
    condition."""
    plan = Plan()
    scans: List[JobScan] = []
    order = {e: i for i, e in enumerate(ps.entities)}

    def rec(node, required: List[Attr], workers: int):
        kind = node["Node Type"]
        if kind in PASS_NODES:
            if kind == "Gather":
                workers = node.get("Workers Planned", 0)
            return rec(node["Plans"][0], required, workers)
        if kind in JOIN_NODES:
            lnode, rnode = _sides(node)
            build_left = node["Plans"][0]["Node Type"] == "Hash"
            lents, rents = _entities(lnode, ps), _entities(rnode, ps)
            cond = None
            for le in sorted(lents, key=order.get):
                for re_ in sorted(rents, key=order.get):
                    if re_ in ps.join_graph.get(le, {}):
                        cond = (le, re_, *ps.join_graph[le][re_])
                        break
                if cond:
                    break
            if cond is None:
                raise ValueError(f"{name}: no join condition between {lents} and {rents}")
            le, re_, lc, rc = cond
            lreq = [a for a in required if a[0] in lents]
            rreq = [a for a in required if a[0] in rents]
            if len(lreq) + len(rreq) != len(required):
                raise ValueError(f"{name}: a required column is on neither side")
            if (le, lc) not in lreq:
                lreq.append((le, lc))
            if (re_, rc) not in rreq:
                rreq.append((re_, rc))
            lid, lcols = rec(lnode, lreq, workers)
            rid, rcols = rec(rnode, rreq, workers)
            both = lcols + rcols
            out = [(both.index(a), _type(schema, a)) for a in required]
            return plan.new_join_node(build_left, lid, rid, lcols.index((le, lc)), rcols.index((re_, rc)),
                                      out), list(required)
        if kind in SCAN_NODES:
            ent = _scan_entity(node, ps)
            cols: List[str] = []
            for e, c in required:
                if e != ent:
                    raise ValueError(f"{name}: required column {e}.{c} is not in scan {ent}")
                if c not in cols:
                    cols.append(c)
            rows = float(node.get("Plan Rows", 1))
            if node.get("Parallel Aware"):
                # per-worker estimate: times the parallel divisor (workers plus
                # the leader's share, 1 - 0.3 x workers when positive)
                rows *= workers + max(0.0, 1.0 - 0.3 * workers)
            input_id = len(scans)
            scans.append(JobScan(ent, cols, rows, "Filter" in node or ent in ps.filtered))
            nid = plan.new_scan_node(input_id, [(cols.index(c), _type(schema, (e, c))) for e, c in required])
            return nid, list(required)
        raise ValueError(f"{name}: unsupported plan node {kind}")

    plan.root, _ = rec(root["Plan"] if "Plan" in root else root, list(ps.outputs), 0)
    return JobPlan(name, plan, scans, [_type(schema, a) for a in ps.outputs])


def _type(schema, attr: Attr) -> int:
    (table, _), col = attr
    t = schema[table][col]["type"]
    if t not in SQL_TYPES:
        raise ValueError(f"unsupported column type {t}")
    return SQL_TYPES[t]


# ---------------------------------------------------------------------------
# synthetic inputs
# ---------------------------------------------------------------------------
@dataclass
class JobSet:
    """A set of JOB queries, their plans and what synthesizing their inputs needs."""
    schema: Dict[str, Dict[str, dict]]
    names: List[str]
    parsed: List[ParsedSQL]
    plans: List[JobPlan]
    domains: Dict[str, int]                  # table -> key domain (rows of the unfiltered table)
    references: Dict[Tuple[str, str], str]   # (table, column) -> table whose id it references


def build_set(names: Sequence[str], sqls: Sequence[str], pg_plans: Sequence[dict],
              schema: Dict[str, Dict[str, dict]]) -> JobSet:
    parsed = [parse_sql(s, schema) for s in sqls]
    plans = [load_join_pipeline(n, p, ps, schema) for n, p, ps in zip(names, pg_plans, parsed)]
    domains: Dict[str, float] = {}
    unfiltered: Dict[str, float] = {}
    for jp in plans:
        for s in jp.scans:
            t = s.entity[0]
            domains[t] = max(domains.get(t, 0.0), s.rows_estimate)
            if not s.filtered:
                unfiltered[t] = max(unfiltered.get(t, 0.0), s.rows_estimate)
    dom = {t: max(1, int(round(max(v, unfiltered.get(t, 0.0))))) for t, v in domains.items()}
    refs: Dict[Tuple[str, str], str] = {}
    for ps in parsed:
        for g in ps.classes:
            targets = sorted({e[0] for e, c in g if c == "id" and "id" in schema[e[0]]})
            for (t, _), c in g:
                if c != "id" and targets:
                    refs.setdefault((t, c), next((x for x in targets if x != t), targets[0]))
    return JobSet(schema, list(names), parsed, plans, dom, refs)


def load_set(path_or_dict) -> JobSet:
    """A JOB set from the fixture format written by tests/golden/make_job_fixture.py:
    {"schema": ..., "queries": [{"name", "sql", "plan"}, ...]}."""
    d = path_or_dict if isinstance(path_or_dict, dict) else json.load(open(path_or_dict))
    qs = d["queries"]
    return build_set([q["name"] for q in qs], [q["sql"] for q in qs], [q["plan"] for q in qs], d["schema"])


def scan_rows(js: JobSet, scan: JobScan, scale: float) -> int:
    return max(1, min(int(round(scan.rows_estimate * scale)), _domain(js, scan.entity[0], scale)))


def _domain(js: JobSet, table: str, scale: float) -> int:
    return max(1, int(round(js.domains[table] * scale)))


def scan_columns(js: JobSet, qi: int, si: int, scale: float, null_frac: float = 0.02) -> List[GenColumn]:
    """Generators of input si of query qi (one stream per query, scan and column)."""
    scan = js.plans[qi].scans[si]
    table = scan.entity[0]
    out = []
    for ci, c in enumerate(scan.columns):
        spec = js.schema[table][c]
        stream = 100_000 * (qi + 1) + 100 * si + ci
        nulls = 0.0 if spec.get("not_null") else null_frac
        if c == "id" and spec.get("primary_key"):
            g = GenColumn(type=INT32, kind=GEN_PERM_KEY, stream=stream, domain=_domain(js, table, scale))
        elif (table, c) in js.references:
            g = GenColumn(type=INT32, kind=GEN_FK_ZIPF, stream=stream, null_stream=stream + 50, null_frac=nulls,
                          domain=_domain(js, js.references[(table, c)], scale), zipf_s=0.5)
        elif SQL_TYPES[spec["type"]] == VARCHAR:
            g = GenColumn(type=VARCHAR, kind=GEN_ASCII, stream=stream, null_stream=stream + 50, null_frac=nulls,
                          lo=3, hi=40)
        else:
            g = GenColumn(type=INT32, kind=GEN_UNIFORM, stream=stream, null_stream=stream + 50, null_frac=nulls,
                          lo=0, hi=1 << 31)
        out.append(g)
    return out


def host_inputs(js: JobSet, qi: int, scale: float, seed: int = 0x10B) -> list:
    """The query's inputs generated on the host (pages as the contest writer packs them)."""
    from . import ColumnarTable, host_generate
    out = []
    for si, scan in enumerate(js.plans[qi].scans):
        n = scan_rows(js, scan, scale)
        out.append(ColumnarTable(n, [host_generate(seed, 0, n, g) for g in scan_columns(js, qi, si, scale)]))
    return out


def device_inputs(ctx, js: JobSet, qi: int, scale: float, seed: int = 0x10B) -> list:
    """The same inputs in device memory: generated there when every column is
    fixed-width, else generated on the host and uploaded."""
    out = []
    for si, scan in enumerate(js.plans[qi].scans):
        n = scan_rows(js, scan, scale)
        gens = scan_columns(js, qi, si, scale)
        if any(g.type == VARCHAR for g in gens):
            from . import ColumnarTable, host_generate
            out.append(ctx.upload(ColumnarTable(n, [host_generate(seed, 0, n, g) for g in gens])))
        else:
            out.append(ctx.generate(seed, n, gens))
    return out


def with_inputs(jp: JobPlan, inputs) -> Plan:
    """The query's Plan with its inputs attached (input i to scan i)."""
    jp.plan.inputs = list(inputs)
    return jp.plan
