"""Shared test utilities: golden-case loading, random plans, brute-force joins
and row-multiset comparison."""
from __future__ import annotations

import json
import os
import random
import struct
from typing import List, Optional

import numpy as np

import gpujoin as gj
from gpujoin.pages import FP64, INT32, INT64, VARCHAR

TYPES = {"INT32": INT32, "INT64": INT64, "FP64": FP64, "VARCHAR": VARCHAR}
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_cases():
    with open(os.path.join(GOLDEN, "unit_cases.json")) as f:
        return json.load(f)["cases"]


def case_plan(case) -> gj.Plan:
    plan = gj.Plan()
    for tab in case["tables"]:
        types = [TYPES[t] for t in tab["types"]]
        rows = tab["rows"]
        cols = []
        for ci, t in enumerate(types):
            vals = [r[ci] for r in rows]
            valid = [v is not None for v in vals]
            if t == VARCHAR:
                vals = [v if v is not None else "" for v in vals]
            else:
                vals = [v if v is not None else 0 for v in vals]
            cols.append(gj.encode_column(t, vals, valid))
        plan.new_input(gj.ColumnarTable(len(rows), cols))
    for nd in case["nodes"]:
        out = [(i, TYPES[t]) for i, t in nd["out"]]
        if "scan" in nd:
            plan.new_scan_node(nd["scan"], out)
        else:
            bl, l, r, la, ra = nd["join"]
            plan.new_join_node(bl, l, r, la, ra, out)
    plan.root = case["root"]
    return plan


def _key(v):
    if v is None:
        return (0, b"")
    if isinstance(v, bytes):
        return (1, v)
    if isinstance(v, float):
        return (1, struct.pack(">d", v))
    if isinstance(v, tuple):  # NaN, by bit pattern
        return (1, b"nan" + v[1])
    return (1, v)


def _canon(v):
    """Strings as bytes; a NaN (which never equals itself) as its bit pattern,
    so rows holding the same NaN bits compare equal."""
    if isinstance(v, str):
        return v.encode()
    if isinstance(v, float) and v != v:
        return ("nan", struct.pack(">d", v))
    return v


def canon_rows(rows) -> List[tuple]:
    """Rows with strings as bytes and NaNs as bit patterns, sorted with NULLs
    first per column."""
    out = []
    for r in rows:
        out.append(tuple(_canon(v) for v in r))
    return sorted(out, key=lambda r: tuple(_key(v) for v in r))


def table_multiset(t: gj.ColumnarTable) -> List[tuple]:
    return canon_rows(gj.table_rows(t))


def check_table(t: gj.ColumnarTable, types):
    """Validates the result pages and returns its sorted rows."""
    assert [c.type for c in t.columns] == list(types)
    for c in t.columns:
        rows = gj.pages.validate_column(c.pages, c.type)
        assert rows == t.num_rows, f"column has {rows} rows, table {t.num_rows}"
    return table_multiset(t)


# ---------------------------------------------------------------------------
# random plans
# ---------------------------------------------------------------------------
def rand_values(rng: random.Random, t: int, n: int, domain: int, null_frac: float):
    valid = [rng.random() >= null_frac for _ in range(n)]
    if t == INT32:
        vals = [rng.randrange(domain) for _ in range(n)]
    elif t == INT64:
        vals = [rng.randrange(domain) * 1_000_000_007 - 5 for _ in range(n)]
    elif t == FP64:
        vals = [rng.choice([0.5, -1.25, 3.0, 1e300, -0.0]) + rng.randrange(domain) for _ in range(n)]
    else:
        vals = ["".join(rng.choice("abcxyz") for _ in range(rng.choice([0, 1, 3, 7, 20, 70]))) for _ in range(n)]
    return vals, valid


def random_plan(rng: random.Random, ntables: int, max_rows: int, key_type=INT32, domain=12,
                null_frac=0.15, long_strings=False):
    """Random tree-shaped join of `ntables` tables. Each table has a key column
    per edge, plus payload columns; the plan joins them left-deep in a random
    order with random build sides and projects a random subset."""
    plan = gj.Plan()
    edges = [(rng.randrange(i), i) for i in range(1, ntables)]  # parent, child
    ncols = [1] * ntables
    colidx = {}
    for ei, (p, c) in enumerate(edges):
        colidx[(ei, p)] = ncols[p]
        ncols[p] += 1
        colidx[(ei, c)] = ncols[c]
        ncols[c] += 1
    tables = []
    for t in range(ntables):
        n = rng.randrange(0 if rng.random() < 0.05 else 1, max_rows + 1)
        cols = []
        types = []
        for c in range(ncols[t]):
            if c == 0:
                ty = rng.choice([INT32, INT64, FP64, VARCHAR])
            else:
                ty = key_type
            vals, valid = rand_values(rng, ty, n, domain, null_frac)
            if long_strings and ty == VARCHAR and n and rng.random() < 0.5:
                vals[rng.randrange(n)] = "L" * rng.choice([8186, 9000, 20000])
            cols.append(gj.encode_column(ty, vals, valid))
            types.append(ty)
        plan.new_input(gj.ColumnarTable(n, cols))
        tables.append(types)
    # scans output every column
    scan_of = []
    for t in range(ntables):
        scan_of.append(plan.new_scan_node(t, [(c, tables[t][c]) for c in range(ncols[t])]))
    # join tables in edge order: left = accumulated, right = new child
    layout = [(0, c) for c in range(ncols[0])]  # (table, column) per output position
    cur = scan_of[0]
    joined = {0}
    pending = list(range(len(edges)))
    while pending:
        for k in list(pending):
            p, c = edges[k]
            if p in joined and c not in joined:
                new_t, old_t = c, p
            elif c in joined and p not in joined:
                new_t, old_t = p, c
            else:
                continue
            pending.remove(k)
            la = layout.index((old_t, colidx[(k, old_t)]))
            ra = colidx[(k, new_t)]
            right_layout = [(new_t, cc) for cc in range(ncols[new_t])]
            full = layout + right_layout
            keep = list(range(len(full)))
            out = [(i, tables[full[i][0]][full[i][1]]) for i in keep]
            cur = plan.new_join_node(rng.random() < 0.5, cur, scan_of[new_t], la, ra, out)
            layout = full
            joined.add(new_t)
            break
    # final projection: a random subset (at least one column), in random order
    root = plan.nodes[cur]
    k = rng.randrange(1, min(len(root.output_attrs), 6) + 1)
    pick = rng.sample(range(len(root.output_attrs)), k)
    # a join's output indices refer to its children's concatenated columns (`layout`);
    # a lone scan's refer to the base table's columns
    root.output_attrs = [(i, tables[layout[i][0]][layout[i][1]]) for i in pick]
    plan.root = cur
    return plan


def brute_force(plan: gj.Plan) -> List[tuple]:
    """Nested-loop evaluation of the plan tree (independent of the oracle)."""

    def decode_table(t):
        cols = [gj.decode_column(c.pages, c.type) for c in t.columns]
        rows = []
        for r in range(t.num_rows):
            row = []
            for (v, m), c in zip(cols, t.columns):
                if not m[r]:
                    row.append(None)
                else:
                    x = v[r]
                    row.append(x if isinstance(x, bytes) else x.item())
            rows.append(row)
        return rows

    tabs = [decode_table(t) for t in plan.inputs]

    def null_sentinel(v, t):
        # 64-bit slot semantics: INT64 -1 / FP64 all-ones bits read back as NULL
        if v is None:
            return None
        if t == INT64 and v == -1:
            return None
        if t == FP64 and struct.pack("<d", v) == b"\xff" * 8:
            return None
        return v

    def rec(i):
        nd = plan.nodes[i]
        if nd.kind == gj.NODE_SCAN:
            rows = tabs[nd.base_table_id]
            return [[null_sentinel(r[c], t) for c, t in nd.output_attrs] for r in rows]
        L = rec(nd.left)
        R = rec(nd.right)
        out = []
        for lr in L:
            a = lr[nd.left_attr]
            if a is None:
                continue
            for rr in R:
                b = rr[nd.right_attr]
                if b is None or a != b:
                    continue
                full = lr + rr
                out.append([full[j] for j, _ in nd.output_attrs])
        return out

    return canon_rows(rec(plan.root))
