"""The CPU oracle pinned against the reference's golden vectors and brute force.

Oracle = test infrastructure (oracle/); these tests check it before it is
used to check the GPU path.
"""
import random

import numpy as np
import pytest

import gpujoin as gj
import oracle
from gpujoin.pages import FP64, INT32, INT64, VARCHAR
from helpers import TYPES, brute_force, canon_rows, case_plan, check_table, load_cases

CASES = load_cases()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_oracle_golden(case):
    """Reference tests/unit_tests.cpp ground truths (transcribed fixtures)."""
    res = oracle.execute(case_plan(case))
    rows = check_table(res, [TYPES[t] for t in case["expect_types"]])
    assert res.num_rows == len(case["expect_rows"])
    assert rows == canon_rows(case["expect_rows"])


@pytest.mark.parametrize("seed", range(40))
def test_oracle_vs_brute_force(seed):
    rng = random.Random(seed)
    plan = __import__("helpers").random_plan(rng, ntables=rng.randint(1, 4), max_rows=rng.choice([3, 12, 40]),
                                              key_type=rng.choice([INT32, INT64]), long_strings=seed % 5 == 0)
    res = oracle.execute(plan)
    types = [t for _, t in plan.nodes[plan.root].output_attrs]
    assert check_table(res, types) == brute_force(plan)


def test_oracle_page_capacities():
    """Reference TableTarget bit budget: 1984 INT32 / 1007 INT64 rows per page
    without NULLs (closed forms 8*(8192-4)/33 and 8*(8192-8)/65)."""
    for t, per_page in ((INT32, 1984), (INT64, 1007), (FP64, 1007)):
        n = per_page * 3 + 5
        vals = list(range(n)) if t != FP64 else [float(i) for i in range(n)]
        plan = gj.Plan()
        plan.new_input(gj.ColumnarTable(n, [gj.encode_column(t, vals)]))
        plan.root = plan.new_scan_node(0, [(0, t)])
        res = oracle.execute(plan)
        counts = [gj.pages.page_header(p)[0] for p in res.columns[0].pages]
        assert counts == [per_page] * 3 + [5]


def test_oracle_long_string_pages():
    """A 42,174-character string spans ceil(42174/8188) = 6 long-string pages
    (first 0xffff, then 0xfffe); its neighbours stay on normal pages."""
    s = "q" * 42174
    plan = gj.Plan()
    plan.new_input(gj.ColumnarTable(3, [gj.encode_column(VARCHAR, ["a", s, "b"])]))
    plan.root = plan.new_scan_node(0, [(0, VARCHAR)])
    res = oracle.execute(plan)
    tags = [gj.pages.page_header(p)[0] for p in res.columns[0].pages]
    assert tags.count(0xFFFF) == 1 and tags.count(0xFFFE) == 5
    vals, valid = gj.decode_column(res.columns[0].pages, VARCHAR)
    assert vals == [b"a", s.encode(), b"b"]


def test_oracle_null_sentinel():
    """64-bit slot semantics of the reference: INT64 -1 reads back as NULL."""
    plan = gj.Plan()
    plan.new_input(gj.ColumnarTable(3, [gj.encode_column(INT64, [5, -1, 7])]))
    plan.root = plan.new_scan_node(0, [(0, INT64)])
    rows = check_table(oracle.execute(plan), [INT64])
    assert rows == canon_rows([(None,), (5,), (7,)])


def test_oracle_decode_and_join_count():
    rng = np.random.default_rng(3)
    vals = rng.integers(-50, 50, 5000)
    valid = rng.random(5000) > 0.2
    col = gj.encode_column(INT64, vals, valid)
    slots = oracle.decode_fixed(col, 5000)
    # independent reader: the numpy page decoder
    v, m = gj.decode_column(col.pages, INT64)
    exp = np.where(m, v.view(np.uint64), np.uint64(~np.uint64(0)))
    assert np.array_equal(slots, exp)
    # Against the raw values, the only differences are the contest writer's
    # overlap: its fit test (plan.h:204-213) reserves 4 bytes for any value, so
    # on 8-byte pages with NULLs the bitmap can overwrite the high 4 bytes of
    # the page's last value. Every reader sees the overwritten bytes.
    raw = np.where(valid, vals.astype(np.int64).view(np.uint64), np.uint64(~np.uint64(0)))
    diff = np.nonzero(slots != raw)[0]
    ends = np.cumsum([gj.pages.page_header(p)[0] for p in col.pages]) - 1
    assert set(diff.tolist()) <= set(ends.tolist())
    assert all((slots[i] & np.uint64(0xFFFFFFFF)) == (raw[i] & np.uint64(0xFFFFFFFF)) for i in diff)
    b = rng.integers(0, 30, 400).astype(np.uint64)
    p = rng.integers(0, 40, 900).astype(np.uint64)
    brute = int(sum((b == x).sum() for x in p))
    assert oracle.join_count(b, p) == brute
