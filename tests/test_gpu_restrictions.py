"""Scan restrictions, singleton elimination and semi-joins on the GPU against
the CPU oracle (results must not depend on them), each also run with the
feature switched off.

* After a build of at most 32 tuples, the incremental loop restricts the
  unjoined producers of the key by its keys (reference
  engine/storage/RestrictionLogic.cpp:397-447): one key (Eq), two (Eq2), a
  contiguous run (range), up to 16 (tiny table), else none.
* A duplicate-free table carrying only its key becomes an exact restriction
  of its key's producers instead of a join input
  (engine/query/QueryPlan.cpp:602-625).
* One-row relations are removed before planning; their values restrict the
  other members of their classes, and outputs only they produce become
  constant columns, VARCHAR and NULL included (QueryPlan.cpp:680-717)."""
import random

import pytest

import gpujoin as gj
import oracle
from gpujoin.pages import INT32, INT64, VARCHAR
from helpers import check_table

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def star(dim_keys, fact_rows=4000, fk_domain=200, seed=0, key_type=INT32, dim_out=True):
    """dim(k, label) <- fact(fk, g) -> other(g, v): output dim label (or not),
    fact row id and other's value."""
    rng = random.Random(seed)
    dim = gj.make_table([(key_type, list(dim_keys), None),
                         (VARCHAR, [f"d{k}" for k in dim_keys], None)], len(dim_keys))
    fk = [rng.randrange(fk_domain) for _ in range(fact_rows)]
    fvalid = [rng.random() > 0.05 for _ in range(fact_rows)]
    g = [rng.randrange(300) for _ in range(fact_rows)]
    fact = gj.make_table([(key_type, fk, fvalid), (INT32, g, None), (INT64, list(range(fact_rows)), None)], fact_rows)
    other = gj.make_table([(INT32, list(range(300)), None), (INT64, [rng.randrange(1 << 40) for _ in range(300)], None)],
                          300)
    plan = gj.Plan()
    plan.new_input(dim)
    plan.new_input(fact)
    plan.new_input(other)
    d = plan.new_scan_node(0, [(0, key_type), (1, VARCHAR)])
    f = plan.new_scan_node(1, [(0, key_type), (1, INT32), (2, INT64)])
    o = plan.new_scan_node(2, [(0, INT32), (1, INT64)])
    j1 = plan.new_join_node(True, d, f, 0, 0, [(1, VARCHAR), (3, INT32), (4, INT64)])
    outs = [(0, VARCHAR), (2, INT64), (4, INT64)] if dim_out else [(2, INT64), (4, INT64)]
    plan.root = plan.new_join_node(False, j1, o, 1, 0, outs)
    types = [t for _, t in plan.nodes[plan.root].output_attrs]
    return plan, types


def run_both(ctx, plan, types, option):
    exp = check_table(oracle.execute(plan), types)
    got = check_table(ctx.execute(plan), types)
    st = ctx.stats()
    ctx.set_option(option, 0)
    try:
        off = check_table(ctx.execute(plan), types)
    finally:
        ctx.set_option(option, 1)
    assert got == exp
    assert off == exp
    return st


KIND_CASES = {
    "eq": [17],
    "eq2": [17, 140],
    "range": list(range(40, 52)),
    "range_from_0": [0, 1, 2, 3],
    "tiny_mod": [1, 3, 5, 9, 100],
    "tiny_hashed": [0, 32, 64, 96, 128],
    "none_17_sparse": [7 * i for i in range(17)],
    "none_40": list(range(0, 200, 5))[:40],
}


@pytest.mark.parametrize("name", sorted(KIND_CASES))
@pytest.mark.parametrize("key_type", [INT32, INT64])
def test_small_table_restrictions(ctx, name, key_type):
    keys = KIND_CASES[name]
    say(f"restriction case {name}, key type {key_type}")
    plan, types = star(keys, seed=len(keys), key_type=key_type)
    st = run_both(ctx, plan, types, "restrictions")
    if len(keys) > 1 and name.startswith("none"):
        return
    # a one-key dimension is a singleton (removed before planning); the others
    # restrict the fact table after their build
    assert st["restrictions"] >= 1, st


def test_semijoin_table(ctx):
    """The dimension (1,000 unique keys, only its key needed) becomes an exact
    restriction of the fact table."""
    say("semi-join dimension")
    plan, types = star(list(range(0, 2000, 2)), fact_rows=6000, fk_domain=2500, dim_out=False)
    st = run_both(ctx, plan, types, "restrictions")
    assert st["semijoins"] >= 1, st


def single_row_plan(dim_key, dim_label, label_valid=True, bridge=False):
    """S(k, label, n) has one row; F(fk, x) 3,000 rows; O(x, v) 50 rows. S
    joins F on k (S.k = F.fk). Normally F joins O (F.x = O.x), and S's label
    and n are outputs only S produces (constant columns once S is removed).
    With `bridge`, S joins O instead (S.n = O.x) and F does not: removing S
    would disconnect F from O, so S stays a relation."""
    rng = random.Random(5)
    s = gj.make_table([(INT32, [0 if dim_key is None else dim_key], [dim_key is not None]),
                       (VARCHAR, [dim_label], [label_valid]), (INT32, [7], None)], 1)
    fk = [rng.randrange(20) for _ in range(3000)]
    x = [rng.randrange(50) for _ in range(3000)]
    f = gj.make_table([(INT32, fk, None), (INT32, x, None)], 3000)
    o = gj.make_table([(INT32, list(range(50)), None), (INT64, [rng.randrange(1 << 30) for _ in range(50)], None)], 50)
    plan = gj.Plan()
    plan.new_input(s)
    plan.new_input(f)
    plan.new_input(o)
    sn = plan.new_scan_node(0, [(0, INT32), (1, VARCHAR), (2, INT32)])
    fn = plan.new_scan_node(1, [(0, INT32), (1, INT32)])
    on = plan.new_scan_node(2, [(0, INT32), (1, INT64)])
    # j1 outputs: label, n, F.x
    j1 = plan.new_join_node(True, sn, fn, 0, 0, [(1, VARCHAR), (2, INT32), (4, INT32)])
    # root inputs: label, n, F.x, O.x, O.v
    on_attr = 1 if bridge else 2
    plan.root = plan.new_join_node(False, j1, on, on_attr, 0, [(0, VARCHAR), (1, INT32), (4, INT64)])
    return plan, [VARCHAR, INT32, INT64]


@pytest.mark.parametrize("case", ["plain", "long_label", "null_label", "null_key", "absent_key", "bridge"])
def test_singleton_elimination(ctx, case):
    say(f"singleton case {case}")
    label = "x" * 9000 if case == "long_label" else "dimension-label"
    key = None if case == "null_key" else (999 if case == "absent_key" else 4)
    plan, types = single_row_plan(key, label, label_valid=case != "null_label", bridge=case == "bridge")
    st = run_both(ctx, plan, types, "singletons")
    assert st["singletons"] == (0 if case == "bridge" else 1), st
    if case in ("plain", "long_label", "null_label"):
        assert st["output_rows"] > 0, st


def test_all_singletons(ctx):
    say("every relation one row")
    a = gj.make_table([(INT32, [3], None), (INT64, [11], None)], 1)
    b = gj.make_table([(INT32, [3], None), (VARCHAR, ["b"], None)], 1)
    plan = gj.Plan()
    plan.new_input(a)
    plan.new_input(b)
    an = plan.new_scan_node(0, [(0, INT32), (1, INT64)])
    bn = plan.new_scan_node(1, [(0, INT32), (1, VARCHAR)])
    plan.root = plan.new_join_node(True, an, bn, 0, 0, [(1, INT64), (3, VARCHAR), (0, INT32)])
    types = [INT64, VARCHAR, INT32]
    st = run_both(ctx, plan, types, "singletons")
    assert st["singletons"] == 1 and st["output_rows"] == 1, st


@pytest.mark.parametrize("seed", range(16))
def test_small_random_plans_with_restrictions(ctx, seed):
    """Random plans over tiny tables (many one-row and few-row relations)."""
    from helpers import random_plan
    rng = random.Random(4000 + seed)
    plan = random_plan(rng, ntables=rng.randint(2, 6), max_rows=rng.choice([1, 3, 8, 30]),
                       key_type=rng.choice([INT32, INT64]), domain=rng.choice([3, 10, 40]))
    types = [t for _, t in plan.nodes[plan.root].output_attrs]
    exp = check_table(oracle.execute(plan), types)
    assert check_table(ctx.execute(plan), types) == exp
    for opt in ("singletons", "restrictions"):
        ctx.set_option(opt, 0)
    try:
        assert check_table(ctx.execute(plan), types) == exp
    finally:
        for opt in ("singletons", "restrictions"):
            ctx.set_option(opt, 1)
