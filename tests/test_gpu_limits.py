"""Plans at the engine's shape limits against the oracle, and one past them:
16 scanned columns and 16 output columns in one pipeline, hash-table entries
of 16 words (key + 15 attributes), and a 17-column output that must fail
with a clear error and leave the context usable."""
import random

import pytest

import gpujoin as gj
import oracle
from gpujoin.pages import INT32, INT64
from helpers import check_table

pytestmark = pytest.mark.gpu


def wide_tables(rng, ncols_wide, ncols_dim, rows=3000, dim_rows=400):
    wide = gj.make_table([(INT32, [rng.randrange(dim_rows * 2) for _ in range(rows)], None)] +
                         [(INT64, [rng.randrange(1 << 40) for _ in range(rows)],
                           [rng.random() > 0.1 for _ in range(rows)]) for _ in range(ncols_wide - 1)], rows)
    dim = gj.make_table([(INT32, list(range(dim_rows)), None)] +
                        [(INT32, [rng.randrange(1 << 30) for _ in range(dim_rows)],
                          [rng.random() > 0.1 for _ in range(dim_rows)]) for _ in range(ncols_dim - 1)], dim_rows)
    return wide, dim


def join_plan(wide, dim, ncols_wide, ncols_dim, outs):
    plan = gj.Plan()
    plan.new_input(dim)
    plan.new_input(wide)
    d = plan.new_scan_node(0, [(0, INT32)] + [(c, INT32) for c in range(1, ncols_dim)])
    w = plan.new_scan_node(1, [(0, INT32)] + [(c, INT64) for c in range(1, ncols_wide)])
    types = [INT32] * ncols_dim + [INT32] + [INT64] * (ncols_wide - 1)
    plan.root = plan.new_join_node(True, d, w, 0, 0, [(i, types[i]) for i in outs])
    return plan, [types[i] for i in outs]


def test_sixteen_scanned_and_output_columns(ctx):
    rng = random.Random(1)
    wide, dim = wide_tables(rng, 16, 1)
    # the driver's 16 columns are all scanned and all output (the key once)
    plan, types = join_plan(wide, dim, 16, 1, list(range(1, 17)))
    assert check_table(ctx.execute(plan), types) == check_table(oracle.execute(plan), types)


def test_sixteen_word_hash_table_entries(ctx):
    rng = random.Random(2)
    wide, dim = wide_tables(rng, 1, 16, dim_rows=5000, rows=20000)
    # build side: key + 15 attributes, all carried through the table
    plan, types = join_plan(wide, dim, 1, 16, list(range(0, 16)))
    assert check_table(ctx.execute(plan), types) == check_table(oracle.execute(plan), types)


def test_past_the_output_limit_raises(ctx):
    rng = random.Random(3)
    wide, dim = wide_tables(rng, 16, 2)
    plan, types = join_plan(wide, dim, 16, 2, list(range(0, 18)))  # 17 distinct columns + the key twice
    with pytest.raises(RuntimeError, match="too many|wider than supported"):
        ctx.execute(plan)
    ok, ok_types = join_plan(wide, dim, 16, 2, [1, 3, 5])
    assert check_table(ctx.execute(ok), ok_types) == check_table(oracle.execute(ok), ok_types)
