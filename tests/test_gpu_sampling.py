"""The sampling estimate and the incremental plan-execute loop (the reference's
planner: engine/query/QueryPlan.cpp:66-72, 719-746, 748-782).

* The sampled cardinality of every relation equals a CPU restatement: the
  same sample rows (drawn with replacement from a stream fixed by the
  relation's size and position), decoded by the oracle, rows x (sample rows
  with no NULL join column) / (sample rows), rows / (2 x sample rows) when
  none passes.
* The loop plans once per hash-table build plus once for the output
  pipeline, and its rows equal the oracle's and those of planning once."""
import random

import pytest

import gpujoin as gj
import oracle
from gpujoin.pages import INT32, INT64
from helpers import check_table, random_plan

pytestmark = pytest.mark.gpu

M64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15


def sm_mix(z):
    z &= M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def sample_rows(rows, r):
    size = min(rows, 64)
    seed = sm_mix(rows * GOLDEN + r + 1)
    return sorted(sm_mix(seed + (i + 1) * GOLDEN) % rows for i in range(size))


@pytest.mark.parametrize("seed", range(8))
def test_sample_cardinalities_restate_the_sampler(ctx, seed):
    rng = random.Random(300 + seed)
    plan = random_plan(rng, ntables=rng.choice([2, 3, 4]), max_rows=rng.choice([40, 900, 5000]),
                       key_type=rng.choice([INT32, INT64]), domain=200, null_frac=rng.choice([0.1, 0.5, 0.9]))
    tables = [ctx.upload(t) for t in plan.inputs]
    got = ctx.sample_cardinalities(plan, tables)
    # scans in depth-first order (left first), and the (scan, column) pairs
    # that join conditions name: the members of the join classes
    order, joined = [], set()

    def rec(n):
        nd = plan.nodes[n]
        if nd.kind == gj.NODE_SCAN:
            order.append(nd)
            return [(len(order) - 1, c) for c, _ in nd.output_attrs]
        left, right = rec(nd.left), rec(nd.right)
        joined.add(left[nd.left_attr])
        joined.add(right[nd.right_attr])
        both = left + right
        return [both[i] for i, _ in nd.output_attrs]
    rec(plan.root)
    scans = [(nd.base_table_id, nd.output_attrs) for nd in order]
    assert len(got) == len(scans)
    if any(plan.inputs[tid].num_rows == 0 for tid, _ in scans):
        assert all(g == 0 for g in got)  # an empty input empties the join: nothing is sampled
        return
    for r, (tid, attrs) in enumerate(scans):
        rows = plan.inputs[tid].num_rows
        if rows == 0:
            assert got[r] == 0
            continue
        ids = sample_rows(rows, r)
        cols = [c for (rr, c) in joined if rr == r]
        decoded = {c: oracle.decode_fixed(plan.inputs[tid].columns[c], rows) for c in cols}
        passing = sum(all(int(decoded[c][i]) != M64 for c in cols) for i in ids)
        exp = rows * (passing / len(ids) if passing else 1.0 / (2 * len(ids)))
        assert got[r] == pytest.approx(exp, rel=1e-12), (r, got[r], exp)


@pytest.mark.parametrize("seed", range(10))
def test_incremental_loop_vs_oracle(ctx, seed):
    rng = random.Random(500 + seed)
    ntables = rng.randint(2, 6)
    max_rows = rng.choice([50, 400, 3000])
    plan = random_plan(rng, ntables=ntables, max_rows=max_rows, key_type=rng.choice([INT32, INT64]),
                       domain=max(max_rows, 30))
    tables = [ctx.upload(t) for t in plan.inputs]
    types = [t for _, t in plan.nodes[plan.root].output_attrs]
    exp = check_table(oracle.execute(plan), types)
    got = check_table(ctx.execute_device(plan, tables).download(), types)
    st = ctx.stats()
    assert got == exp
    if st["output_rows"] and all(t.num_rows for t in plan.inputs):
        # one plan per hash table built, one for the output pipeline
        assert st["replans"] == len(plan.inputs), st
    ctx.set_option("incremental", 0)
    try:
        assert check_table(ctx.execute_device(plan, tables).download(), types) == exp
        assert ctx.stats()["replans"] == 0
    finally:
        ctx.set_option("incremental", 1)
