"""GPU tests of cgj_execute_many: independent plans run concurrently on worker
streams give the rows of the oracle and of one execute_device per plan."""
import random
import sys

import pytest

import gpujoin as gj
import oracle
from gpujoin.pages import INT32, INT64
from gpujoin import workloads as wl
from helpers import check_table, random_plan

pytestmark = pytest.mark.gpu


def assert_join_result(pages, tables, outputs, label):
    """The downloaded result pages of a build/probe join of two device tables
    (key in column 0 of both) hold the rows of the oracle's CPU join over the
    same input pages."""
    got = oracle.table_fingerprint(pages.host_table())
    pages.release()
    b = tables[0].download_pages([0, 1])
    p = tables[1].download_pages([0, 1])
    exp = oracle.join_fingerprint(b.host_table(), 0, p.host_table(), 0, outputs)
    b.release()
    p.release()
    assert got == exp, (label, got, exp)
    return got[0]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _plans(n, seed):
    rng = random.Random(seed)
    out = []
    for i in range(n):
        ntables = rng.choice([2, 2, 3, 4])
        max_rows = rng.choice([50, 400, 3000, 20000])
        domain = max(rng.choice([30, 1000]), max_rows // 4 if ntables >= 3 else 4)
        out.append(random_plan(rng, ntables=ntables, max_rows=max_rows, key_type=rng.choice([INT32, INT64]),
                               domain=domain))
    return out


@pytest.mark.parametrize("workers", [1, 4, 7])
def test_execute_many_matches_oracle_and_single_runs(ctx, workers):
    plans = _plans(30, 77 + workers)
    tables = [[ctx.upload(t) for t in p.inputs] for p in plans]
    say(f"execute_many: {len(plans)} plans on {workers} workers")
    res = ctx.execute_many(plans, tables, workers=workers)
    assert len(res) == len(plans)
    for i, (p, r) in enumerate(zip(plans, res)):
        types = [t for _, t in p.nodes[p.root].output_attrs]
        got = check_table(r.download(), types)
        assert got == check_table(ctx.execute_device(p, tables[i]).download(), types), i
        assert got == check_table(oracle.execute(p), types), i


def test_execute_many_c4_recipe_vs_oracle(ctx):
    """Sixteen c4-shaped joins (generated in HBM) at 1/100 scale in one
    execute_many call: every job's result rows against the oracle, as the
    oracle's fingerprint of the downloaded result pages and of a CPU hash join
    over the same input pages."""
    bw = wl.c4(scale=0.01, njobs=16)
    plans = [w.make_plan() for w in bw.jobs]
    tables = [[ctx.generate(w.seed, t.rows, t.columns) for t in w.tables] for w in bw.jobs]
    say("execute_many: 16 generated c4 joins vs the oracle")
    res = ctx.execute_many(plans, tables, workers=4)
    for j, (ts, r) in enumerate(zip(tables, res)):
        assert_join_result(r.download_pages(), ts, [(0, 1), (1, 1)], j)
        r.free()


def test_execute_many_reports_errors(ctx):
    """A plan whose scan type does not match its input fails the call with the
    engine's message; the context stays usable."""
    plans = _plans(6, 5)
    tables = [[ctx.upload(t) for t in p.inputs] for p in plans]
    bad = plans[2]
    # declare a wrong type for a column of a non-empty scanned table
    scan = next(n for n in bad.nodes if n.kind == gj.NODE_SCAN and bad.inputs[n.base_table_id].num_rows > 0)
    col, ty = scan.output_attrs[0]
    scan.output_attrs[0] = (col, INT64 if ty != INT64 else INT32)
    with pytest.raises(RuntimeError):
        ctx.execute_many(plans, tables, workers=3)
    ok = ctx.execute_many(plans[:2], tables[:2], workers=2)
    assert len(ok) == 2


@pytest.mark.parametrize("workers", [4, 12])
@pytest.mark.parametrize("mode", [gj.EST_INDEPENDENCE, gj.EST_LEARNED])
def test_execute_many_planned_here_matches_single_runs(ctx, mode, workers):
    """Under a model other than the max rule, each worker plans its queries
    with the context's estimator (shared statistics catalog, inference on the
    worker's stream) and executes them: the rows equal those of one
    execute_device per plan under the max rule."""
    from helpers import check_table as ct
    ps = wl.c5(scale=0.002, nplans=24)
    pool = [ctx.generate(ps.seed, t.rows, t.columns) for t in ps.pool]
    plans = [pp.plan for pp in ps.plans]
    tables = [[pool[t] for t in pp.tables] for pp in ps.plans]
    expect = []
    for p, ts in zip(plans, tables):
        types = [t for _, t in p.nodes[p.root].output_attrs]
        expect.append(ct(ctx.execute_device(p, ts).download(), types))
    mlp = ctx.mlp(ctx.num_features(), seed=4) if mode == gj.EST_LEARNED else None
    ctx.set_estimator(mode, mlp, max_sigma=100.0)
    try:
        say(f"execute_many planned by the workers, estimator {mode}, {workers} workers")
        res = ctx.execute_many(plans, tables, workers=workers)
    finally:
        ctx.set_estimator(gj.EST_SAMPLE)
    for p, r, e in zip(plans, res, expect):
        types = [t for _, t in p.nodes[p.root].output_attrs]
        assert ct(r.download(), types) == e
