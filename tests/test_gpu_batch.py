"""GPU parity of batched joins (one persistent kernel launch over a device work
queue, replayed from a hipGraph): every job's rows equal the oracle's and the
one-launch-per-join path's, over mixed shapes and through graph replays."""
import sys

import numpy as np
import pytest

import gpujoin as gj
import oracle
from gpujoin.pages import FP64, INT32, INT64
from helpers import check_table
from test_gpu_many import assert_join_result

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _two_table_plan(build, probe, bkey, pkey, out, build_left=True):
    plan = gj.Plan()
    plan.new_input(gj.make_table(build, len(build[0][1])))
    plan.new_input(gj.make_table(probe, len(probe[0][1])))
    b = plan.new_scan_node(0, [(i, c[0]) for i, c in enumerate(build)])
    q = plan.new_scan_node(1, [(i, c[0]) for i, c in enumerate(probe)])
    plan.root = plan.new_join_node(build_left, b, q, bkey, pkey, out)
    return plan


def _mixed_job(seed):
    """A two-table join of varied shape: key width, NULLs, duplicates, sizes
    (either side may be the smaller), and 1-4 outputs drawn from the probe
    key and both sides' payloads."""
    rng = np.random.default_rng(900 + seed)
    kt = [INT32, INT64][seed % 2]
    nb = int(rng.choice([0, 1, 70, 3000, 40_000]))
    npr = int(rng.choice([1, 500, 9000, 120_000]))
    dup = seed % 3 == 0
    nullf = [0.0, 0.1, 0.5][seed % 3]
    dt = np.int32 if kt == INT32 else np.int64
    dom = max(nb // (5 if dup else 1), 1)
    bkeys = (rng.integers(0, dom, nb) if dup else rng.permutation(nb)).astype(dt)
    pkeys = rng.integers(0, int(dom * 1.1) + 1, npr).astype(dt)

    def pay(n):
        t = [INT32, INT64, FP64][int(rng.integers(0, 3))]
        v = rng.random(n) * 1e6 if t == FP64 else rng.integers(-(1 << 30), 1 << 30, n)
        return (t, v, rng.random(n) >= nullf)

    build = [(kt, bkeys, rng.random(nb) >= nullf / 2)] + [pay(nb) for _ in range(int(rng.integers(1, 4)))]
    probe = [(kt, pkeys, rng.random(npr) >= nullf / 2)] + [pay(npr) for _ in range(int(rng.integers(1, 3)))]
    layout = [c[0] for c in build] + [c[0] for c in probe]
    nbc = len(build)
    choices = [nbc] + list(range(1, nbc)) + list(range(nbc + 1, len(layout)))
    k = int(rng.integers(1, 5))
    picks = [int(x) for x in rng.choice(choices, size=k, replace=len(choices) < k)]
    return _two_table_plan(build, probe, 0, 0, [(i, layout[i]) for i in picks], build_left=bool(seed % 2))


def _run_and_check(ctx, plans, label):
    tables = [[ctx.upload(t) for t in p.inputs] for p in plans]
    batch = ctx.batch(plans, tables)
    for rep in range(3):  # graph replays, each checked against the oracle
        batch.run()
        say(f"{label}: run {rep} {batch.info()}")
        for j, plan in enumerate(plans):
            types = [t for _, t in plan.nodes[plan.root].output_attrs]
            got = check_table(batch.download(j), types)
            if rep == 0:
                exp = check_table(oracle.execute(plan), types)
                plan._expected = exp
            assert got == plan._expected, (label, rep, j)
            assert batch.num_rows(j) == len(plan._expected)
    return batch, tables


def test_batch_mixed_shapes_vs_oracle(ctx):
    plans = [_mixed_job(s) for s in range(24)]
    say("batch of 24 mixed joins")
    batch, _ = _run_and_check(ctx, plans, "mixed")
    info = batch.info()
    assert info["jobs"] == 24 and info["tasks"] > 0 and info["waves"] > 0


def test_batch_matches_one_launch_per_join(ctx):
    """The c4 recipe at 1% scale: each job's rows as the plain path returns them."""
    from gpujoin import workloads

    wl = workloads.c4(scale=0.01, njobs=40)
    tables = [[ctx.generate(j.seed, t.rows, t.columns) for t in j.tables] for j in wl.jobs]
    plans = [j.make_plan() for j in wl.jobs]
    batch = ctx.batch(plans, tables)
    batch.run()
    say(f"c4 at 1%: {batch.info()}")
    for j, plan in enumerate(plans):
        types = [t for _, t in plan.nodes[plan.root].output_attrs]
        r = ctx.execute_device(plan, tables[j])
        plain = check_table(r.download(), types)
        r.free()
        assert check_table(batch.download(j), types) == plain, j
        # every job against the oracle's CPU join over the same input pages
        assert_join_result(batch.download_pages(j), tables[j], [(0, 1), (1, 1)], j)
    # and one job against the oracle on host-generated pages
    job = wl.jobs[3]
    host = job.make_plan()
    for t in job.tables:
        host.new_input(t.host_table(job.seed))
    types = [t for _, t in host.nodes[host.root].output_attrs]
    assert check_table(batch.download(3), types) == check_table(oracle.execute(host), types)


def test_batch_autotune_keeps_results(ctx):
    plans = [_mixed_job(s) for s in range(6, 14)]
    tables = [[ctx.upload(t) for t in p.inputs] for p in plans]
    batch = ctx.batch(plans, tables)
    batch.run()
    before = [batch.download(j) for j in range(len(plans))]
    batch.autotune()
    batch.run()
    say(f"autotuned: {batch.info()}")
    for j, plan in enumerate(plans):
        types = [t for _, t in plan.nodes[plan.root].output_attrs]
        assert check_table(batch.download(j), types) == check_table(before[j], types)


def test_batch_rejects_other_shapes(ctx):
    keys = np.arange(100, dtype=np.int32)
    plan = gj.Plan()
    for _ in range(3):
        plan.new_input(gj.make_table([(INT32, keys, None)], 100))
    a = plan.new_scan_node(0, [(0, INT32)])
    b = plan.new_scan_node(1, [(0, INT32)])
    c = plan.new_scan_node(2, [(0, INT32)])
    ab = plan.new_join_node(True, a, b, 0, 0, [(0, INT32)])
    plan.root = plan.new_join_node(True, ab, c, 0, 0, [(0, INT32)])
    tables = [[ctx.upload(t) for t in plan.inputs]]
    with pytest.raises(RuntimeError, match="two scans"):
        ctx.batch([plan], tables)
