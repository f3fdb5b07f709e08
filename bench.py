#!/usr/bin/env python3
"""Join-pipeline benchmark on one MI355X.

Default workload: 50M-row INT64 build x 500M-row Zipf probe with 5% NULLs
(c2), generated directly in HBM. A step is one full query over the
device-resident inputs: hash-table build pipeline, insert, fused
probe/output pipeline writing contest pages into HBM. Prints one JSON line.

Algorithmic bytes per query: every referenced 8 KiB input page read once,
every output page written once (counted at the reference's page packing), and
the compact hash-table tuples written once and read once.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "contest-sigmod2025_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

METRIC = "probe Gtuples/s & %HBM peak; batched+learned vs plain plan time at q-err≤4"
HBM_PEAK_GBS = 8000.0
PAGE = 8192


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(name, target_seconds=15.0, max_fraction=0.2):
    """Times the CPU oracle (single thread) on a bounded sample of the workload:
    the same recipe with every table scaled down so the timed run takes
    ~10-30 s (the sample grows until it does)."""
    import oracle
    import gpujoin as gj
    from gpujoin import workloads

    frac = 0.002
    t_run = 0.0
    rows = 0
    while True:
        wl = workloads.WORKLOADS[name](frac)
        plan = wl.make_plan()
        for t in wl.tables:
            plan.new_input(t.host_table(wl.seed))
        rows = sum(t.rows for t in wl.tables)
        t0 = time.perf_counter()
        res = oracle.execute(plan)
        t_run = time.perf_counter() - t0
        log(f"cpu baseline: sample {frac:.4f} of {name} ({rows} tuples, {res.num_rows} out) took {t_run:.2f} s")
        if t_run >= 10.0 or frac >= max_fraction:
            break
        frac = min(max_fraction, frac * max(1.5, min(8.0, target_seconds / max(t_run, 1e-3))))
    return {
        "value": rows / t_run / 1e9,
        "unit": "Gtuples/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{name} recipe at {frac:.4f} scale: {rows} input tuples, CPU oracle execute() incl. page decode and page output",
    }


def load_traffic(name):
    path = os.path.join(ROOT, "profiles", f"{name}_probe_traffic.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return None


def pcie_inclusive(ctx, name, scale, reps=3):
    """Host-to-host rate of the contest entry point: each call uploads the
    scanned columns' pages, runs the query and downloads the result into host
    pages (one operator-new block per page), as Contest::execute does. The
    library times its own upload and download; the Python wall time adds the
    bindings' per-page marshalling, which a C++ caller does not pay."""
    from gpujoin import workloads
    wl = workloads.WORKLOADS[name](scale)
    plan = wl.make_plan()
    for t in wl.tables:
        plan.new_input(t.host_table(wl.seed))
    tuples = sum(t.rows for t in wl.tables)
    res = ctx.execute(plan)  # warm-up
    rows = res.num_rows
    del res
    times, lib_ms = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        res = ctx.execute(plan)
        times.append(time.perf_counter() - t0)
        st = ctx.stats()
        lib_ms.append((st["ms_host_upload"], st["ms_total"], st["ms_host_download"]))
        del res
    t = min(times)
    up, dev, down = min(lib_ms, key=sum)
    t_lib = (up + dev + down) / 1000.0
    log(f"pcie-inclusive: {name} at scale {scale}: {t * 1000:.1f} ms through Python; library: upload {up:.1f} ms, "
        f"device {dev:.1f} ms, download {down:.1f} ms")
    return {"value": round(tuples / t_lib / 1e9, 4), "unit": "Gtuples/s", "scale": scale, "tuples": tuples,
            "output_rows": rows, "ms_library": round(t_lib * 1000.0, 2), "ms_upload": round(up, 2),
            "ms_device": round(dev, 2), "ms_download": round(down, 2), "ms_through_python": round(t * 1000.0, 2),
            "path": "host pages in, host pages out (upload, execute, download) as timed inside cgj_execute; "
                    "best of 3"}


def run_job(args):
    """The 113 JOB queries (the reference's plans.json through the JOB
    adapter) over synthetic inputs at `scale` x PostgreSQL's row estimates of
    the filtered IMDB scans. A step runs every query once, device-resident,
    with the per-query column indexing; inputs are generated before the timed
    region. The CPU baseline is the oracle executing the queries one after
    another until ~15 s of work."""
    import gpujoin as gj
    from gpujoin import job

    fixture = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "job_queries.json")
    js = job.load_set(fixture)
    ctx = gj.Context(0)
    t0 = time.perf_counter()
    inputs = [job.device_inputs(ctx, js, qi, args.scale) for qi in range(len(js.plans))]
    rows_in = sum(t.num_rows for ins in inputs for t in ins)
    log(f"job: {len(js.plans)} queries, {rows_in} input rows generated in {time.perf_counter() - t0:.1f} s")
    for _ in range(args.warmup):
        for jp, ins in zip(js.plans, inputs):
            ctx.execute_device(jp.plan, ins).free()
    per_step, per_query = [], [0.0] * len(js.plans)
    out_rows = pipelines = 0
    for step in range(args.steps):
        t_start = time.perf_counter()
        out_rows = pipelines = 0
        for qi, (jp, ins) in enumerate(zip(js.plans, inputs)):
            tq = time.perf_counter()
            for t in ins:
                t.prepare()
            r = ctx.execute_device(jp.plan, ins)
            out_rows += r.num_rows
            pipelines += ctx.stats()["pipelines"]
            r.free()
            per_query[qi] += time.perf_counter() - tq
        per_step.append(time.perf_counter() - t_start)
        log(f"job step {step}: {per_step[-1] * 1000:.1f} ms, {out_rows} result rows")
    s_step = min(per_step)
    slowest = sorted(range(len(js.plans)), key=lambda i: -per_query[i])[:5]
    line = {
        "metric": "JOB: 113 queries end-to-end (column indexing + planning + device-resident execution)",
        "value": round(len(js.plans) / s_step, 2),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(s_step * 1000.0, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32 joins, varchar outputs",
        "data": "synthetic",
        "config": {"workload": f"job: 113 JOB plans (PostgreSQL hash-join plans of the reference's plans.json), "
                               f"inputs at {args.scale} x the estimated filtered scan sizes ({rows_in} rows)"},
        "input_rows": rows_in,
        "result_rows_per_step": out_rows,
        "pipelines_per_step": pipelines,
        "slowest_ms": {js.names[i]: round(per_query[i] * 1000.0 / args.steps, 2) for i in slowest},
    }
    for ins in inputs:
        for t in ins:
            t.free()
    ctx.close()
    if not args.no_cpu_baseline:
        import oracle
        t_run, n = 0.0, 0
        for qi, jp in enumerate(js.plans):
            plan = job.with_inputs(jp, job.host_inputs(js, qi, args.scale))
            t1 = time.perf_counter()
            oracle.execute(plan)
            t_run += time.perf_counter() - t1
            n += 1
            plan.inputs = []
            if t_run >= 15.0:
                break
        log(f"cpu baseline: {n} JOB queries in {t_run:.2f} s")
        line["cpu_baseline"] = {"value": round(n / t_run, 3), "unit": "queries/s", "cores": 1, "kind": "port",
                                "sample": f"the first {n} queries in order, CPU oracle execute(), host inputs"}
    print(json.dumps(line), flush=True)


def run_c4(args):
    """256 independent INT32 joins (build log-uniform [1K, 2M] rows, probe 10x)
    on device-resident tables. The batched form runs all of them from one
    hipGraph: fill kernels, one persistent build launch over a device queue
    of every job's build tasks, one persistent probe launch over the probe
    tasks; task sizes, load factor and bloom bits come from the batch
    autotuner (tuned once before the timed steps). Beside it: the same joins
    as one execute_device call each (the plain one-join-per-launch path, as
    the reference runs one Contest::execute per query), and all joins in one
    cgj_execute_many call on worker streams. A step is one pass over all
    joins; every form's row counts must agree."""
    import gpujoin as gj
    from gpujoin import workloads

    bw = workloads.c4(args.scale)
    ctx = gj.Context(0)
    for o in args.option:
        name, value = o.split("=", 1)
        ctx.set_option(name, float(value))
    t0 = time.perf_counter()
    jobs = []
    for w in bw.jobs:
        jobs.append((w.make_plan(), [ctx.generate(w.seed, t.rows, t.columns) for t in w.tables]))
    tuples = sum(t.rows for w in bw.jobs for t in w.tables)
    log(f"c4: {len(jobs)} joins, {tuples} tuples generated in {time.perf_counter() - t0:.1f} s")
    plans = [p for p, _ in jobs]
    tabs_all = [t for _, t in jobs]

    # batched: one graph launch per step
    t0 = time.perf_counter()
    batch = ctx.batch(plans, tabs_all)
    batch.run()
    t_first = time.perf_counter() - t0
    t_tune = 0.0
    if not args.no_autotune:
        t0 = time.perf_counter()
        batch.autotune()
        t_tune = time.perf_counter() - t0
        log(f"c4 batch autotuned in {t_tune:.1f} s: {batch.info()}")
    for _ in range(args.warmup):
        batch.run()
    per_step, dev_ms = [], []
    for step in range(args.steps):
        t_start = time.perf_counter()
        batch.run()
        per_step.append(time.perf_counter() - t_start)
        dev_ms.append(batch.info()["ms_last"])
        log(f"c4 batch step {step}: {per_step[-1] * 1000:.2f} ms host, {dev_ms[-1]:.2f} ms graph")
    s_batch = sum(per_step) / len(per_step)
    info = batch.info()
    rows_batch = [batch.num_rows(j) for j in range(len(jobs))]

    # one execute_device per join, one after another (plain path)
    seq = []
    ms_dev = 0.0
    rows_plain = []
    for step in range(max(1, args.steps // 2)):
        t_start = time.perf_counter()
        rows_plain = []
        for plan, tabs in jobs:
            r = ctx.execute_device(plan, tabs)
            ms_dev += ctx.stats()["ms_total"]
            rows_plain.append(r.num_rows)
            r.free()
        seq.append(time.perf_counter() - t_start)
    s_seq = sum(seq) / len(seq)
    log(f"c4 one join per call: {s_seq * 1000:.1f} ms")
    if rows_plain != rows_batch:
        raise SystemExit("c4: batched row counts differ from one call per join")

    # all joins in one cgj_execute_many call on worker streams
    for r in ctx.execute_many(plans, tabs_all, workers=args.workers):
        r.free()
    many = []
    for step in range(max(1, args.steps // 2)):
        t_start = time.perf_counter()
        res = ctx.execute_many(plans, tabs_all, workers=args.workers)
        many.append(time.perf_counter() - t_start)
        for r in res:
            r.free()
    s_many = sum(many) / len(many)
    log(f"c4 execute_many ({args.workers} workers): {s_many * 1000:.1f} ms")

    # algorithmic bytes: input pages read once, output pages at the reference's
    # packing written once, compact build tuples (INT32 key + INT32 payload)
    # written and read once
    budget32 = 8 * (PAGE - 4)
    in_pages = sum(t.num_pages(c) for _, tabs in jobs for t in tabs for c in range(2))
    out_pages = sum(2 * ((r + 32 * r + budget32 - 1) // budget32) for r in rows_batch)
    build_rows = sum(w.tables[0].rows for w in bw.jobs)
    b_alg = (in_pages + out_pages) * PAGE + 2 * 8 * build_rows
    g_ms = sum(dev_ms) / len(dev_ms)
    achieved = b_alg / (g_ms / 1000.0) / 1e9
    line = {
        "metric": "c4 batched joins: Gtuples/s over 256 independent joins in one graph launch",
        "value": round(tuples / s_batch / 1e9, 4),
        "unit": "Gtuples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(s_batch * 1000.0, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"c4: {bw.description}",
                   "path": "cgj_batch: one hipGraph per step (fills, persistent build launch, persistent probe "
                           "launch over device task queues), autotuned task sizes / load factor / bloom bits"},
        "joins_per_s": round(len(jobs) / s_batch, 1),
        "output_rows_per_step": int(sum(rows_batch)),
        "batch": {"graph_ms": round(g_ms, 3), "tasks": info["tasks"], "waves": info["waves"],
                  "launches": info["launches"], "s_first_run": round(t_first, 2), "s_autotune": round(t_tune, 2)},
        "one_call_per_join": {"ms_per_step": round(s_seq * 1000.0, 2), "joins_per_s": round(len(jobs) / s_seq, 1),
                              "ms_device_per_step": round(ms_dev / len(seq), 2)},
        "execute_many": {"ms_per_step": round(s_many * 1000.0, 2), "workers": args.workers},
        "speedup_vs_one_call_per_join": round(s_seq / s_batch, 3),
        "speedup_vs_execute_many": round(s_many / s_batch, 3),
        "roofline": {"kernel": "the batch graph (k_batch_build + k_batch_probe<2>)", "bound": "hbm",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "algorithmic_bytes": int(b_alg),
                     "kernel_ms": round(g_ms, 3)},
    }
    print(json.dumps(line), flush=True)
    batch.free()
    for _, tabs in jobs:
        for t in tabs:
            t.free()
    ctx.close()


def c5_cpu_baseline(ps, eval_ids, target_seconds=15.0):
    """The CPU oracle (one thread) executing evaluated plans in their given
    join order until ~15 s of work; host generation of the pool tables is not
    timed."""
    import oracle
    host = {}
    t_run = 0.0
    n = 0
    for pid in eval_ids:
        pp = ps.plans[pid]
        for t in pp.tables:
            if t not in host:
                host[t] = ps.pool[t].host_table(ps.seed)
        plan = pp.plan
        plan.inputs = [host[t] for t in pp.tables]
        t0 = time.perf_counter()
        oracle.execute(plan)
        t_run += time.perf_counter() - t0
        plan.inputs = []
        n += 1
        if t_run >= target_seconds:
            break
    log(f"cpu baseline: {n} c5 plans in {t_run:.2f} s")
    return {"value": n / t_run, "unit": "plans/s", "cores": 1, "kind": "port",
            "sample": f"the first {n} evaluated plans, CPU oracle execute() in each plan's given join order"}


def run_c5(args):
    """Plan set (3-12-table chains/stars/trees over a shared table pool):
    trains the estimator on the first `train_plans` plans' connected sets
    (exact counts), then executes the next `plans` plans device-resident:
    sequentially, planned the reference's way (the max-of-inputs rule over
    64-row samples, re-sampled and re-planned after every hash-table build:
    the plain path), sequentially under the learned estimator, and all in one
    cgj_execute_many call under the learned estimator (each worker plans and
    runs its queries on its own stream: the batched + learned path, reported
    as `value`). A step is one pass over the evaluated plans. Planning
    (sampling, statistics, inference, DP) is inside the timed region: the
    learned steps start from an empty catalog of pairwise join counts, so the
    exact 2-way counts the learned features read are counted and timed in
    every step."""
    import numpy as np
    import gpujoin as gj
    from gpujoin import estimation as est
    from gpujoin import workloads

    ps = workloads.c5(args.scale, nplans=args.train_plans + args.plans)
    ctx = gj.Context(0)
    t0 = time.perf_counter()
    tabs = [ctx.generate(ps.seed, t.rows, t.columns) for t in ps.pool]
    log(f"c5 pool: {len(tabs)} tables, {sum(t.rows for t in ps.pool)} rows in {time.perf_counter() - t0:.1f} s")
    counts = {}
    t0 = time.perf_counter()

    def progress(what):
        last = [time.perf_counter()]

        def f(done):
            if time.perf_counter() - last[0] > 20.0:
                last[0] = time.perf_counter()
                log(f"{what}: {done} plans, {time.perf_counter() - t0:.0f} s")
        return f

    train_s = est.collect(ctx, ps, tabs, range(args.train_plans), counts=counts, progress=progress("training sets"))
    t_collect = time.perf_counter() - t0
    log(f"training sets: {sum(len(s.rows) for s in train_s)} from {args.train_plans} plans in {t_collect:.1f} s")
    t0 = time.perf_counter()
    mlp, loss = est.train(ctx, train_s, epochs=args.epochs)
    t_train = time.perf_counter() - t0
    log(f"trained in {t_train:.1f} s, loss {loss:.4f}")
    eval_ids = list(range(args.train_plans, args.train_plans + args.plans))
    t0 = time.perf_counter()
    held = est.collect(ctx, ps, tabs, eval_ids[:min(100, len(eval_ids))], counts=counts, seed=8,
                       progress=progress("held-out sets"))
    if args.dump:
        est.save(args.dump + "_train.npz", train_s)
        est.save(args.dump + "_held.npz", held)
    true = np.concatenate([s.rows for s in held])
    q = {}
    for name, mode in (("sample", gj.EST_SAMPLE), ("max", gj.EST_MAX), ("independence", gj.EST_INDEPENDENCE),
                       ("learned", gj.EST_LEARNED)):
        qe = est.qerror(est.estimates(held, ps, mode, mlp), true)
        q[name] = {"median": round(float(np.median(qe)), 3), "p90": round(float(np.percentile(qe, 90)), 3)}
    log(f"held-out q-error: {q}")

    def sequential(name, mode):
        ctx.set_estimator(mode, mlp if mode == gj.EST_LEARNED else None)
        for pid in eval_ids[:args.warmup * 20]:
            ctx.execute_device(ps.plans[pid].plan, est.plan_inputs(tabs, ps.plans[pid])).free()
        t_rows = 0
        ms_plan = ms_stats = 0.0
        pipelines = retries = retries_build = replans = stat_joins = 0
        ms_build = ms_probe = 0.0
        per_step = []
        for step in range(args.steps):
            ctx.clear_statistics()
            t_start = time.perf_counter()
            for k, pid in enumerate(eval_ids):
                pp = ps.plans[pid]
                r = ctx.execute_device(pp.plan, est.plan_inputs(tabs, pp))
                st = ctx.stats()
                ms_plan += st["ms_plan"]
                ms_stats += st["ms_stats"]
                stat_joins += st["est_stat_joins"]
                pipelines += st["pipelines"]
                replans += st["replans"]
                ms_build += st["ms_build"]
                ms_probe += st["ms_probe"]
                retries += st["retries"]
                retries_build += st["retries_build"]
                t_rows += r.num_rows
                r.free()
                if k % 250 == 249:
                    log(f"{name} step {step}: {k + 1} plans, {time.perf_counter() - t_start:.1f} s")
            per_step.append(time.perf_counter() - t_start)
            log(f"{name} step {step}: {per_step[-1]:.2f} s")
        n = args.steps
        return {"s_per_step": round(min(per_step), 4), "ms_plan_per_step": round(ms_plan / n, 1),
                "ms_statistics_per_step": round(ms_stats / n, 1), "stat_joins_per_step": stat_joins // n,
                "rows_per_step": t_rows // n, "pipelines_per_step": pipelines // n,
                "replans_per_step": replans // n, "retries_per_step": retries // n,
                "build_retries_per_step": retries_build // n,
                "ms_device_build_per_step": round(ms_build / n, 1),
                "ms_device_probe_per_step": round(ms_probe / n, 1)}

    results = {}
    for name, mode in (("sample", gj.EST_SAMPLE), ("learned", gj.EST_LEARNED)):
        results[name] = sequential(name, mode)
        log(f"{name}: {results[name]}")
    # batched + learned: every evaluated plan in one cgj_execute_many call;
    # each worker plans its queries with the learned estimator (pairwise
    # counts from an empty catalog, shared between the workers) and runs them
    # on its own stream
    ctx.set_estimator(gj.EST_LEARNED, mlp)
    all_plans = [ps.plans[pid].plan for pid in eval_ids]
    all_tabs = [est.plan_inputs(tabs, ps.plans[pid]) for pid in eval_ids]
    for r in ctx.execute_many(all_plans[:args.warmup * 20], all_tabs[:args.warmup * 20], workers=args.workers):
        r.free()
    per_step = []
    for step in range(args.steps):
        ctx.clear_statistics()
        t_start = time.perf_counter()
        res = ctx.execute_many(all_plans, all_tabs, workers=args.workers)
        per_step.append(time.perf_counter() - t_start)
        rows = sum(r.num_rows for r in res)
        for r in res:
            r.free()
        log(f"batched+learned step {step}: {per_step[-1]:.2f} s, {rows} rows")
    s_batched = min(per_step)
    results["batched_learned"] = {"s_per_step": round(s_batched, 4), "workers": args.workers, "rows_per_step": rows,
                                  "s_steps": [round(x, 3) for x in per_step],
                                  "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or 4)}
    if rows != results["sample"]["rows_per_step"] or rows != results["learned"]["rows_per_step"]:
        raise RuntimeError(f"c5: result rows differ between estimators: {rows} (batched + learned), "
                           f"{results['sample']['rows_per_step']} (sampling), {results['learned']['rows_per_step']} (learned)")
    ctx.set_estimator(gj.EST_SAMPLE)
    s_plain = results["sample"]["s_per_step"]
    line = {
        "metric": "c5 plan set: plans/s end-to-end (planning + device-resident execution), batched + learned",
        "value": round(args.plans / s_batched, 2),
        "unit": "plans/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(s_batched * 1000.0, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32 joins; fp32/bf16 estimator",
        "data": "synthetic",
        "config": {"workload": f"c5: {ps.description}; evaluated plans {args.plans}, training plans {args.train_plans}",
                   "plain": "one query at a time, sampling-planned (max rule over 64-row samples, "
                            "re-planned after every build)",
                   "statistics": "pairwise join counts of the learned features counted inside every timed step"},
        "speedup_batched_learned_vs_plain": round(s_plain / s_batched, 3),
        "speedup_learned_vs_plain_sequential": round(s_plain / results["learned"]["s_per_step"], 3),
        "plain_sampling": results["sample"],
        "learned": results["learned"],
        "batched_learned": results["batched_learned"],
        "qerror_held_out": q,
        "training": {"sets": int(sum(len(s.rows) for s in train_s)), "s_collect": round(t_collect, 1),
                     "s_train": round(t_train, 1), "epochs": args.epochs, "final_loss": round(loss, 4)},
    }
    for t in tabs:
        t.free()
    ctx.close()
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = c5_cpu_baseline(ps, eval_ids)
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--scale", type=float, default=None,
                    help="table scale (default 1: the stated table sizes)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="c4: skip the batch autotuner")
    ap.add_argument("--option", action="append", default=[], metavar="NAME=VALUE",
                    help="execution option (cgj_set_option), repeatable")
    ap.add_argument("--workers", type=int, default=None,
                    help="c4/c5: concurrent worker streams of cgj_execute_many (default: c4 6, c5 8)")
    ap.add_argument("--plans", type=int, default=2000, help="c5: evaluated plans")
    ap.add_argument("--train-plans", type=int, default=150, help="c5: training plans")
    ap.add_argument("--epochs", type=int, default=150, help="c5: training epochs")
    ap.add_argument("--dump", default="", help="c5: write training/held-out sets to PREFIX_{train,held}.npz")
    ap.add_argument("--pcie", type=float, default=0.0, metavar="SCALE",
                    help="also time the host-pages-in/host-pages-out path at this table scale")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the join pipeline runs on a single MI355X "
              f"(--gpus {args.gpus} is not supported: sharding across GPUs is a design, not built)",
              file=sys.stderr)
        sys.exit(2)

    if args.scale is None:
        args.scale = 1.0
    if args.workers is None:  # the fastest counts measured on one MI355X
        args.workers = 8 if args.workload == "c5" else 6
    if args.workload in ("c4", "c5"):
        # One hardware queue per worker stream (plus the calling context's):
        # with fewer, HIP maps several streams onto one queue and their
        # kernels run one after another (c5: 2.0-3.5 s per step at HIP's
        # default of 4 queues, 1.8-2.0 s at 12). Read at HIP's initialisation,
        # which has not happened yet.
        want = min(32, args.workers + 2)
        if int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or 4) < want:
            os.environ["GPU_MAX_HW_QUEUES"] = str(want)
    if args.workload == "c5":
        return run_c5(args)
    if args.workload == "c4":
        return run_c4(args)
    if args.workload == "job":
        return run_job(args)

    import gpujoin as gj
    from gpujoin import workloads

    wl = workloads.WORKLOADS[args.workload](args.scale)
    ctx = gj.Context(0)
    for o in args.option:
        name, value = o.split("=", 1)
        ctx.set_option(name, float(value))
    t0 = time.perf_counter()
    tables = []
    for t in wl.tables:
        if t.host:
            log(f"generating {t.name}: {t.rows} rows on the host, uploading")
            tables.append(ctx.upload(t.host_table(wl.seed)))
        else:
            log(f"generating {t.name}: {t.rows} rows in HBM")
            tables.append(ctx.generate(wl.seed, t.rows, t.columns))
    log(f"generated in {time.perf_counter() - t0:.1f} s")
    plan = wl.make_plan()
    tuples = sum(t.rows for t in wl.tables)

    for i in range(args.warmup):
        r = ctx.execute_device(plan, tables)
        log(f"warmup {i}: {r.num_rows} rows, {ctx.stats()['ms_total']:.2f} ms device")
        r.free()

    # A step indexes every scanned column (rows-per-page prefix and 64-row
    # block index: the search vectors the reference builds inside every
    # execute()) and runs the query. execute_device returns after the query's
    # last kernel has completed (the library synchronises its stream), so the
    # host clock brackets whole steps.
    stats = []
    prep = []
    t_start = time.perf_counter()
    for i in range(args.steps):
        tp = time.perf_counter()
        for t in tables:
            t.prepare()
        prep.append(time.perf_counter() - tp)
        r = ctx.execute_device(plan, tables)
        s = ctx.stats()
        s["rows"] = r.num_rows
        stats.append(s)
        r.free()
    elapsed = time.perf_counter() - t_start
    ms_step = elapsed * 1000.0 / args.steps
    for i, s in enumerate(stats):
        part = f", probe partition {s['ms_probe_partition']:.2f} ms" if s["ms_probe_partition"] else ""
        log(f"step {i}: build {s['ms_build']:.2f} ms, probe kernel {s['ms_probe_kernel']:.2f} ms{part}, rows {s['rows']}")

    # achievable streaming bandwidth on this GPU: a plain 2 GiB copy (reads + writes)
    copy_bytes = 2 << 30
    copy_ms = ctx.copy_probe(copy_bytes, 5)
    copy_gbs = 2 * copy_bytes / (copy_ms / 1000.0) / 1e9

    s = stats[-1]
    b_alg = (s["input_pages"] + s["output_ref_pages"]) * PAGE + 2 * s["build_bytes"]
    b_kernel = (s["probe_input_pages"] + s["output_ref_pages"]) * PAGE + s["build_bytes"]
    k_ms = sum(x["ms_probe_kernel"] for x in stats) / len(stats)
    achieved = b_kernel / (k_ms / 1000.0) / 1e9
    traffic = load_traffic(wl.name)
    if s["probe_kernel"] == 1:
        kernel_name = "k_lean (one-probe pipeline: page decode, hash probe, output pages)"
        if s["probe_packed"]:
            kernel_name += "; hash table read packed, 8 bytes per slot"
    else:
        kernel_name = "k_pipeline (fused probe chain: page decode, probes, output pages)"
    notes = {}
    if wl.name == "c2":
        # the join tuner turns bloom filters on for builds of at most 16M keys
        # probed by at least 8x as many rows (planner.cpp autotune_join)
        notes["bloom"] = ("off: the build has 50M keys, above the 16M-key limit of the join tuner's bloom "
                          "filters; the 10% missing probe keys walk the hash table to an empty slot")
    line = {
        "metric": METRIC,
        "value": round(tuples / (ms_step / 1000.0) / 1e9, 4),
        "unit": "Gtuples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": {
            "workload": f"{wl.name}: {wl.description}",
            "tuples": tuples,
            "output_rows": int(s["rows"]),
            "device_resident": True,
            "step": "index every scanned column (row prefix + block index), then run the query",
            **notes,
        },
        "ms_prepare_per_step": round(1000.0 * sum(prep) / len(prep), 3),
        "hbm_frac_job": round(b_alg / (ms_step / 1000.0) / (HBM_PEAK_GBS * 1e9), 4),
        "b_alg_bytes": int(b_alg),
        "roofline": {
            "kernel": kernel_name,
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic["bytes_per_launch"] if traffic else None,
            "algorithmic_bytes": int(b_kernel),
            "kernel_ms": round(k_ms, 3),
            "copy_gbs": round(copy_gbs, 1),
        },
    }
    for t in tables:
        t.free()
    if args.pcie:
        line["pcie_inclusive"] = pcie_inclusive(ctx, args.workload, args.pcie)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(wl.name)
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
