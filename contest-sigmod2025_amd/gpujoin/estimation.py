"""Training data, training and evaluation of the learned cardinality estimator
over a plan set whose plans share one pool of device tables.

A training row is one connected relation set of a plan: its features
(Context.plan_features) and the log10 of its exact cardinality
(Context.count_subset, one count-only join pipeline per set). The network
predicts (mu, log sigma) of the residual of that log10 cardinality over the
independence estimate (feature 10 x 8, at least 0): the engine adds mu to
the independence estimate, and keeps the independence estimate where sigma
exceeds the confidence bar.
"""
from __future__ import annotations

import math
import random
from dataclasses import dataclass, field
from typing import Dict, List, Sequence, Tuple

import numpy as np

from . import EST_INDEPENDENCE, EST_LEARNED, EST_MAX, EST_SAMPLE, Context, DeviceTable, Mlp


@dataclass
class SetSamples:
    plan: int
    sets: np.ndarray       # uint64 relation sets
    features: np.ndarray   # n x num_features float32
    rows: np.ndarray       # exact cardinalities (float64)
    sampled: np.ndarray = None  # the sampling estimate of each set (largest sampled relation)


def plan_inputs(pool_tables: Sequence[DeviceTable], pp) -> List[DeviceTable]:
    return [pool_tables[t] for t in pp.tables]


def collect(ctx: Context, planset, pool_tables: Sequence[DeviceTable], plan_ids: Sequence[int],
            max_sets_per_plan: int = 48, seed: int = 7, counts: Dict = None, progress=None) -> List[SetSamples]:
    """Features and exact cardinalities of up to `max_sets_per_plan` connected
    sets per plan (all sets of the full plan included). `counts` caches exact
    counts by (plan, set) across calls; `progress(plans_done)` is called after
    every plan."""
    rng = random.Random(seed)
    out = []
    for done, pid in enumerate(plan_ids):
        if progress is not None and done:
            progress(done)
        pp = planset.plans[pid]
        tabs = plan_inputs(pool_tables, pp)
        sets, feats = ctx.plan_features(pp.plan, tabs)
        rel_cards = ctx.sample_cardinalities(pp.plan, tabs)
        full = (1 << len(pp.tables)) - 1
        idx = list(range(len(sets)))
        if len(idx) > max_sets_per_plan:
            keep = [i for i in idx if int(sets[i]) == full]
            rest = [i for i in idx if int(sets[i]) != full]
            keep += rng.sample(rest, max_sets_per_plan - len(keep))
            idx = sorted(keep)
        rows = []
        for i in idx:
            key = (pid, int(sets[i]))
            if counts is not None and key in counts:
                c = counts[key]
            else:
                c = ctx.count_subset(pp.plan, tabs, int(sets[i]))
                if counts is not None:
                    counts[key] = c
            rows.append(float(c))
        sampled = np.array([max(rel_cards[r] for r in range(len(rel_cards)) if (int(s) >> r) & 1)
                            for s in sets[idx]])
        out.append(SetSamples(pid, sets[idx], feats[idx], np.array(rows), sampled))
    return out


def save(path: str, samples: Sequence[SetSamples]):
    """Writes samples as arrays (plan, set, features, rows, sampled) to an .npz file."""
    np.savez(path, plan=np.concatenate([np.full(len(s.rows), s.plan) for s in samples]),
             sets=np.concatenate([s.sets for s in samples]), features=np.concatenate([s.features for s in samples]),
             rows=np.concatenate([s.rows for s in samples]), sampled=np.concatenate([s.sampled for s in samples]))


def base_log10(features: np.ndarray) -> np.ndarray:
    """The independence estimate in log10 units, at least 0 (the engine's base)."""
    return np.maximum(features[:, 10].astype(np.float64) * 8.0, 0.0)


def labels(rows: np.ndarray, features: np.ndarray) -> np.ndarray:
    return (np.log10(np.maximum(rows, 1.0)) - base_log10(features)).astype(np.float32)


def stack(samples: Sequence[SetSamples]) -> Tuple[np.ndarray, np.ndarray]:
    x = np.concatenate([s.features for s in samples]).astype(np.float32)
    y = np.concatenate([labels(s.rows, s.features) for s in samples])
    return x, y


def train(ctx: Context, samples: Sequence[SetSamples], epochs: int = 150, batch: int = 256, lr: float = 1e-3,
          weight_decay: float = 1e-4, seed: int = 1) -> Tuple[Mlp, float]:
    """Trains the residual network. The defaults won a float64 sweep over
    c5 sets (held-out q-error p90 equal to the independence rule's, p99 lower)."""
    x, y = stack(samples)
    m = ctx.mlp(Context.num_features(), seed=seed)
    loss = m.train(x, y, epochs=epochs, batch=batch, lr=lr, seed=seed + 1, weight_decay=weight_decay)
    return m, loss


def qerror(est: np.ndarray, true: np.ndarray) -> np.ndarray:
    e = np.maximum(np.asarray(est, dtype=np.float64), 1.0)
    t = np.maximum(np.asarray(true, dtype=np.float64), 1.0)
    return np.maximum(e / t, t / e)


def estimates(samples: Sequence[SetSamples], planset, mode: int, mlp: Mlp = None,
              max_sigma: float = 0.5) -> np.ndarray:
    """Estimated cardinalities of the samples' sets under one model, computed
    on the host from the features (they hold the same inputs the engine uses):
    EST_SAMPLE = largest sampled relation (gathered on the device by collect),
    EST_MAX = largest relation, EST_INDEPENDENCE = feature 10 (x 8 = log10 of
    the independence estimate), EST_LEARNED = that estimate corrected by the
    network's residual (BF16 inference), the sampling estimate where the
    network's sigma exceeds `max_sigma`."""
    x = np.concatenate([s.features for s in samples])
    indep = 10.0 ** (x[:, 10].astype(np.float64) * 8.0)
    sampled = np.concatenate([s.sampled for s in samples])
    if mode == EST_SAMPLE:
        return sampled
    if mode == EST_MAX:
        return 10.0 ** (x[:, 11].astype(np.float64) * 8.0)
    if mode == EST_INDEPENDENCE:
        return indep
    out = mlp.predict(x)
    mu, s = out[:, 0].astype(np.float64), out[:, 1].astype(np.float64)
    base = base_log10(x)
    learned = 10.0 ** np.minimum(base + mu, 300.0)
    return np.where(np.exp(s) <= max_sigma, learned, sampled)


def true_cost(ctx: Context, planset, pool_tables, pid: int, counts: Dict) -> float:
    """Cost of the join order the context's current estimator picks, priced
    with exact cardinalities: per join, result rows + 2 x build rows (the
    planner's own cost function, engine/query/QueryGraph.cpp:46-50 restated
    for a hash join whose table is written and read once more)."""
    pp = planset.plans[pid]
    tabs = plan_inputs(pool_tables, pp)

    def card(s):
        if s & (s - 1) == 0:
            return float(tabs[s.bit_length() - 1].num_rows)
        key = (pid, s)
        if key not in counts:
            counts[key] = ctx.count_subset(pp.plan, tabs, s)
        return float(counts[key])

    cost = 0.0
    for b, p, _ in ctx.plan_joins(pp.plan, tabs):
        cost += card(b | p) + 2.0 * card(b)
    return cost


@dataclass
class Evaluation:
    qerror_median: Dict[str, float] = field(default_factory=dict)
    qerror_p90: Dict[str, float] = field(default_factory=dict)
    cost: Dict[str, float] = field(default_factory=dict)


def evaluate(ctx: Context, planset, pool_tables, samples: Sequence[SetSamples], mlp: Mlp, counts: Dict,
             max_sigma: float = 0.5) -> Evaluation:
    ev = Evaluation()
    true = np.concatenate([s.rows for s in samples])
    for name, mode in (("sample", EST_SAMPLE), ("max", EST_MAX), ("independence", EST_INDEPENDENCE),
                       ("learned", EST_LEARNED)):
        q = qerror(estimates(samples, planset, mode, mlp, max_sigma), true)
        ev.qerror_median[name] = float(np.median(q))
        ev.qerror_p90[name] = float(np.percentile(q, 90))
        ctx.set_estimator(mode, mlp if mode == EST_LEARNED else None, max_sigma)
        ev.cost[name] = sum(true_cost(ctx, planset, pool_tables, s.plan, counts) for s in samples)
    ctx.set_estimator(EST_SAMPLE)
    return ev
