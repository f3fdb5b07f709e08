"""Synthetic join workloads (seeded, generated directly in HBM or on the host).

Every column draws from its own counter-based SplitMix64 stream
(seed = 0x5EED0000 + config number), so the device generator and the host
generator produce the same rows and any row range can be generated alone.
"""
from __future__ import annotations

import math
import random
from dataclasses import dataclass
from typing import Callable, List

from . import (FP64, GEN_ASCII, GEN_FK_UNIFORM, GEN_FK_ZIPF, GEN_PERM_KEY, GEN_UNIFORM, GEN_UNIQUE_KEY,
               GEN_UNIT_F64, INT32, INT64, VARCHAR, ColumnarTable, GenColumn, Plan, host_generate)


@dataclass
class TableSpec:
    name: str
    rows: int
    columns: List[GenColumn]
    host: bool = False  # generated on the host (VARCHAR) and uploaded

    def host_table(self, seed) -> ColumnarTable:
        return ColumnarTable(self.rows, [host_generate(seed, 0, self.rows, c) for c in self.columns])


@dataclass
class Workload:
    name: str
    seed: int
    tables: List[TableSpec]
    make_plan: Callable[[], Plan]
    description: str


def c1(scale: float = 1.0) -> Workload:
    """1M-row INT32 build (shuffled unique keys + INT32 payload) joined with a
    10M-row probe (uniform FK + INT64 payload); no NULLs; output
    {build payload, probe payload}."""
    nb = max(int(1_000_000 * scale), 1)
    npr = max(int(10_000_000 * scale), 1)
    build = TableSpec("build", nb, [
        GenColumn(type=INT32, kind=GEN_PERM_KEY, stream=1, domain=nb),
        GenColumn(type=INT32, kind=GEN_UNIFORM, stream=2, lo=0, hi=1 << 31),
    ])
    probe = TableSpec("probe", npr, [
        GenColumn(type=INT32, kind=GEN_FK_UNIFORM, stream=1, domain=nb),
        GenColumn(type=INT64, kind=GEN_UNIFORM, stream=4, lo=0, hi=1 << 63),
    ])

    def plan():
        p = Plan()
        b = p.new_scan_node(0, [(0, INT32), (1, INT32)])
        q = p.new_scan_node(1, [(0, INT32), (1, INT64)])
        p.root = p.new_join_node(True, b, q, 0, 0, [(1, INT32), (3, INT64)])
        return p

    return Workload("c1", 0x5EED0001, [build, probe], plan,
                    f"INT32 equi-join, build {nb} unique keys, probe {npr} uniform FKs, no NULLs")


def c2(scale: float = 1.0) -> Workload:
    """50M-row INT64 build (unique shuffled keys, FP64 payload) joined with a
    500M-row probe (90% Zipf(1.0)-over-rank FKs, 10% misses from a disjoint
    range, FP64 payload); 5% Bernoulli NULLs in every column; output
    {build payload, probe payload, probe key}."""
    nb = max(int(50_000_000 * scale), 1)
    npr = max(int(500_000_000 * scale), 1)
    build = TableSpec("build", nb, [
        GenColumn(type=INT64, kind=GEN_UNIQUE_KEY, stream=1, null_stream=11, null_frac=0.05, domain=nb),
        GenColumn(type=FP64, kind=GEN_UNIT_F64, stream=2, null_stream=12, null_frac=0.05),
    ])
    probe = TableSpec("probe", npr, [
        GenColumn(type=INT64, kind=GEN_FK_ZIPF, stream=1, null_stream=13, null_frac=0.05, domain=nb, zipf_s=1.0,
                  miss_frac=0.1),
        GenColumn(type=FP64, kind=GEN_UNIT_F64, stream=4, null_stream=14, null_frac=0.05),
    ])

    def plan():
        p = Plan()
        b = p.new_scan_node(0, [(0, INT64), (1, FP64)])
        q = p.new_scan_node(1, [(0, INT64), (1, FP64)])
        p.root = p.new_join_node(True, b, q, 0, 0, [(1, FP64), (3, FP64), (2, INT64)])
        return p

    return Workload("c2", 0x5EED0002, [build, probe], plan,
                    f"INT64 equi-join, build {nb} unique keys, probe {npr} Zipf(1.0) FKs + 10% misses, 5% NULLs")


def c3(scale: float = 1.0) -> Workload:
    """Five-table snowflake: a 36M-row fact table with four INT32 foreign keys
    (Zipf s=0.8 over key rank) into dimensions of 2.5M, 4M, 1.5M and 0.3M keys
    that arrive pre-filtered to 1%, 5%, 20% and 100% of their rows; the last
    dimension carries a VARCHAR column (ASCII, 5-60 bytes, 0.1% long strings).
    Output {dimension VARCHAR, fact FP64}; all four joins fuse into one probe
    chain over the fact table."""
    nf = max(int(36_000_000 * scale), 1)
    dims = [(2_500_000, 0.01), (4_000_000, 0.05), (1_500_000, 0.20), (300_000, 1.0)]
    fact_cols = []
    tables = []
    for i, (n, sel) in enumerate(dims):
        dn = max(int(n * scale), 16)
        rows = max(int(dn * sel), 1)
        cols = [GenColumn(type=INT32, kind=GEN_PERM_KEY, stream=31 + i, domain=dn)]
        if i == 3:
            cols.append(GenColumn(type=VARCHAR, kind=GEN_ASCII, stream=35, lo=5, hi=60, miss_frac=0.001))
        tables.append(TableSpec(f"dim{i + 1}", rows, cols, host=True))
        # the fact side ranks keys with its own permutation, so the hot keys are
        # independent of which keys survived the dimension's filter
        fact_cols.append(GenColumn(type=INT32, kind=GEN_FK_ZIPF, stream=21 + i, domain=dn, zipf_s=0.8))
    fact_cols.append(GenColumn(type=FP64, kind=GEN_UNIT_F64, stream=25))
    tables.insert(0, TableSpec("fact", nf, fact_cols))

    def plan():
        p = Plan()
        f = p.new_scan_node(0, [(i, INT32) for i in range(4)] + [(4, FP64)])
        d = [p.new_scan_node(1 + i, [(0, INT32)]) for i in range(3)]
        d4 = p.new_scan_node(4, [(0, INT32), (1, VARCHAR)])
        keep = [(i, INT32) for i in range(4)] + [(4, FP64)]
        cur = f
        for i in range(3):
            cur = p.new_join_node(False, cur, d[i], i, 0, keep)
        p.root = p.new_join_node(False, cur, d4, 3, 0, [(6, VARCHAR), (4, FP64)])
        return p

    return Workload("c3", 0x5EED0003, tables, plan,
                    f"snowflake: fact {nf} rows x 4 pre-filtered dimensions, fused 4-probe chain with bloom filters")


@dataclass
class BatchWorkload:
    name: str
    jobs: List[Workload]  # independent two-table joins
    description: str


def c4(scale: float = 1.0, njobs: int = 256) -> BatchWorkload:
    """`njobs` independent INT32 equi-joins submitted together. Build rows are
    log-uniform in [1K, 2M] (times `scale`; unique shuffled keys + INT32
    payload); the probe side has 10x the build rows (FK uniform over the build
    keys, 5% non-matching keys, INT32 payload). Output {build payload, probe
    payload}. Sizes come from a seeded generator; every column has its own
    stream."""
    rng = random.Random(0x5EED0004)
    jobs = []

    def plan():
        p = Plan()
        b = p.new_scan_node(0, [(0, INT32), (1, INT32)])
        q = p.new_scan_node(1, [(0, INT32), (1, INT32)])
        p.root = p.new_join_node(True, b, q, 0, 0, [(1, INT32), (3, INT32)])
        return p

    for j in range(njobs):
        nb = max(int(round(math.exp(rng.uniform(math.log(1000), math.log(2_000_000))) * scale)), 64)
        npr = 10 * nb
        s = 1000 + 8 * j
        build = TableSpec(f"build{j}", nb, [
            GenColumn(type=INT32, kind=GEN_PERM_KEY, stream=s, domain=nb),
            GenColumn(type=INT32, kind=GEN_UNIFORM, stream=s + 1, lo=0, hi=1 << 31),
        ])
        probe = TableSpec(f"probe{j}", npr, [
            GenColumn(type=INT32, kind=GEN_FK_UNIFORM, stream=s + 2, domain=nb, miss_frac=0.05),
            GenColumn(type=INT32, kind=GEN_UNIFORM, stream=s + 3, lo=0, hi=1 << 31),
        ])
        jobs.append(Workload(f"c4.{j}", 0x5EED0004, [build, probe], plan, f"INT32 join {nb} x {npr}"))
    total_b = sum(w.tables[0].rows for w in jobs)
    total_p = sum(w.tables[1].rows for w in jobs)
    return BatchWorkload("c4", jobs, f"{njobs} batched INT32 equi-joins, build log-uniform [1K, 2M] x {scale:g} "
                                     f"({total_b} rows), probe 10x build ({total_p} rows), 5% misses")


@dataclass
class PoolPlan:
    plan: Plan
    tables: List[int]  # pool table of each plan input (scan base_table_id i -> pool table tables[i])
    shape: str         # "chain", "star" or "tree"
    edges: List[tuple] = None  # (source member, FK column, target member): source.FK = target.key

    def subplan(self, relset: int) -> "PoolPlan":
        """The plan of a connected set of members (bit k = member k, the plan's
        k-th scan): the same joins restricted to those members. Members join in
        insertion order, and every member's parent edge comes from an earlier
        member, so a connected set keeps each non-first member's parent."""
        keep = [k for k in range(len(self.tables)) if (relset >> k) & 1]
        idx = {k: i for i, k in enumerate(keep)}
        edges = [(idx[a], c, idx[b]) for a, c, b in self.edges if a in idx and b in idx]
        if len(edges) != len(keep) - 1:
            raise ValueError("relation set is not connected")
        return PoolPlan(_c5_plan(len(keep), edges), [self.tables[k] for k in keep], self.shape, edges)


def _c5_plan(nmembers: int, edges) -> Plan:
    """Left-deep contest plan over members 0..n-1 joined by `edges` in order;
    output: member 0's key and the last joined member's key (member 0's key
    alone for a single member)."""
    p = Plan()
    # columns each member scans: its key (joined or output) and its FK columns used
    scans = {k: [0] for k in range(nmembers)}
    for sp, c, _ in edges:
        scans[sp].append(c)
    for k in scans:
        scans[k] = sorted(set(scans[k]))
    nodes = [p.new_scan_node(k, [(c, INT32) for c in scans[k]]) for k in range(nmembers)]
    pos = {(0, c): i for i, c in enumerate(scans[0])}
    width = len(scans[0])
    cur = nodes[0]
    for e, (sp, c, tp) in enumerate(edges):
        right_cols = scans[tp]
        la = pos[(sp, c)]
        ra = right_cols.index(0)
        if e == len(edges) - 1:
            out = [(pos[(0, 0)], INT32), (width + ra, INT32)]
        else:
            out = [(i, INT32) for i in range(width + len(right_cols))]
        cur = p.new_join_node(False, cur, nodes[tp], la, ra, out)
        for i, c2 in enumerate(right_cols):
            pos[(tp, c2)] = width + i
        width += len(right_cols)
    p.root = cur
    return p


@dataclass
class PlanSet:
    name: str
    seed: int
    pool: List[TableSpec]  # shared by all plans
    plans: List[PoolPlan]
    description: str


C5_FKS = 11  # foreign-key columns per pool table (a star of 12 tables needs 11)


def c5(scale: float = 1.0, nplans: int = 2000, pool: int = 32, seed: int = 0x5EED0005) -> PlanSet:
    """`nplans` seeded join plans of 3-12 tables (chains, stars and mixed
    trees) over one shared pool of `pool` tables. Pool table i has
    log-uniform [10K, 40M] x `scale` rows; column 0 is its key: the table holds
    a log-uniform [1e-4, 1] fraction (its pre-filter selectivity) of a key
    domain, as unique INT32 keys. Columns 1..11 are INT32 foreign keys into
    the key domains of other pool tables, Zipf-distributed over key rank with
    s ~ U[0, 1.5]. Every join of a plan is foreign key = key of a table not yet
    in the plan, so a plan's result has at most as many rows as its first
    table. Plans output the first table's key and the last table's key."""
    rng = random.Random(seed)
    specs = []
    refs = []  # refs[i][c] = pool table that FK column c+1 of table i references
    lo, hi = math.log(10_000), math.log(40_000_000)
    domains = []
    for i in range(pool):
        n = max(int(round(math.exp(rng.uniform(lo, hi)) * scale)), 16)
        sel = math.exp(rng.uniform(math.log(1e-4), 0.0))
        d = min(max(int(math.ceil(n / sel)), n), 1 << 30)
        domains.append(d)
        specs.append((n, d))
    for i in range(pool):
        others = [j for j in range(pool) if j != i]
        refs.append([rng.choice(others) for _ in range(C5_FKS)])
    for i, (n, d) in enumerate(specs):
        cols = [GenColumn(type=INT32, kind=GEN_PERM_KEY, stream=5000 + 16 * i, domain=d)]
        for c in range(C5_FKS):
            cols.append(GenColumn(type=INT32, kind=GEN_FK_ZIPF, stream=5001 + 16 * i + c, domain=domains[refs[i][c]],
                                  zipf_s=rng.uniform(0.0, 1.5)))
        specs[i] = TableSpec(f"pool{i}", n, cols)
    pool_specs = specs

    def one_plan(prng):
        ntab = prng.randint(3, 12)
        shape = prng.choice(["chain", "star", "tree"])
        for _ in range(100):
            start = prng.randrange(pool)
            members = [start]
            edges = []  # (source position, fk column, target position)
            used = set()
            ok = True
            while len(members) < ntab:
                if shape == "chain":
                    srcs = [len(members) - 1]
                elif shape == "star":
                    srcs = [0]
                else:
                    srcs = list(range(len(members)))
                cands = [(sp, c) for sp in srcs for c in range(1, C5_FKS + 1)
                         if (sp, c) not in used and refs[members[sp]][c - 1] not in members]
                if not cands:
                    ok = False
                    break
                sp, c = prng.choice(cands)
                used.add((sp, c))
                members.append(refs[members[sp]][c - 1])
                edges.append((sp, c, len(members) - 1))
            if ok:
                return members, edges, shape
        raise RuntimeError("could not draw a plan from the pool")

    prng = random.Random(seed + 1)
    plans = []
    for _ in range(nplans):
        members, edges, shape = one_plan(prng)
        plans.append(PoolPlan(_c5_plan(len(members), edges), members, shape, edges))
    total = sum(t.rows for t in pool_specs)
    return PlanSet("c5", seed, pool_specs, plans,
                   f"{nplans} plans of 3-12 tables (chain/star/tree) over a pool of {pool} tables "
                   f"({total} rows, log-uniform [10K, 40M] x {scale:g}), Zipf FKs s~U[0,1.5], "
                   f"pre-filter selectivity log-uniform [1e-4, 1]")


WORKLOADS = {"c1": c1, "c2": c2, "c3": c3}
BATCH_WORKLOADS = {"c4": c4}
PLAN_SETS = {"c5": c5}
