"""Packed hash tables on the one-probe kernel: a table of an INT32 key and one
INT32 attribute is read as one 8-byte word per slot (key | attr << 32). The
rows must be the oracle's with packing on and off; entries that do not fit
(a NULL attribute, the pair (-1, -1) whose packed word reads as an empty
slot) must make the kernel read the two-word table instead."""
import numpy as np
import pytest

import gpujoin as gj
import oracle
from gpujoin.pages import INT32, INT64
from helpers import check_table

pytestmark = pytest.mark.gpu


def _plan(bkeys, bvals, bvalid, pkeys, ppay, out):
    plan = gj.Plan()
    plan.new_input(gj.make_table([(INT32, bkeys, None), (INT32, bvals, bvalid)], len(bkeys)))
    plan.new_input(gj.make_table([(INT32, pkeys, None), (INT64, ppay, None)], len(pkeys)))
    b = plan.new_scan_node(0, [(0, INT32), (1, INT32)])
    q = plan.new_scan_node(1, [(0, INT32), (1, INT64)])
    plan.root = plan.new_join_node(True, b, q, 0, 0, out)
    return plan


def _run(ctx, plan, packed):
    types = [t for _, t in plan.nodes[plan.root].output_attrs]
    ctx.set_option("packed_tables", 1 if packed else 0)
    try:
        rows = check_table(ctx.execute(plan), types)
        stats = ctx.stats()
        tables = [ctx.upload(t) for t in plan.inputs]
        count = ctx.count_device(plan, tables)
    finally:
        ctx.set_option("packed_tables", 1)
    return rows, stats, count


def _check(ctx, plan, expect_packed):
    types = [t for _, t in plan.nodes[plan.root].output_attrs]
    exp = check_table(oracle.execute(plan), types)
    rows, stats, count = _run(ctx, plan, True)
    assert stats["probe_kernel"] == 1
    assert stats["probe_packed"] == (1 if expect_packed else 0), stats
    assert rows == exp
    assert count == len(exp)
    rows0, stats0, count0 = _run(ctx, plan, False)
    assert stats0["probe_packed"] == 0
    assert rows0 == exp and count0 == len(exp)


@pytest.mark.parametrize("nb,npr,dup", [(2, 64, False), (900, 5_000, False), (20_000, 200_000, True),
                                         (300_000, 1_000_000, False)])
def test_packed_table_rows_match_oracle(ctx, nb, npr, dup):
    """Unique and duplicate keys, misses, negative keys and attributes (the
    high half of a zero-extended INT32 is zero), small and partitioned builds."""
    rng = np.random.default_rng(nb)
    dom = max(nb // 3, 1) if dup else nb
    bk = (rng.integers(0, dom, nb) if dup else rng.permutation(nb)).astype(np.int64) - dom // 2
    bv = rng.integers(-(1 << 31), 1 << 31, nb)
    pk = rng.integers(0, int(dom * 1.25) + 1, npr) - dom // 2
    pp = rng.integers(-(1 << 62), 1 << 62, npr)
    plan = _plan(bk, bv, None, pk, pp, [(1, INT32), (3, INT64), (0, INT32)])
    _check(ctx, plan, True)


def test_packed_table_key_and_attr_all_ones_fall_back(ctx):
    """Key -1 with attribute -1 packs to all ones, the empty-slot word: the
    table is read unpacked. Key -1 alone, or attribute -1 alone, still packs."""
    n = 5_000
    keys = np.arange(n, dtype=np.int64) - 1          # includes -1
    vals = np.arange(n, dtype=np.int64)
    vals[0] = -1                                      # (-1, -1)
    probe = np.tile(keys, 4)  # the larger side: the planner builds on `keys`
    plan = _plan(keys, vals, None, probe, np.arange(len(probe)), [(1, INT32), (0, INT32)])
    _check(ctx, plan, False)
    vals[0] = 7                                       # key -1 with another attribute
    vals[5] = -1                                      # attribute -1 with key 4
    plan = _plan(keys, vals, None, probe, np.arange(len(probe)), [(1, INT32), (0, INT32)])
    _check(ctx, plan, True)


def test_packed_table_null_attribute_falls_back(ctx):
    """A NULL attribute has no 32-bit form: the table is read unpacked."""
    n = 20_000
    rng = np.random.default_rng(3)
    keys = rng.permutation(n)
    vals = rng.integers(0, 1000, n)
    valid = np.ones(n, bool)
    valid[1234] = False
    probe = rng.integers(0, n, 100_000)
    plan = _plan(keys, vals, valid, probe, np.arange(len(probe)), [(1, INT32), (3, INT64)])
    _check(ctx, plan, False)


def test_packed_table_consecutive_pairs(ctx):
    """With probe steps over the next two slots selected for two-word tables,
    a packed table still reads aligned pairs: same rows either way."""
    rng = np.random.default_rng(11)
    n = 50_000
    keys = rng.permutation(n)
    plan = _plan(keys, rng.integers(0, 1 << 31, n), None, rng.integers(0, n + 500, 300_000),
                 rng.integers(0, 1 << 40, 300_000), [(1, INT32), (3, INT64)])
    ctx.set_option("aligned_pairs", 0)
    try:
        _check(ctx, plan, True)
    finally:
        ctx.set_option("aligned_pairs", 1)
