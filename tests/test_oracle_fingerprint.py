"""Pins the oracle's full-size checks (CPU only).

oracle.join_fingerprint (a threaded, partitioned open-addressing hash join)
must agree with the oracle's plan evaluator (a string-keyed multimap join,
tests/test_oracle.py pins it to the reference's ground truths and brute
force) followed by oracle.table_fingerprint, on random two-table joins with
NULL keys, duplicate keys and misses. table_fingerprint must not depend on
row order or page packing and must change when one value changes."""
import random

import numpy as np
import pytest

import gpujoin as gj
import oracle
from gpujoin.pages import FP64, INT32, INT64


def _host(table):
    m = gj._Marshal(gj.Plan(), [table])
    return m, m.tables[0]


def _table(rng, n, key_type, domain, null_frac, dup):
    keys = [rng.randrange(domain if dup else 10 * domain) for _ in range(n)]
    kvalid = [rng.random() >= null_frac for _ in range(n)]
    pay32 = [rng.randrange(-(1 << 31), 1 << 31) for _ in range(n)]
    pay64 = [rng.random() * 1e6 - 5e5 for _ in range(n)]
    pvalid = [rng.random() >= null_frac for _ in range(n)]
    return gj.make_table([(key_type, keys, kvalid), (INT32, pay32, pvalid), (FP64, pay64, None)], n)


@pytest.mark.parametrize("seed", range(12))
def test_join_fingerprint_matches_evaluator(seed):
    rng = random.Random(seed)
    kt = [INT32, INT64][seed % 2]
    nb, npr = rng.choice([0, 1, 50, 700]), rng.choice([1, 90, 2500])
    build = _table(rng, nb, kt, 40, 0.1, dup=seed % 3 == 0)
    probe = _table(rng, npr, kt, 40, 0.2, dup=True)
    outs = [(1, 2), (0, 1), (1, 0), (0, 2)][: 1 + seed % 4]
    plan = gj.Plan()
    plan.new_input(build)
    plan.new_input(probe)
    a = plan.new_scan_node(0, [(0, kt), (1, INT32), (2, FP64)])
    b = plan.new_scan_node(1, [(0, kt), (1, INT32), (2, FP64)])
    types = [kt, INT32, FP64]
    plan.root = plan.new_join_node(True, a, b, 0, 0, [(s * 3 + c, types[c]) for s, c in outs])
    res = oracle.execute(plan)
    mr, hr = _host(res)
    mb, hb = _host(build)
    mp, hp = _host(probe)
    assert oracle.join_fingerprint(hb, 0, hp, 0, outs, threads=1 + seed % 5) == \
        oracle.table_fingerprint(hr, threads=3)


def test_table_fingerprint_order_and_packing_independent():
    rng = np.random.default_rng(5)
    n = 20_000
    # NULLs sit in the INT32 column: on 8-byte pages with NULLs the contest
    # writer's bitmap may overlap the last value, so there a different page
    # packing can change what a value reads back as
    keys = rng.integers(0, 1 << 40, n)
    vals = rng.integers(-(1 << 31), 1 << 31, n)
    valid = rng.random(n) > 0.3

    def tab(k, v, m):
        return gj.make_table([(INT64, k, None), (INT32, v, m)], len(k))

    t1 = tab(keys, vals, valid)
    perm = rng.permutation(n)
    t2 = tab(keys[perm], vals[perm], valid[perm])
    # the same rows split over two tables sum to the whole
    h = n // 3
    t3 = tab(keys[:h], vals[:h], valid[:h])
    t4 = tab(keys[h:], vals[h:], valid[h:])
    f1 = oracle.table_fingerprint(_host(t1)[1], 4)
    f2 = oracle.table_fingerprint(_host(t2)[1], 7)
    f3 = oracle.table_fingerprint(_host(t3)[1], 2)
    f4 = oracle.table_fingerprint(_host(t4)[1], 1)
    assert f1 == f2
    assert (f3[0] + f4[0], (f3[1] + f4[1]) % (1 << 64)) == f1
    # one changed value, one NULL flag flipped, two columns swapped: all differ
    k2 = keys.copy()
    k2[123] ^= 1
    v2 = valid.copy()
    v2[77] = not v2[77]
    for t in (tab(k2, vals, valid), tab(keys, vals, v2),
              gj.make_table([(INT32, vals, valid), (INT64, keys, None)], n)):
        assert oracle.table_fingerprint(_host(t)[1], 4) != f1
