"""The c5 plan set on the CPU: plan sizes and shapes, bounded results (every
join is foreign key = key of a new table), sub-plans, determinism."""
from collections import Counter

import pytest

from gpujoin import workloads as wl


def test_plan_set_shapes_and_sizes():
    ps = wl.c5(scale=0.001, nplans=300)
    sizes = Counter(len(p.tables) for p in ps.plans)
    assert min(sizes) == 3 and max(sizes) == 12
    assert set(p.shape for p in ps.plans) == {"chain", "star", "tree"}
    for p in ps.plans:
        assert len(set(p.tables)) == len(p.tables)  # a table at most once per plan
        assert len(p.edges) == len(p.tables) - 1
        for a, c, b in p.edges:
            assert a < b and 1 <= c <= wl.C5_FKS
        if p.shape == "chain":
            assert [a for a, _, _ in p.edges] == list(range(len(p.tables) - 1))
        if p.shape == "star":
            assert all(a == 0 for a, _, _ in p.edges)


def test_plan_set_is_deterministic():
    a = wl.c5(scale=0.001, nplans=20)
    b = wl.c5(scale=0.001, nplans=20)
    assert [p.tables for p in a.plans] == [p.tables for p in b.plans]
    assert [p.edges for p in a.plans] == [p.edges for p in b.plans]
    assert [t.rows for t in a.pool] == [t.rows for t in b.pool]


def test_subplans():
    ps = wl.c5(scale=0.001, nplans=20)
    p = next(p for p in ps.plans if p.shape == "chain" and len(p.tables) >= 4)
    sub = p.subplan(0b0110)
    assert sub.tables == p.tables[1:3] and len(sub.edges) == 1
    with pytest.raises(ValueError):
        p.subplan(0b0101)  # members 0 and 2 of a chain are not adjacent
    # the full set rebuilds the same plan
    full = p.subplan((1 << len(p.tables)) - 1)
    assert full.edges == p.edges
    assert [n.__dict__ for n in full.plan.nodes] == [n.__dict__ for n in p.plan.nodes]
