"""Bloom filters left out where the samples say they would reject little.

The join tuner gives a table a bloom filter when the probe side is at least 8
times the build side. A probe that is estimated at 1M rows or more first
probes the driving relation's 64 sampled keys into the new table; when at
least half of the sampled non-NULL keys match, the pipeline probes without
the filter. The result never changes; the counters show which way it went.
Foreign keys into an unfiltered build side skip the filter (and the one-probe
kernel then reads the packed table); a probe side whose keys mostly miss keeps
it."""
import pytest

import gpujoin as gj
import oracle
from gpujoin.pages import INT32
from helpers import check_table

pytestmark = pytest.mark.gpu

SEED = 0x5EED00B1
NB, NP = 200_000, 2_000_000


def plan_for(miss_frac):
    build = [gj.GenColumn(type=INT32, kind=gj.GEN_PERM_KEY, stream=1, domain=NB),
             gj.GenColumn(type=INT32, kind=gj.GEN_UNIFORM, stream=2, lo=0, hi=1 << 31)]
    probe = [gj.GenColumn(type=INT32, kind=gj.GEN_FK_UNIFORM, stream=1, domain=NB, miss_frac=miss_frac),
             gj.GenColumn(type=INT32, kind=gj.GEN_UNIFORM, stream=4, lo=0, hi=1 << 31)]
    plan = gj.Plan()
    plan.new_input(gj.ColumnarTable(NB, [gj.host_generate(SEED, 0, NB, x) for x in build]))
    plan.new_input(gj.ColumnarTable(NP, [gj.host_generate(SEED, 0, NP, x) for x in probe]))
    b = plan.new_scan_node(0, [(0, INT32), (1, INT32)])
    q = plan.new_scan_node(1, [(0, INT32), (1, INT32)])
    plan.root = plan.new_join_node(True, b, q, 0, 0, [(1, INT32), (3, INT32)])
    return plan


@pytest.fixture(autouse=True)
def sampler(ctx):
    ctx.set_estimator(gj.EST_SAMPLE)
    yield


def test_foreign_keys_into_unfiltered_build_skip_the_filter(ctx):
    print("200K x 2M INT32 join, every probe key matching", flush=True)
    plan = plan_for(0.0)
    got = ctx.execute(plan)
    st = ctx.stats()
    assert got.num_rows == NP
    assert st["blooms_skipped"] == 1, st
    assert st["probe_packed"] == 1, st
    assert check_table(got, [INT32, INT32]) == check_table(oracle.execute(plan), [INT32, INT32])


def test_mostly_missing_keys_keep_the_filter(ctx):
    print("200K x 2M INT32 join, 99.9% of the probe keys missing", flush=True)
    plan = plan_for(0.999)
    got = ctx.execute(plan)
    st = ctx.stats()
    assert 0 < got.num_rows < NP // 100
    assert st["blooms_skipped"] == 0, st
    assert check_table(got, [INT32, INT32]) == check_table(oracle.execute(plan), [INT32, INT32])


def test_without_the_filter_rows_match_with_it(ctx):
    """The same join with bloom filters off altogether and on (the skip is
    only a cost decision)."""
    print("200K x 2M INT32 join, bloom filters off, then automatic", flush=True)
    plan = plan_for(0.05)
    ctx.set_option("bloom_bits", 0)
    try:
        off = check_table(ctx.execute(plan), [INT32, INT32])
        assert ctx.stats()["blooms_skipped"] == 0
    finally:
        ctx.set_option("bloom_bits", -1)
    on = check_table(ctx.execute(plan), [INT32, INT32])
    assert ctx.stats()["blooms_skipped"] == 1
    assert on == off
