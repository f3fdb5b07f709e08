"""Results far smaller than their inputs. An output page pool is sized for an
upper bound of the result (the largest input of the joined set); a result that
fills little of it is moved into a pool of exactly its pages, so results held
together (execute_many over many plans) do not keep the unused slots. Rows stay
identical to the oracle's, through host pages, device tables and
execute_many, and the moved pool's directory still orders every column.

These joins are planned with the max rule over exact sizes: the sampler would
see that few probe keys match and size the pool close to the result."""
import pytest

import gpujoin as gj
import oracle
from gpujoin.pages import FP64, INT64
from helpers import check_table

pytestmark = pytest.mark.gpu

SEED = 0x5EED00C7
NB, NP = 100_000, 8_000_000


def gens():
    build = [gj.GenColumn(type=INT64, kind=gj.GEN_UNIQUE_KEY, stream=1, null_stream=11, null_frac=0.05, domain=NB),
             gj.GenColumn(type=FP64, kind=gj.GEN_UNIT_F64, stream=2, null_stream=12, null_frac=0.05)]
    # 0.05% of the probe keys hit the build side: a ~3,500-row result from an 8M-row scan
    # (the foreign keys follow the build keys' stream)
    probe = [gj.GenColumn(type=INT64, kind=gj.GEN_FK_ZIPF, stream=1, null_stream=13, null_frac=0.05, domain=NB,
                          zipf_s=0.5, miss_frac=0.9995),
             gj.GenColumn(type=INT64, kind=gj.GEN_UNIFORM, stream=4, null_stream=14, null_frac=0.05, lo=0,
                          hi=1 << 40)]
    return build, probe


def plan_for(build=None, probe=None):
    """The c2 shape; without inputs for device tables."""
    plan = gj.Plan()
    if build is not None:
        plan.new_input(build)
        plan.new_input(probe)
    b = plan.new_scan_node(0, [(0, INT64), (1, FP64)])
    q = plan.new_scan_node(1, [(0, INT64), (1, INT64)])
    plan.root = plan.new_join_node(True, b, q, 0, 0, [(1, FP64), (3, INT64), (2, INT64)])
    return plan


TYPES = [FP64, INT64, INT64]


@pytest.fixture(autouse=True)
def max_rule(ctx):
    ctx.set_estimator(gj.EST_MAX)
    yield
    ctx.set_estimator(gj.EST_SAMPLE)


def host_plan():
    g, p = gens()
    build = gj.ColumnarTable(NB, [gj.host_generate(SEED, 0, NB, x) for x in g])
    probe = gj.ColumnarTable(NP, [gj.host_generate(SEED, 0, NP, x) for x in p])
    return plan_for(build, probe)


def test_selective_join_moves_its_pool(ctx):
    print("selective join 100K x 8M through host pages", flush=True)
    plan = host_plan()
    exp = check_table(oracle.execute(plan), TYPES)
    got = ctx.execute(plan)
    st = ctx.stats()
    assert 0 < got.num_rows < 10_000
    assert st["pools_compacted"] == 1, st
    assert check_table(got, TYPES) == exp


def test_moved_pools_in_execute_many(ctx):
    print("selective join 100K x 8M, device tables, execute_many", flush=True)
    exp = check_table(oracle.execute(host_plan()), TYPES)
    g, p = gens()
    tabs = [ctx.generate(SEED, NB, g), ctx.generate(SEED, NP, p)]
    plans = [plan_for() for _ in range(6)]
    one = ctx.execute_device(plans[0], tabs)
    assert ctx.stats()["pools_compacted"] == 1
    assert check_table(one.download(), TYPES) == exp
    for r in ctx.execute_many(plans, [tabs] * len(plans), workers=3):
        assert check_table(r.download(), TYPES) == exp


def test_large_result_keeps_its_pool(ctx):
    """A result that fills most of its pool is left where it is."""
    print("join 100K x 400K with every probe key matching", flush=True)
    g, _ = gens()
    p = [gj.GenColumn(type=INT64, kind=gj.GEN_FK_ZIPF, stream=1, null_stream=13, null_frac=0.05, domain=NB,
                      zipf_s=0.5, miss_frac=0.0),
         gj.GenColumn(type=INT64, kind=gj.GEN_UNIFORM, stream=4, lo=0, hi=1 << 40)]
    build = gj.ColumnarTable(NB, [gj.host_generate(SEED, 0, NB, x) for x in g])
    probe = gj.ColumnarTable(400_000, [gj.host_generate(SEED, 0, 400_000, x) for x in p])
    plan = plan_for(build, probe)
    got = ctx.execute(plan)
    assert ctx.stats()["pools_compacted"] == 0
    assert check_table(got, TYPES) == check_table(oracle.execute(plan), TYPES)
