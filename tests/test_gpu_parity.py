"""GPU parity: the HIP path through the C ABI against the oracle and the
reference's golden vectors. Bit-exact row multisets (NULLs and VARCHAR bytes
included); every result page is validated against the page format."""
import random
import sys

import numpy as np
import pytest

import gpujoin as gj
import oracle
from gpujoin.pages import FP64, INT32, INT64, VARCHAR
from helpers import TYPES, canon_rows, case_plan, check_table, load_cases, random_plan

pytestmark = pytest.mark.gpu
CASES = load_cases()


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_golden_cases(ctx, case):
    say(f"golden case {case['name']}")
    res = ctx.execute(case_plan(case))
    rows = check_table(res, [TYPES[t] for t in case["expect_types"]])
    assert res.num_rows == len(case["expect_rows"])
    assert rows == canon_rows(case["expect_rows"])


@pytest.mark.parametrize("seed", range(60))
def test_random_plans_vs_oracle(ctx, seed):
    rng = random.Random(1000 + seed)
    plan = random_plan(rng, ntables=rng.randint(1, 6), max_rows=rng.choice([5, 50, 400, 3000]),
                       key_type=rng.choice([INT32, INT64]), domain=rng.choice([4, 30, 1000]),
                       long_strings=seed % 7 == 0)
    say(f"random plan seed={seed} tables={len(plan.inputs)}")
    types = [t for _, t in plan.nodes[plan.root].output_attrs]
    got = check_table(ctx.execute(plan), types)
    exp = check_table(oracle.execute(plan), types)
    assert got == exp


def slots_of(table):
    """Fixed-width result columns as uint64 slot arrays (NULL = all ones)."""
    out = []
    for c in table.columns:
        v, m = gj.decode_column(c.pages, c.type)
        if c.type == INT32:
            bits = v.astype(np.uint32).astype(np.uint64)
        else:
            bits = v.view(np.uint64)
        out.append(np.where(m, bits, np.uint64(~np.uint64(0))))
    return out


def sorted_rows(cols):
    order = np.lexsort(cols[::-1])
    return [c[order] for c in cols]


@pytest.mark.parametrize("null_frac", [0.0, 0.05, 0.5, 0.97])
@pytest.mark.parametrize("t", [INT32, INT64, FP64])
def test_decode_kernel_vs_oracle(ctx, t, null_frac):
    import torch

    n = 123_457
    rng = np.random.default_rng(int(null_frac * 100) + t)
    valid = rng.random(n) >= null_frac
    vals = rng.random(n) if t == FP64 else rng.integers(0, 1 << 31, n)
    col = gj.encode_column(t, vals, valid)
    say(f"decode t={t} null_frac={null_frac} pages={len(col.pages)}")
    dt = ctx.upload(gj.ColumnarTable(n, [col]))
    out = torch.empty(n, dtype=torch.int64, device="cuda")
    dt.decode(0, out.data_ptr())
    got = out.cpu().numpy().view(np.uint64)
    assert np.array_equal(got, oracle.decode_fixed(col, n))


def c1_plan(nb, np_, seed=0x5EED0001):
    """Config-1 shape: INT32 unique keys + INT32 payload joined with an INT32 FK
    + INT64 payload probe side; output {build payload, probe payload}."""
    g = [gj.GenColumn(type=INT32, kind=gj.GEN_PERM_KEY, stream=1, domain=nb),
         gj.GenColumn(type=INT32, kind=gj.GEN_UNIFORM, stream=2, lo=0, hi=1 << 31)]
    p = [gj.GenColumn(type=INT32, kind=gj.GEN_FK_UNIFORM, stream=1, domain=nb),
         gj.GenColumn(type=INT64, kind=gj.GEN_UNIFORM, stream=4, lo=0, hi=1 << 63)]
    return g, p, seed


def test_c1_shape_vs_oracle(ctx):
    g, p, seed = c1_plan(20_000, 200_000)
    build = gj.ColumnarTable(20_000, [gj.host_generate(seed, 0, 20_000, x) for x in g])
    probe = gj.ColumnarTable(200_000, [gj.host_generate(seed, 0, 200_000, x) for x in p])
    plan = gj.Plan()
    plan.new_input(build)
    plan.new_input(probe)
    b = plan.new_scan_node(0, [(0, INT32), (1, INT32)])
    q = plan.new_scan_node(1, [(0, INT32), (1, INT64)])
    plan.root = plan.new_join_node(True, b, q, 0, 0, [(1, INT32), (3, INT64)])
    say("c1 shape 20K x 200K")
    got = ctx.execute(plan)
    exp = oracle.execute(plan)
    assert got.num_rows == exp.num_rows == 200_000
    for c in got.columns:
        gj.pages.validate_column(c.pages, c.type)
    for a, b_ in zip(sorted_rows(slots_of(got)), sorted_rows(slots_of(exp))):
        assert np.array_equal(a, b_)


def test_device_generator_matches_host(ctx):
    """Pages generated in HBM decode to the same values as the host recipe."""
    import torch

    n = 300_001
    gens = [gj.GenColumn(type=INT64, kind=gj.GEN_FK_ZIPF, stream=7, null_stream=8, null_frac=0.05,
                         domain=100_000, zipf_s=1.0, miss_frac=0.1),
            gj.GenColumn(type=FP64, kind=gj.GEN_UNIT_F64, stream=9, null_stream=10, null_frac=0.05)]
    dt = ctx.generate(0x5EED0002, n, gens)
    for ci, g in enumerate(gens):
        host = gj.host_generate(0x5EED0002, 0, n, g)
        assert dt.num_pages(ci) == len(host.pages)
        out = torch.empty(n, dtype=torch.int64, device="cuda")
        dt.decode(ci, out.data_ptr())
        assert np.array_equal(out.cpu().numpy().view(np.uint64), oracle.decode_fixed(host, n))
