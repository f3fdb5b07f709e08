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
    ntables = rng.randint(1, 6)
    max_rows = rng.choice([5, 50, 400, 3000])
    key_type = rng.choice([INT32, INT64])
    # keep the result bounded: with d key values per join column each join
    # multiplies the row count by about rows/d
    domain = max(rng.choice([4, 30, 1000]), max_rows if ntables >= 3 else max_rows // 8, 4)
    plan = random_plan(rng, ntables=ntables, max_rows=max_rows, key_type=key_type, domain=domain,
                       long_strings=seed % 7 == 0)
    say(f"random plan seed={seed} tables={len(plan.inputs)}")
    types = [t for _, t in plan.nodes[plan.root].output_attrs]
    got = check_table(ctx.execute(plan), types)
    exp = check_table(oracle.execute(plan), types)
    assert got == exp


def test_result_beyond_device_memory_is_refused(ctx):
    """A five-table plan with a 4-value join key domain whose result has
    11,109,749,415,203 rows (about 10^14 bytes of pages; counted on the CPU by
    tree message passing). The pipeline that stages or writes it stops once
    its reservations pass what device memory holds, and execute raises a clear
    error within seconds instead of enumerating the whole result first."""
    import time

    rng = random.Random(1002)
    plan = random_plan(rng, ntables=rng.randint(1, 6), max_rows=rng.choice([5, 50, 400, 3000]),
                       key_type=rng.choice([INT32, INT64]), domain=rng.choice([4, 30, 1000]))
    assert [t.num_rows for t in plan.inputs] == [2545, 1462, 2832, 1495, 677]
    say("result beyond device memory")
    t0 = time.perf_counter()
    with pytest.raises(RuntimeError, match="too large|out of memory"):
        ctx.execute(plan)
    assert time.perf_counter() - t0 < 60
    # the context stays usable
    small = random_plan(random.Random(3), ntables=2, max_rows=50, domain=100)
    types = [t for _, t in small.nodes[small.root].output_attrs]
    assert check_table(ctx.execute(small), types) == check_table(oracle.execute(small), types)


@pytest.mark.parametrize("seed", range(16))
def test_block_prefilter_with_bloom_filters_on_every_join(ctx, seed):
    """Every join gets a bloom filter, so pipelines over pages with eight
    tuple slots test their early filters on eight blocks at a time before the
    per-block work. Sparse key domains leave most blocks without a match;
    odd seeds have no NULLs, so their fixed-width columns are read as dense
    columns (no block index). Rows equal the oracle's and those with the block
    pre-filter off."""
    rng = random.Random(7000 + seed)
    ntables = rng.randint(3, 6)
    max_rows = rng.choice([3000, 20000, 60000])
    key_type = rng.choice([INT32, INT64])
    domain = rng.choice([max_rows, 4 * max_rows, 40 * max_rows])
    plan = random_plan(rng, ntables=ntables, max_rows=max_rows, key_type=key_type, domain=domain,
                       null_frac=0.0 if seed % 2 else 0.15)
    types = [t for _, t in plan.nodes[plan.root].output_attrs]
    say(f"block pre-filter seed={seed} tables={ntables} rows<={max_rows} domain={domain}")
    ctx.set_option("bloom_bits", 16)
    try:
        got = check_table(ctx.execute(plan), types)
        ctx.set_option("block_prefilter", 0)
        off = check_table(ctx.execute(plan), types)
    finally:
        ctx.set_option("block_prefilter", 1)
        ctx.set_option("bloom_bits", -1)
    assert got == off
    assert got == check_table(oracle.execute(plan), types)


@pytest.fixture
def grouped_ctx(ctx):
    """The context with every hash table built grouped by key."""
    ctx.set_option("force_grouped", 1)
    yield ctx
    ctx.set_option("force_grouped", 0)


@pytest.mark.parametrize("seed", range(0, 60, 3))
def test_random_plans_grouped_tables_vs_oracle(grouped_ctx, seed):
    """The random plans again, every hash table grouped by key (one slot per
    distinct key, the key's tuples in a payload array): residual equalities,
    probe chains, counts and every output type."""
    rng = random.Random(1000 + seed)
    ntables = rng.randint(1, 6)
    max_rows = rng.choice([5, 50, 400, 3000])
    key_type = rng.choice([INT32, INT64])
    domain = max(rng.choice([4, 30, 1000]), max_rows if ntables >= 3 else max_rows // 8, 4)
    plan = random_plan(rng, ntables=ntables, max_rows=max_rows, key_type=key_type, domain=domain,
                       long_strings=seed % 7 == 0)
    say(f"grouped random plan seed={seed} tables={len(plan.inputs)}")
    types = [t for _, t in plan.nodes[plan.root].output_attrs]
    got = check_table(grouped_ctx.execute(plan), types)
    exp = check_table(oracle.execute(plan), types)
    assert got == exp
    tables = [grouped_ctx.upload(t) for t in plan.inputs]
    assert grouped_ctx.count_device(plan, tables) == len(exp)


def _hot_key_plan(n_build, n_probe, hot, seed):
    """Build side of `n_build` rows where key 7 takes a `hot` share of the
    rows; probe side with a uniform key over the same domain."""
    rng = np.random.default_rng(seed)
    domain = 5000
    bk = rng.integers(0, domain, n_build).astype(np.int64)
    bk[rng.random(n_build) < hot] = 7
    bv = rng.integers(0, 1 << 30, n_build)
    pk = rng.integers(0, domain, n_probe).astype(np.int64)
    pv = rng.integers(0, 1 << 30, n_probe)
    # the planner hashes the smaller input: callers make it the build side
    return _join_plan([(INT64, bk, None), (INT32, bv, None)], [(INT64, pk, None), (INT32, pv, None)],
                      0, 0, [(1, INT32), (3, INT32)])


@pytest.mark.parametrize("n_build,n_probe,hot", [(40_000, 300_000, 0.5), (200_000, 400_000, 0.3)])
def test_heavily_repeated_build_key_takes_grouped_table(ctx, n_build, n_probe, hot):
    """A key repeated tens of thousands of times on the (smaller, hashed)
    build side, below and above the partitioned build's size threshold: the
    plain insert stops at the walk cap and the table is rebuilt grouped; same
    rows as the oracle and an exact count (the count target adds whole groups
    at once)."""
    plan = _hot_key_plan(n_build, n_probe, hot, 5)
    say(f"hot build key {n_build} x {n_probe}, share {hot}")
    got = ctx.execute(plan)
    st = ctx.stats()
    exp = oracle.execute(plan)
    assert got.num_rows == exp.num_rows
    assert st["retries"] >= 1, st  # the plain build was abandoned
    for a, b_ in zip(sorted_rows(slots_of(got)), sorted_rows(slots_of(exp))):
        assert np.array_equal(a, b_)
    tables = [ctx.upload(t) for t in plan.inputs]
    assert ctx.count_device(plan, tables) == exp.num_rows


def test_grouped_key_column_remembered_across_queries(ctx):
    """A build on a base table's heavily repeated key column that had to be
    rebuilt grouped is remembered per (table, column): the next query over the
    same device tables builds it grouped at once (fewer retries), same rows."""
    # about 2.4M result rows, compared as sorted numpy slot columns
    plan = _hot_key_plan(100_000, 120_000, 0.3, 6)
    tables = [ctx.upload(t) for t in plan.inputs]
    say("hot build key, two queries over the same device tables")
    first = sorted_rows(slots_of(ctx.execute_device(plan, tables).download()))
    r1 = ctx.stats()["retries_build"]
    second = sorted_rows(slots_of(ctx.execute_device(plan, tables).download()))
    r2 = ctx.stats()["retries_build"]
    assert r1 >= 1 and r2 < r1, (r1, r2)
    exp = sorted_rows(slots_of(oracle.execute(plan)))
    for a, b_, c in zip(first, second, exp):
        assert np.array_equal(a, b_)
        assert np.array_equal(a, c)


def test_grouped_and_plain_tables_agree_on_a_star_plan(ctx):
    """A c5 star plan (foreign keys with Zipf skew up to s = 1.5) run with
    grouped tables and with plain ones: identical rows."""
    from gpujoin import workloads as wl
    from gpujoin import estimation as est
    ps = wl.c5(scale=0.002, nplans=40)
    pp = next(p for p in ps.plans if p.shape == "star" and len(p.tables) >= 5)
    tabs = [ctx.generate(ps.seed, ps.pool[t].rows, ps.pool[t].columns) for t in pp.tables]
    types = [t for _, t in pp.plan.nodes[pp.plan.root].output_attrs]
    say(f"star plan over {len(pp.tables)} tables, grouped vs plain")
    plain = check_table(ctx.execute_device(pp.plan, tabs).download(), types)
    ctx.set_option("force_grouped", 1)
    try:
        grouped = check_table(ctx.execute_device(pp.plan, tabs).download(), types)
    finally:
        ctx.set_option("force_grouped", 0)
    assert plain == grouped


@pytest.mark.parametrize("insert_max", [8 << 20, 0, 1 << 30], ids=["default", "no-insert", "insert-all"])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_scan_build_matches_staged_build(ctx, seed, insert_max):
    """Hash tables on a plain scan built straight from the pages (inserted by
    one kernel up to `insert_max` bytes of table, partitioned above) against
    the staged path and the oracle: NULL keys, a second join column with NULLs
    in the build tuple, a VARCHAR payload, and a three-table chain."""
    rng = np.random.default_rng(seed)
    nb, npr, nd = 120_000, 300_000, 90_000
    bkey = rng.permutation(np.arange(nb, dtype=np.int64) * 3 + 1)
    bv = np.array([f"s{v % 977}" for v in range(nb)], dtype=object)
    bfk = rng.integers(0, nd, nb).astype(np.int32)
    pkey = np.where(rng.random(npr) < 0.8, bkey[rng.integers(0, nb, npr)], -7)
    dkey = rng.permutation(np.arange(nd, dtype=np.int32))
    plan = gj.Plan()
    plan.new_input(gj.make_table([(INT64, bkey, rng.random(nb) > 0.05), (VARCHAR, bv, rng.random(nb) > 0.1),
                                  (INT32, bfk, rng.random(nb) > 0.05)], nb))
    plan.new_input(gj.make_table([(INT64, pkey, rng.random(npr) > 0.05), (FP64, rng.random(npr), None)], npr))
    plan.new_input(gj.make_table([(INT32, dkey, rng.random(nd) > 0.02), (INT32, rng.integers(0, 9, nd), None)], nd))
    b = plan.new_scan_node(0, [(0, INT64), (1, VARCHAR), (2, INT32)])
    q = plan.new_scan_node(1, [(0, INT64), (1, FP64)])
    d = plan.new_scan_node(2, [(0, INT32), (1, INT32)])
    j = plan.new_join_node(True, b, q, 0, 0, [(1, VARCHAR), (2, INT32), (4, FP64)])
    plan.root = plan.new_join_node(True, j, d, 1, 0, [(0, VARCHAR), (2, FP64), (4, INT32)])
    types = [VARCHAR, FP64, INT32]
    say(f"scan build vs staged build, seed {seed}, insert up to {insert_max} bytes")
    ctx.set_option("scan_insert_max_bytes", insert_max)
    try:
        got = check_table(ctx.execute(plan), types)
        ctx.set_option("scan_build", 0)
        staged = check_table(ctx.execute(plan), types)
    finally:
        ctx.set_option("scan_build", 1)
        ctx.set_option("scan_insert_max_bytes", 8 << 20)
    exp = check_table(oracle.execute(plan), types)
    assert got == staged == exp


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
    n = 123_457
    rng = np.random.default_rng(int(null_frac * 100) + t)
    valid = rng.random(n) >= null_frac
    vals = rng.random(n) if t == FP64 else rng.integers(0, 1 << 31, n)
    col = gj.encode_column(t, vals, valid)
    say(f"decode t={t} null_frac={null_frac} pages={len(col.pages)}")
    dt = ctx.upload(gj.ColumnarTable(n, [col]))
    assert np.array_equal(dt.decode(0), oracle.decode_fixed(col, n))


@pytest.mark.parametrize("n", [0, 1, 63, 64, 65, 1023, 1024, 1025, 4097])
def test_decode_ragged_sizes(ctx, n):
    """Morsel and wave boundaries: sizes around 64 rows and 1024-row morsels."""
    rng = np.random.default_rng(n)
    valid = rng.random(n) > 0.3
    col = gj.encode_column(INT32, rng.integers(0, 1000, n), valid)
    dt = ctx.upload(gj.ColumnarTable(n, [col]))
    assert np.array_equal(dt.decode(0), oracle.decode_fixed(col, n))


@pytest.mark.parametrize("n", [1, 64, 1006, 1007, 1008, 1983, 1984, 1985, 4097, 70_001])
@pytest.mark.parametrize("t", [INT32, INT64])
def test_decode_dense_columns(ctx, t, n):
    """NULL-free columns with full pages are read without the block index
    (row r is value r % d of page r / d): sizes around the 1984-row INT32 and
    1007-row INT64 pages, and blocks that cross a page boundary."""
    rng = np.random.default_rng(n + 7 * t)
    col = gj.encode_column(t, rng.integers(0, 1 << 31, n), np.ones(n, dtype=bool))
    dt = ctx.upload(gj.ColumnarTable(n, [col]))
    assert np.array_equal(dt.decode(0), oracle.decode_fixed(col, n))


def test_decode_uneven_pages_not_taken_as_dense(ctx):
    """NULL-free pages of unequal fill (a 100-row page before full ones) must
    be read through the block index, not as fixed rows per page."""
    rng = np.random.default_rng(3)
    a = gj.encode_column(INT32, rng.integers(0, 1 << 31, 100), np.ones(100, dtype=bool))
    b = gj.encode_column(INT32, rng.integers(0, 1 << 31, 5000), np.ones(5000, dtype=bool))
    col = gj.Column(INT32, a.pages + b.pages)
    dt = ctx.upload(gj.ColumnarTable(5100, [col]))
    assert np.array_equal(dt.decode(0), oracle.decode_fixed(col, 5100))


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
    n = 300_001
    gens = [gj.GenColumn(type=INT64, kind=gj.GEN_FK_ZIPF, stream=7, null_stream=8, null_frac=0.05,
                         domain=100_000, zipf_s=1.0, miss_frac=0.1),
            gj.GenColumn(type=FP64, kind=gj.GEN_UNIT_F64, stream=9, null_stream=10, null_frac=0.05)]
    dt = ctx.generate(0x5EED0002, n, gens)
    for ci, g in enumerate(gens):
        host = gj.host_generate(0x5EED0002, 0, n, g)
        assert dt.num_pages(ci) == len(host.pages)
        assert np.array_equal(dt.decode(ci), oracle.decode_fixed(host, n))


def _join_plan(build_cols, probe_cols, bkey, pkey, out):
    plan = gj.Plan()
    nb = len(build_cols[0][1])
    npr = len(probe_cols[0][1])
    plan.new_input(gj.make_table(build_cols, nb))
    plan.new_input(gj.make_table(probe_cols, npr))
    b = plan.new_scan_node(0, [(i, c[0]) for i, c in enumerate(build_cols)])
    q = plan.new_scan_node(1, [(i, c[0]) for i, c in enumerate(probe_cols)])
    plan.root = plan.new_join_node(True, b, q, bkey, pkey, out)
    return plan


def _compare_large(ctx, plan):
    got = ctx.execute(plan)
    exp = oracle.execute(plan)
    assert got.num_rows == exp.num_rows
    for c in got.columns:
        assert gj.pages.validate_column(c.pages, c.type) == got.num_rows
    for a, b_ in zip(sorted_rows(slots_of(got)), sorted_rows(slots_of(exp))):
        assert np.array_equal(a, b_)
    return ctx.stats()


def test_partitioned_build_unique_int64(ctx):
    """>= 64K build tuples take the radix-partitioned LDS build."""
    rng = np.random.default_rng(11)
    nb, npr = 300_000, 1_000_000
    keys = rng.permutation(np.arange(nb, dtype=np.int64) * 7919 + 3)
    bval = rng.random(nb)
    pk = np.where(rng.random(npr) < 0.9, keys[rng.integers(0, nb, npr)], -keys[rng.integers(0, nb, npr)] - 10)
    plan = _join_plan([(INT64, keys, rng.random(nb) > 0.05), (FP64, bval, rng.random(nb) > 0.05)],
                      [(INT64, pk, rng.random(npr) > 0.05), (FP64, rng.random(npr), rng.random(npr) > 0.05)],
                      0, 0, [(1, FP64), (3, FP64), (2, INT64)])
    say("partitioned build 300K x 1M")
    s = _compare_large(ctx, plan)
    assert s["retries"] == 0, s


def test_partitioned_build_large_regions(ctx):
    """Few regions of 8192 slots (a 128 KiB LDS table per workgroup), as the
    largest tables use; same rows as the oracle."""
    rng = np.random.default_rng(21)
    nb, npr = 300_000, 900_000
    keys = rng.permutation(np.arange(nb, dtype=np.int64) * 31 + 7)
    pk = np.where(rng.random(npr) < 0.8, keys[rng.integers(0, nb, npr)], -keys[rng.integers(0, nb, npr)] - 3)
    plan = _join_plan([(INT64, keys, rng.random(nb) > 0.05), (INT32, rng.integers(0, 1 << 30, nb), None)],
                      [(INT64, pk, rng.random(npr) > 0.05)], 0, 0, [(1, INT32), (2, INT64)])
    ctx.set_option("build_regions_log", 7)
    try:
        say("partitioned build with 2^7 regions")
        s = _compare_large(ctx, plan)
    finally:
        ctx.set_option("build_regions_log", 14)
    assert s["retries"] == 0, s


def test_partitioned_build_wide_entries(ctx):
    """Eight-word entries shrink the LDS region; results stay exact."""
    rng = np.random.default_rng(12)
    nb, npr = 150_000, 400_000
    keys = rng.permutation(nb).astype(np.int32)
    cols = [(INT32, keys, None)] + [(INT64, rng.integers(0, 1 << 40, nb), rng.random(nb) > 0.1) for _ in range(7)]
    pk = rng.integers(0, nb + nb // 10, npr).astype(np.int32)
    out = [(i, INT64) for i in range(1, 8)] + [(8, INT32)]
    plan = _join_plan(cols, [(INT32, pk, rng.random(npr) > 0.02)], 0, 0, out)
    say("wide entries 150K x 400K")
    _compare_large(ctx, plan)


def test_heavy_duplicate_build_falls_back(ctx):
    """A key repeated 20K times overflows its region: the build falls back to
    global inserts and the join still expands every duplicate."""
    rng = np.random.default_rng(13)
    nb = 100_000
    keys = rng.integers(0, 50_000, nb).astype(np.int32)
    keys[:20_000] = 7
    probe = rng.integers(60_000, 1 << 30, 300_000).astype(np.int32)  # misses
    probe[[5, 77, 1234]] = 7
    probe[10:400] = rng.integers(0, 50_000, 390)
    plan = _join_plan([(INT32, keys, None), (INT32, np.arange(nb, dtype=np.int32), None)],
                      [(INT32, probe, None)], 0, 0, [(1, INT32), (2, INT32)])
    say("heavy duplicates")
    s = _compare_large(ctx, plan)
    assert s["retries"] >= 1


def _lean_shape_plan(seed):
    """Two-table join whose output mixes the probe key, build and probe payloads
    of every fixed-width type; duplicates, NULLs and misses as the seed says."""
    rng = np.random.default_rng(700 + seed)
    nb = int(rng.choice([1, 900, 20_000, 90_000]))
    npr = int(rng.choice([64, 5_000, 70_000, 200_000]))
    kt = [INT32, INT64][seed % 2]
    dup = seed % 3 == 0
    dom = max(nb // (4 if dup else 1), 1)
    bkeys = rng.integers(0, dom, nb) if dup else rng.permutation(nb)
    bkeys = bkeys.astype(np.int32 if kt == INT32 else np.int64) * (1 if kt == INT32 else 1_000_003)
    pk = rng.integers(0, int(dom * 1.2) + 1, npr).astype(bkeys.dtype) * (1 if kt == INT32 else 1_000_003)
    nullf = [0.0, 0.05, 0.6][seed % 3]

    def col(t, n):
        v = rng.random(n) * 1e6 if t == FP64 else rng.integers(-(1 << 30), 1 << 30, n)
        return (t, v, rng.random(n) >= nullf)

    bpay = [col(t, nb) for t in rng.choice([INT32, INT64, FP64], int(rng.integers(1, 4)))]
    ppay = [col(t, npr) for t in rng.choice([INT32, INT64, FP64], int(rng.integers(1, 3)))]
    build = [(kt, bkeys, rng.random(nb) >= nullf / 2)] + bpay
    probe = [(kt, pk, rng.random(npr) >= nullf / 2)] + ppay
    layout = [c[0] for c in build] + [c[0] for c in probe]
    nbc = len(build)
    choices = [nbc] + list(range(1, nbc)) + list(range(nbc + 1, len(layout)))  # probe key, payloads
    k = int(rng.integers(1, 5))
    picks = [int(x) for x in rng.choice(choices, size=k, replace=len(choices) < k)]
    out = [(i, layout[i]) for i in picks]
    return _join_plan(build, probe, 0, 0, out)


@pytest.mark.parametrize("seed", range(18))
def test_lean_kernel_shapes(ctx, seed):
    """One-probe pipelines on the lean kernel: the oracle's rows, and the same
    rows as the general pipeline kernel; the count target agrees."""
    plan = _lean_shape_plan(seed)
    say(f"lean shape seed={seed}")
    types = [t for _, t in plan.nodes[plan.root].output_attrs]
    lean = check_table(ctx.execute(plan), types)
    exp = check_table(oracle.execute(plan), types)
    assert lean == exp
    ctx.set_option("lean_kernel", 0)
    try:
        general = check_table(ctx.execute(plan), types)
    finally:
        ctx.set_option("lean_kernel", 1)
    assert general == exp
    tables = [ctx.upload(t) for t in plan.inputs]
    assert ctx.count_device(plan, tables) == len(exp)


@pytest.mark.parametrize("seed", range(0, 18, 2))
def test_lean_kernel_consecutive_pair_probes(ctx, seed):
    """Probe steps over the next two slots (the default reads 2-slot groups
    aligned to two slots, where the slot before an odd start slot is read but
    never decides the lookup): the oracle's rows either way."""
    plan = _lean_shape_plan(seed)
    say(f"lean consecutive pairs seed={seed}")
    types = [t for _, t in plan.nodes[plan.root].output_attrs]
    exp = check_table(oracle.execute(plan), types)
    ctx.set_option("aligned_pairs", 0)
    try:
        got = check_table(ctx.execute(plan), types)
        tables = [ctx.upload(t) for t in plan.inputs]
        count = ctx.count_device(plan, tables)
    finally:
        ctx.set_option("aligned_pairs", 1)
    assert got == exp
    assert count == len(exp)


def test_lean_kernel_all_null_int32_rows_cap(ctx):
    """An all-NULL INT32 output column costs one bit per row: pages close at the
    4096-row bitmap capacity instead of the bit budget."""
    n = 50_000
    keys = np.arange(n, dtype=np.int32)
    plan = _join_plan([(INT32, keys, None), (INT32, keys, np.zeros(n, bool))],
                      [(INT32, keys[::-1].copy(), None)], 0, 0, [(1, INT32), (2, INT32)])
    say("lean all-NULL INT32")
    _compare_large(ctx, plan)


def _workload_plan(wl):
    plan = wl.make_plan()
    for t in wl.tables:
        plan.new_input(t.host_table(wl.seed))
    return plan


@pytest.fixture
def radix_ctx(ctx):
    """The context with the probe-side radix partitioning forced on for any size."""
    ctx.set_option("radix_min_table_bytes", 0)
    ctx.set_option("radix_min_rows", 0)
    yield ctx
    ctx.set_option("radix_min_table_bytes", float(2**64 - 1))
    ctx.set_option("radix_min_rows", 4 << 20)


def test_radix_probe_c2_recipe(radix_ctx):
    """Partitioned probe side + XCD-ordered walk: same rows as the oracle."""
    from gpujoin import workloads

    plan = _workload_plan(workloads.c2(0.005))
    say("radix probe on c2 recipe at 0.5% scale")
    s = _compare_large(radix_ctx, plan)
    assert s["ms_probe_partition"] > 0, s


@pytest.mark.parametrize("seed", range(12))
def test_radix_probe_random_plans(radix_ctx, seed):
    """Multi-way plans (VARCHAR, duplicates, NULLs) with every probe side partitioned."""
    rng = random.Random(5000 + seed)
    ntables = rng.randint(2, 5)
    max_rows = rng.choice([50, 400, 3000])
    plan = random_plan(rng, ntables=ntables, max_rows=max_rows, key_type=rng.choice([INT32, INT64]),
                       domain=max(max_rows, 30), long_strings=seed % 4 == 0)
    say(f"radix random plan seed={seed}")
    types = [t for _, t in plan.nodes[plan.root].output_attrs]
    assert check_table(radix_ctx.execute(plan), types) == check_table(oracle.execute(plan), types)


def test_c2_recipe_scaled_vs_oracle(ctx):
    """The 50M x 500M INT64 Zipf workload at 1% scale (500K x 5M), 5% NULLs."""
    from gpujoin import workloads

    plan = _workload_plan(workloads.c2(0.01))
    say("c2 recipe at 1% scale")
    s = _compare_large(ctx, plan)
    assert s["output_rows"] > 3_000_000


def test_c2_full_build_sampled_probe_vs_oracle(ctx):
    """The c2 build side at full size (50M INT64 keys, 5% NULLs: the 2^27-slot
    table in 16K LDS regions that the benchmark builds) probed by a 2M-row
    slice from the middle of the 500M-row probe recipe; bit-exact against the
    oracle on the same rows."""
    from gpujoin import workloads

    wl = workloads.c2(1.0)
    bt, pt = wl.tables
    first, n = 123_456_789, 2_000_000
    say("c2 full build (50M) x 2M-row probe slice")
    build_dev = ctx.generate(wl.seed, bt.rows, bt.columns)
    probe = gj.ColumnarTable(n, [gj.host_generate(wl.seed, first, n, c) for c in pt.columns])
    plan = wl.make_plan()
    got = ctx.execute_device(plan, [build_dev, ctx.upload(probe)]).download()
    build = gj.ColumnarTable(bt.rows, [gj.host_generate(wl.seed, 0, bt.rows, c) for c in bt.columns])
    plan.new_input(build)
    plan.new_input(probe)
    exp = oracle.execute(plan)
    assert got.num_rows == exp.num_rows > 1_000_000
    for c in got.columns:
        assert gj.pages.validate_column(c.pages, c.type) == got.num_rows
    for a, b_ in zip(sorted_rows(slots_of(got)), sorted_rows(slots_of(exp))):
        assert np.array_equal(a, b_)


def test_c2_full_size_count_matches_pages(ctx):
    """At full c2 size (50M x 500M) the counting target and the page target
    agree on the result size, and the pages hold that many rows."""
    from gpujoin import workloads

    wl = workloads.c2(1.0)
    say("c2 full size: count vs pages")
    tables = [ctx.generate(wl.seed, t.rows, t.columns) for t in wl.tables]
    plan = wl.make_plan()
    count = ctx.count_device(plan, tables)
    res = ctx.execute_device(plan, tables)
    assert res.num_rows == count > 400_000_000
    st = ctx.stats()
    assert st["output_rows"] == count and st["retries"] == 0, st


def test_c2_full_size_fingerprint_vs_oracle(ctx):
    """Every row of the full c2 query (50M x 500M, ~407M output rows) against
    the oracle: the result pages, downloaded as the contest receives them, are
    fingerprinted by the oracle's page decoder (row count and an order-free
    sum of row hashes), and the expected count and fingerprint come from a CPU
    hash join over the same input pages, downloaded from HBM."""
    from gpujoin import workloads

    wl = workloads.c2(1.0)
    say("c2 full size: result fingerprint vs CPU join over the same pages")
    tables = [ctx.generate(wl.seed, t.rows, t.columns) for t in wl.tables]
    plan = wl.make_plan()
    for t in tables:
        t.prepare()
    res = ctx.execute_device(plan, tables)
    n_dev = res.num_rows
    pages = res.download_pages()
    res.free()
    say(f"c2 full size: {n_dev} rows downloaded, fingerprinting")
    got = oracle.table_fingerprint(pages.host_table())
    pages.release()
    build = tables[0].download_pages([0, 1])
    probe = tables[1].download_pages([0, 1])
    for t in tables:
        t.free()
    say("c2 full size: CPU hash join over the downloaded input pages")
    # plan output: build payload, probe payload, probe key
    exp = oracle.join_fingerprint(build.host_table(), 0, probe.host_table(), 0, [(0, 1), (1, 1), (1, 0)])
    build.release()
    probe.release()
    say(f"c2 full size: GPU {got}, CPU {exp}")
    assert got[0] == n_dev > 400_000_000
    assert got == exp


def test_c3_recipe_scaled_vs_oracle(ctx):
    """Snowflake: fact x 4 dimensions in one fused probe chain with bloom
    filters; the output carries a VARCHAR column with long strings."""
    from gpujoin import workloads

    plan = _workload_plan(workloads.c3(0.02))
    say("c3 recipe at 2% scale")
    types = [t for _, t in plan.nodes[plan.root].output_attrs]
    got = check_table(ctx.execute(plan), types)
    s = ctx.stats()
    exp = check_table(oracle.execute(plan), types)
    assert got == exp
    # the four small dimension tables are inserted straight from their pages
    # (no staging pipelines); the four joins run as one fused probe chain
    assert s["pipelines"] == 1, s


def test_c1_full_size_vs_oracle(ctx):
    """c1 at full size: 1M-row INT32 build x 10M-row probe, bit-exact."""
    from gpujoin import workloads

    plan = _workload_plan(workloads.c1(1.0))
    say("c1 at full size")
    s = _compare_large(ctx, plan)
    assert s["output_rows"] == 10_000_000


def test_c3_full_size_vs_oracle(ctx):
    """c3 at full size: the 36M-row fact table through the fused 4-probe chain
    (bloom filters, lazily decoded columns) with the VARCHAR output, bit-exact."""
    from gpujoin import workloads

    plan = _workload_plan(workloads.c3(1.0))
    say("c3 at full size")
    types = [t for _, t in plan.nodes[plan.root].output_attrs]
    got = check_table(ctx.execute(plan), types)
    exp = check_table(oracle.execute(plan), types)
    assert got == exp and len(got) > 1000


def test_tables_outlive_their_context():
    """Device tables and results freed after their context is destroyed."""
    ctx2 = gj.Context(0)
    keys = np.arange(1000, dtype=np.int32)
    t = ctx2.upload(gj.make_table([(INT32, keys, None)], 1000))
    plan = _join_plan([(INT32, keys, None)], [(INT32, keys, None)], 0, 0, [(0, INT32)])
    r = ctx2.execute_device(plan, [ctx2.upload(tb) for tb in plan.inputs])
    assert r.num_rows == 1000
    ctx2.close()
    r.free()
    t.free()
