"""The no-allocation count entry point (cgj_count_device_noalloc), called
through the C ABI: its counts equal cgj_count_device's on the same plans and
device tables, on consecutive calls (the second smaller, so the shared scratch
array holds the previous call's contents), without changing the device's free
memory; plans outside its limits are rejected with a message."""
import ctypes as C

import pytest

import gpujoin as gj
from gpujoin.pages import INT32, INT64

pytestmark = pytest.mark.gpu

SEED = 0x5EED00AA


def _noalloc(ctx, plan, tables):
    m = gj._Marshal(plan, [])
    arr = (C.c_void_p * max(len(tables), 1))(*[t.handle for t in tables])
    out = C.c_uint64()
    rc = gj.lib().cgj_count_device_noalloc(ctx.handle, C.byref(m.plan), arr, len(tables), C.byref(out))
    if rc != 0:
        raise RuntimeError(ctx.error())
    return out.value


def _free_bytes():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def _join(ctx, nb, npr, key_type, dup, stream):
    """A build table (unique keys, or keys repeated about 4 times) and a probe
    table of foreign keys with 10% misses, both with NULL keys and a payload."""
    bkind = gj.GEN_FK_UNIFORM if dup else (gj.GEN_PERM_KEY if key_type == INT32 else gj.GEN_UNIQUE_KEY)
    dom = max(nb // 4, 1) if dup else nb
    build = ctx.generate(SEED, nb, [
        gj.GenColumn(type=key_type, kind=bkind, stream=stream, null_stream=stream + 1, null_frac=0.05, domain=dom),
        gj.GenColumn(type=INT32, kind=gj.GEN_UNIFORM, stream=stream + 2, lo=0, hi=1000)])
    probe = ctx.generate(SEED, npr, [
        gj.GenColumn(type=key_type, kind=gj.GEN_FK_UNIFORM, stream=stream, null_stream=stream + 3, null_frac=0.1,
                     domain=dom, miss_frac=0.1),
        gj.GenColumn(type=INT64, kind=gj.GEN_UNIFORM, stream=stream + 4, lo=0, hi=1 << 40)])
    plan = gj.Plan()
    a = plan.new_scan_node(0, [(0, key_type), (1, INT32)])
    b = plan.new_scan_node(1, [(0, key_type), (1, INT64)])
    plan.root = plan.new_join_node(True, a, b, 0, 0, [(1, INT32), (3, INT64)])
    return plan, [build, probe]


@pytest.mark.parametrize("key_type,dup", [(INT32, False), (INT64, False), (INT32, True)])
def test_noalloc_count_matches_count_device(ctx, key_type, dup):
    cases = [_join(ctx, 1_500_000, 9_000_000, key_type, dup, 40),
             _join(ctx, 3_000, 50_000, key_type, dup, 60)]
    expected = [ctx.count_device(p, t) for p, t in cases]
    before = _free_bytes()
    got = [_noalloc(ctx, p, t) for p, t in cases]  # second call smaller than the first
    after = _free_bytes()
    assert got == expected and expected[0] > 1_000_000
    assert after == before
    # the same plan again, after the smaller one, over the dirty scratch
    assert _noalloc(ctx, *cases[0]) == expected[0]


def test_noalloc_rejects_what_it_cannot_do(ctx):
    big = _join(ctx, 5_000_000, 5_000_000, INT32, False, 80)
    with pytest.raises(RuntimeError, match="scratch"):
        _noalloc(ctx, *big)
    # a three-table plan
    t = ctx.generate(SEED, 1000, [gj.GenColumn(type=INT32, kind=gj.GEN_PERM_KEY, stream=90, domain=1000)])
    plan = gj.Plan()
    s = [plan.new_scan_node(i, [(0, INT32)]) for i in range(3)]
    ab = plan.new_join_node(True, s[0], s[1], 0, 0, [(0, INT32)])
    plan.root = plan.new_join_node(True, ab, s[2], 0, 0, [(0, INT32)])
    with pytest.raises(RuntimeError, match="two scans"):
        _noalloc(ctx, plan, [t, t, t])
    # the entry point still works afterwards
    small = _join(ctx, 2_000, 20_000, INT64, False, 100)
    assert _noalloc(ctx, *small) == ctx.count_device(*small)
