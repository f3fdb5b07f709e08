"""The one-probe kernel is compiled per launch shape: with the page driver,
no bloom filter, no selection words and 2-word entries read as aligned pairs
(LEAN_TWO), with a packed table (LEAN_PACKED), or reading every flag from its
descriptor (LEAN_ANY). The execution options pick the shape: bloom_bits=0 and
aligned_pairs=1 give LEAN_TWO; a bloom filter or aligned_pairs=0 gives
LEAN_ANY. Every shape must give the oracle's rows, with unique and repeated
build keys, NULLs, and outputs from the probe key, a probe column and the
table's attribute."""
import random

import pytest

import gpujoin as gj
import oracle
from gpujoin.pages import FP64, INT32, INT64
from helpers import check_table

pytestmark = pytest.mark.gpu

SHAPES = {"two": {"bloom_bits": 0, "aligned_pairs": 1},
          "any_bloom": {"bloom_bits": 16, "aligned_pairs": 1},
          "any_unaligned": {"bloom_bits": 0, "aligned_pairs": 0}}


def _plan(seed, key_type, dup):
    rng = random.Random(seed)
    nb, npr = 3000, 40000
    dom = nb // 2 if dup else nb
    bkeys = [rng.randrange(dom) for _ in range(nb)] if dup else rng.sample(range(nb * 4), nb)
    bvals = [rng.random() for _ in range(nb)]
    pkeys = [rng.choice(bkeys) if rng.random() < 0.8 else (1 << 30) + rng.randrange(1 << 20) for _ in range(npr)]
    ppay = [rng.randrange(1 << 40) for _ in range(npr)]
    plan = gj.Plan()
    plan.new_input(gj.make_table([(key_type, bkeys, [rng.random() > 0.05 for _ in range(nb)]),
                                  (FP64, bvals, [rng.random() > 0.05 for _ in range(nb)])], nb))
    plan.new_input(gj.make_table([(key_type, pkeys, [rng.random() > 0.05 for _ in range(npr)]),
                                  (INT64, ppay, [rng.random() > 0.05 for _ in range(npr)])], npr))
    b = plan.new_scan_node(0, [(0, key_type), (1, FP64)])
    q = plan.new_scan_node(1, [(0, key_type), (1, INT64)])
    plan.root = plan.new_join_node(True, b, q, 0, 0, [(1, FP64), (3, INT64), (2, key_type)])
    return plan, [FP64, INT64, key_type]


@pytest.mark.parametrize("shape", sorted(SHAPES))
@pytest.mark.parametrize("dup", [False, True], ids=["unique", "repeated"])
@pytest.mark.parametrize("key_type", [INT64, INT32], ids=["int64", "int32"])
def test_lean_shape_vs_oracle(ctx, shape, dup, key_type):
    print(f"lean shape {shape}, {'repeated' if dup else 'unique'} keys, key type {key_type}", flush=True)
    plan, types = _plan(7 + 2 * int(dup) + key_type, key_type, dup)
    exp = check_table(oracle.execute(plan), types)
    # INT32 key with an FP64 attribute is not packable, so every shape runs 2-word tables
    for name, value in SHAPES[shape].items():
        ctx.set_option(name, value)
    try:
        got = check_table(ctx.execute(plan), types)
        st = ctx.stats()
    finally:
        ctx.set_option("bloom_bits", -1)
        ctx.set_option("aligned_pairs", 1)
    assert st["probe_kernel"] == 1, st
    assert got == exp
