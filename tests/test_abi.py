"""The C-ABI library loads and exports every symbol include/gpujoin.h declares;
host-side helpers of the ABI (page writer, host generator) work without a GPU."""
import os
import re

import numpy as np
import pytest

import gpujoin as gj
from gpujoin.pages import FP64, INT32, INT64, VARCHAR

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "gpujoin.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(cgj_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = gj.lib()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), f"{n} is declared in include/gpujoin.h but not exported"
    assert set(names) == set(gj.SIGNATURES), "ctypes signatures and header disagree"


def test_stats_struct_matches_binding():
    import ctypes

    assert gj.lib().cgj_stats_size() == ctypes.sizeof(gj.Stats)


def test_contest_symbols_exported():
    """The C++ contest surface (Contest::build_context/destroy_context/execute)."""
    import subprocess

    out = subprocess.run(["nm", "-D", "-C", gj.LIB_PATH], capture_output=True, text=True).stdout
    for sym in ("Contest::build_context()", "Contest::destroy_context(void*)", "Contest::execute(Plan const&, void*)"):
        assert sym in out, sym


@pytest.mark.parametrize("t", [INT32, INT64, FP64, VARCHAR])
def test_page_writer_round_trip(t):
    rng = np.random.default_rng(t)
    n = 7000
    valid = rng.random(n) > 0.1
    if t == VARCHAR:
        vals = ["x" * int(k) for k in rng.integers(0, 90, n)]
        vals[17] = "L" * 20000  # long string
        valid[17] = True
    elif t == FP64:
        vals = rng.random(n)
    else:
        vals = rng.integers(0, 1 << 30, n)
    col = gj.encode_column(t, vals, valid)
    if t in (INT32, VARCHAR):  # 8-byte pages may carry the contest writer's bitmap overlap
        assert gj.pages.validate_column(col.pages, t) == n
    v, m = gj.decode_column(col.pages, t)
    assert np.array_equal(m, valid)
    if t == VARCHAR:
        assert [x for x, ok in zip(v, m) if ok] == [s.encode() for s, ok in zip(vals, valid) if ok]
    elif t == INT32:
        assert np.array_equal(v[m], np.asarray(vals)[valid])


def test_page_writer_capacities():
    """ColumnInserter packs 1984 INT32 / 1007 INT64 values per page without NULLs
    (the reference's page writer gives 5041 and 9935 pages for 10M rows)."""
    n = 10_000_000 // 100
    c32 = gj.encode_column(INT32, np.arange(n, dtype=np.int32))
    c64 = gj.encode_column(INT64, np.arange(n, dtype=np.int64))
    assert gj.pages.page_header(c32.pages[0]) == (1984, 1984)
    assert gj.pages.page_header(c64.pages[0]) == (1007, 1007)
    assert len(c32.pages) == -(-n // 1984) and len(c64.pages) == -(-n // 1007)


def test_host_generator_matches_spec():
    g = gj.GenColumn(type=INT32, kind=gj.GEN_PERM_KEY, stream=1, null_stream=2, null_frac=0.0, domain=5000)
    col = gj.host_generate(0x5EED0001, 0, 5000, g)
    v, m = gj.decode_column(col.pages, INT32)
    assert m.all() and sorted(v.tolist()) == list(range(5000))  # a permutation
    fk = gj.GenColumn(type=INT32, kind=gj.GEN_FK_UNIFORM, stream=1, null_stream=3, null_frac=0.05, domain=5000,
                      miss_frac=0.1)
    col = gj.host_generate(0x5EED0001, 0, 20000, fk)
    v, m = gj.decode_column(col.pages, INT32)
    hit = (v[m] < 5000).mean()
    assert 0.85 < hit < 0.95 and 0.03 < 1 - m.mean() < 0.07
