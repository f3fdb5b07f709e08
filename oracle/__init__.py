"""CPU oracle for the join pipeline. TEST INFRASTRUCTURE ONLY.

Only the tests, the smoke check and the benchmark's CPU-baseline leg use this
package, as the checker; the product path never imports it. The restatement
itself is plain C++ (oracle.cpp, built into liboracle.so by `make -C oracle`);
see oracle.cpp's header for the reference files and lines each part follows.

Pinning (what each function is checked against, tests/test_oracle.py):
  * oracle_execute: the twelve literal ground truths of the reference's
    contest-level tests (tests/golden/unit_cases.json, transcribed from the
    reference tests/unit_tests.cpp), a brute-force nested-loop join over random
    small plans (a different algorithm), and the reference's page capacities
    (1984 INT32 / 1007 INT64 rows per page) and long-string page counts.
  * oracle_decode_fixed: round trips through the contest page writer.
  * oracle_join_count: brute force on small inputs and the closed form of the
    unique-key joins used at benchmark scale.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

_lib = None


def _gj():
    root = os.path.dirname(_HERE)
    pkg = os.path.join(root, "contest-sigmod2025_amd")
    if pkg not in sys.path:
        sys.path.insert(0, pkg)
    import gpujoin  # noqa: E402

    return gpujoin


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        gj = _gj()
        l = C.CDLL(LIB_PATH)
        l.oracle_execute.restype = C.c_int
        l.oracle_execute.argtypes = [C.POINTER(gj.CPlan), C.POINTER(gj.HostTable), C.c_uint32,
                                     C.POINTER(gj.Result), C.c_char_p, C.c_size_t]
        l.oracle_result_release.restype = None
        l.oracle_result_release.argtypes = [C.POINTER(gj.Result)]
        l.oracle_decode_fixed.restype = C.c_int
        l.oracle_decode_fixed.argtypes = [C.POINTER(gj.HostColumn), C.c_uint64, C.c_void_p, C.c_char_p, C.c_size_t]
        l.oracle_table_fingerprint.restype = C.c_int
        l.oracle_table_fingerprint.argtypes = [C.POINTER(gj.HostTable), C.c_uint32, C.POINTER(C.c_uint64),
                                               C.POINTER(C.c_uint64), C.c_char_p, C.c_size_t]
        l.oracle_join_fingerprint.restype = C.c_int
        l.oracle_join_fingerprint.argtypes = [C.POINTER(gj.HostTable), C.c_uint32, C.POINTER(gj.HostTable),
                                              C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.c_uint32,
                                              C.c_uint32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                              C.c_char_p, C.c_size_t]
        l.oracle_join_count.restype = C.c_uint64
        l.oracle_join_count.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64]
        _lib = l
    return _lib


def execute(plan):
    """Evaluates a gpujoin.Plan on the CPU; returns a gpujoin.ColumnarTable."""
    gj = _gj()
    m = gj._Marshal(plan)
    res = gj.Result()
    err = C.create_string_buffer(512)
    if lib().oracle_execute(C.byref(m.plan), m.tables, m.ntables, C.byref(res), err, 512) != 0:
        raise RuntimeError(err.value.decode())
    try:
        return gj.result_to_table(res)
    finally:
        lib().oracle_result_release(C.byref(res))


def decode_fixed(column, num_rows):
    """Decodes a fixed-width gpujoin.Column into uint64 slots (NULL = all ones)."""
    import numpy as np

    gj = _gj()
    t = gj.ColumnarTable(num_rows, [column])
    m = gj._Marshal(gj.Plan(), [t])
    out = np.zeros(max(num_rows, 1), dtype=np.uint64)
    err = C.create_string_buffer(512)
    if lib().oracle_decode_fixed(m.tables[0].columns, num_rows, out.ctypes.data, err, 512) != 0:
        raise RuntimeError(err.value.decode())
    return out[:num_rows]


def join_count(build_keys, probe_keys):
    """Number of matching pairs of two uint64 key arrays (NULL = all ones never matches)."""
    import numpy as np

    b = np.ascontiguousarray(build_keys, dtype=np.uint64)
    p = np.ascontiguousarray(probe_keys, dtype=np.uint64)
    return int(lib().oracle_join_count(b.ctypes.data, len(b), p.ctypes.data, len(p)))


def _threads(threads):
    return threads or max(1, min(16, os.cpu_count() or 1))


def table_fingerprint(host_table, threads=0):
    """(rows, fingerprint) of a gpujoin.HostTable of fixed-width columns: the
    sum over rows of a hash of the row's slots, independent of row order and
    page packing."""
    n, fp = C.c_uint64(), C.c_uint64()
    err = C.create_string_buffer(512)
    if lib().oracle_table_fingerprint(C.byref(host_table), _threads(threads), C.byref(n), C.byref(fp), err, 512):
        raise RuntimeError(err.value.decode())
    return n.value, fp.value


def join_fingerprint(build, bkey, probe, pkey, outputs, threads=0):
    """(rows, fingerprint) of the inner equi-join build[bkey] = probe[pkey] of
    two gpujoin.HostTables projected on `outputs` ((side, column) pairs, side 0
    = build, 1 = probe), computed by a threaded CPU hash join."""
    side = (C.c_uint32 * max(len(outputs), 1))(*[s for s, _ in outputs])
    col = (C.c_uint32 * max(len(outputs), 1))(*[c for _, c in outputs])
    n, fp = C.c_uint64(), C.c_uint64()
    err = C.create_string_buffer(512)
    if lib().oracle_join_fingerprint(C.byref(build), bkey, C.byref(probe), pkey, side, col, len(outputs),
                                     _threads(threads), C.byref(n), C.byref(fp), err, 512):
        raise RuntimeError(err.value.decode())
    return n.value, fp.value
