"""Python mirror of the contest API over libgpujoin.so (ctypes).

The classes mirror the contest's C++ surface (reference include/plan.h:27-149,
engine/Execute.cpp:28-49): `Plan.new_scan_node / new_join_node / new_input`,
`ColumnarTable`, and `Context.execute(plan)` as `Contest::execute`. Pages are
`bytes` of 8192 bytes in the contest page format. Everything runs through the
C ABI declared in include/gpujoin.h; there is no Python fallback: when the
native library is missing, loading fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import pages as pagecodec
from .pages import INT32, INT64, FP64, VARCHAR, PAGE_SIZE, decode_column, type_name

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libgpujoin.so")

NODE_JOIN = 0
NODE_SCAN = 1

GEN_UNIQUE_KEY = 0
GEN_FK_ZIPF = 1
GEN_UNIFORM = 2
GEN_UNIT_F64 = 3
GEN_PERM_KEY = 4
GEN_FK_UNIFORM = 5
GEN_ASCII = 6


class HostColumn(C.Structure):
    _fields_ = [("pages", C.POINTER(C.c_void_p)), ("num_pages", C.c_uint64), ("type", C.c_int32)]


class HostTable(C.Structure):
    _fields_ = [("num_rows", C.c_uint64), ("columns", C.POINTER(HostColumn)), ("num_columns", C.c_uint32)]


class Attr(C.Structure):
    _fields_ = [("index", C.c_uint64), ("type", C.c_int32)]


class Node(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("build_left", C.c_int32),
        ("base_table_id", C.c_uint64),
        ("left", C.c_uint64),
        ("right", C.c_uint64),
        ("left_attr", C.c_uint64),
        ("right_attr", C.c_uint64),
        ("output_attrs", C.POINTER(Attr)),
        ("num_output_attrs", C.c_uint32),
    ]


class CPlan(C.Structure):
    _fields_ = [("nodes", C.POINTER(Node)), ("num_nodes", C.c_uint32), ("root", C.c_uint64)]


class ResultColumn(C.Structure):
    _fields_ = [("type", C.c_int32), ("num_pages", C.c_uint64), ("pages", C.POINTER(C.c_void_p))]


class Result(C.Structure):
    _fields_ = [("num_rows", C.c_uint64), ("num_columns", C.c_uint32), ("columns", C.POINTER(ResultColumn))]


class Stats(C.Structure):
    _fields_ = [
        ("ms_total", C.c_double),
        ("ms_build", C.c_double),
        ("ms_probe", C.c_double),
        ("input_pages", C.c_uint64),
        ("output_pages", C.c_uint64),
        ("build_tuples", C.c_uint64),
        ("probe_rows", C.c_uint64),
        ("output_rows", C.c_uint64),
        ("pipelines", C.c_uint32),
        ("retries", C.c_uint32),
        ("probe_input_pages", C.c_uint64),
        ("output_ref_pages", C.c_uint64),
        ("build_bytes", C.c_uint64),
        ("ms_probe_kernel", C.c_double),
        ("ms_probe_partition", C.c_double),
        ("probe_kernel", C.c_uint32),
        ("probe_waves", C.c_uint32),
        ("ms_plan", C.c_double),
        ("ms_stats", C.c_double),
        ("est_sets", C.c_uint32),
        ("est_fallbacks", C.c_uint32),
        ("est_stat_joins", C.c_uint32),
        ("retries_build", C.c_uint32),
        ("ms_host_upload", C.c_double),
        ("ms_host_download", C.c_double),
        ("replans", C.c_uint32),
        ("singletons", C.c_uint32),
        ("restrictions", C.c_uint32),
        ("semijoins", C.c_uint32),
        ("probe_packed", C.c_uint32),
        ("pools_compacted", C.c_uint32),
        ("blooms_skipped", C.c_uint32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class GenColumn(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("kind", C.c_int32),
        ("stream", C.c_uint64),
        ("null_stream", C.c_uint64),
        ("null_frac", C.c_double),
        ("domain", C.c_uint64),
        ("zipf_s", C.c_double),
        ("miss_frac", C.c_double),
        ("lo", C.c_uint64),
        ("hi", C.c_uint64),
    ]


class BatchStats(C.Structure):
    _fields_ = [
        ("ms_last", C.c_double),
        ("tasks", C.c_uint64),
        ("waves", C.c_uint32),
        ("launches", C.c_uint32),
        ("retries", C.c_uint32),
        ("jobs", C.c_uint32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# exported symbols of include/gpujoin.h with their ctypes signatures
_F32P = C.POINTER(C.c_float)

SIGNATURES = {
    "cgj_context_create": (C.c_void_p, [C.c_int]),
    "cgj_context_destroy": (None, [C.c_void_p]),
    "cgj_last_error": (C.c_char_p, [C.c_void_p]),
    "cgj_execute": (C.c_int, [C.c_void_p, C.POINTER(CPlan), C.POINTER(HostTable), C.c_uint32, C.POINTER(Result)]),
    "cgj_result_release": (None, [C.POINTER(Result)]),
    "cgj_dtable_upload": (C.c_void_p, [C.c_void_p, C.POINTER(HostTable)]),
    "cgj_dtable_free": (None, [C.c_void_p, C.c_void_p]),
    "cgj_dtable_num_rows": (C.c_uint64, [C.c_void_p]),
    "cgj_dtable_num_pages": (C.c_uint64, [C.c_void_p, C.c_uint32]),
    "cgj_dtable_prepare": (C.c_int, [C.c_void_p, C.c_void_p]),
    "cgj_dtable_download": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(ResultColumn)]),
    "cgj_execute_device": (C.c_int, [C.c_void_p, C.POINTER(CPlan), C.POINTER(C.c_void_p), C.c_uint32,
                                     C.POINTER(C.c_void_p)]),
    "cgj_dresult_num_rows": (C.c_uint64, [C.c_void_p]),
    "cgj_dresult_num_pages": (C.c_uint64, [C.c_void_p]),
    "cgj_dresult_download": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Result)]),
    "cgj_dresult_free": (None, [C.c_void_p, C.c_void_p]),
    "cgj_last_stats": (C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    "cgj_scan_decode": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    "cgj_encode_column": (C.c_int, [C.c_int32, C.c_void_p, C.c_char_p, C.c_void_p, C.c_uint64,
                                    C.POINTER(ResultColumn)]),
    "cgj_column_release": (None, [C.POINTER(ResultColumn)]),
    "cgj_dtable_generate": (C.c_void_p, [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(GenColumn), C.c_uint32]),
    "cgj_host_generate": (C.c_int, [C.c_uint64, C.c_uint64, C.c_uint64, C.POINTER(GenColumn),
                                    C.POINTER(ResultColumn)]),
    "cgj_sample_cardinalities": (C.c_int, [C.c_void_p, C.POINTER(CPlan), C.POINTER(C.c_void_p), C.c_uint32,
                                           C.POINTER(C.c_double), C.c_uint32, C.POINTER(C.c_uint32)]),
    "cgj_count_device_noalloc": (C.c_int, [C.c_void_p, C.POINTER(CPlan), C.POINTER(C.c_void_p), C.c_uint32,
                                           C.POINTER(C.c_uint64)]),
    "cgj_count_device": (C.c_int, [C.c_void_p, C.POINTER(CPlan), C.POINTER(C.c_void_p), C.c_uint32,
                                   C.POINTER(C.c_uint64)]),
    "cgj_copy_probe": (C.c_double, [C.c_void_p, C.c_uint64, C.c_uint32]),
    "cgj_set_option": (C.c_int, [C.c_void_p, C.c_char_p, C.c_double]),
    "cgj_stats_size": (C.c_size_t, []),
    "cgj_batch_create": (C.c_void_p, [C.c_void_p, C.POINTER(CPlan), C.c_uint32, C.POINTER(C.c_void_p),
                                      C.POINTER(C.c_uint32)]),
    "cgj_batch_run": (C.c_int, [C.c_void_p, C.c_void_p]),
    "cgj_batch_autotune": (C.c_int, [C.c_void_p, C.c_void_p]),
    "cgj_batch_info": (C.c_int, [C.c_void_p, C.POINTER(BatchStats)]),
    "cgj_batch_num_rows": (C.c_uint64, [C.c_void_p, C.c_uint32]),
    "cgj_batch_download": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(Result)]),
    "cgj_batch_free": (None, [C.c_void_p, C.c_void_p]),
    "cgj_mlp_create": (C.c_void_p, [C.c_void_p, C.c_uint32, C.c_uint64]),
    "cgj_mlp_num_params": (C.c_uint64, [C.c_void_p]),
    "cgj_mlp_set_params": (C.c_int, [C.c_void_p, C.c_void_p, _F32P]),
    "cgj_mlp_get_params": (C.c_int, [C.c_void_p, C.c_void_p, _F32P]),
    "cgj_mlp_loss_grad": (C.c_int, [C.c_void_p, C.c_void_p, _F32P, _F32P, C.c_uint32, _F32P, _F32P]),
    "cgj_mlp_train": (C.c_int, [C.c_void_p, C.c_void_p, _F32P, _F32P, C.c_uint32, C.c_uint32, C.c_uint32,
                                C.c_float, C.c_float, C.c_uint64, _F32P]),
    "cgj_mlp_predict": (C.c_int, [C.c_void_p, C.c_void_p, _F32P, C.c_uint32, _F32P]),
    "cgj_mlp_predict_f32": (C.c_int, [C.c_void_p, C.c_void_p, _F32P, C.c_uint32, _F32P]),
    "cgj_mlp_free": (None, [C.c_void_p, C.c_void_p]),
    "cgj_execute_many": (C.c_int, [C.c_void_p, C.POINTER(CPlan), C.c_uint32, C.POINTER(C.c_void_p),
                                   C.POINTER(C.c_uint32), C.c_uint32, C.POINTER(C.c_void_p)]),
    "cgj_num_features": (C.c_uint32, []),
    "cgj_set_estimator": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_float]),
    "cgj_clear_statistics": (C.c_int, [C.c_void_p]),
    "cgj_plan_features": (C.c_int64, [C.c_void_p, C.POINTER(CPlan), C.POINTER(C.c_void_p), C.c_uint32,
                                      C.POINTER(C.c_uint64), _F32P, C.c_uint64]),
    "cgj_count_subset": (C.c_int, [C.c_void_p, C.POINTER(CPlan), C.POINTER(C.c_void_p), C.c_uint32, C.c_uint64,
                                   C.POINTER(C.c_uint64)]),
    "cgj_plan_joins": (C.c_int, [C.c_void_p, C.POINTER(CPlan), C.POINTER(C.c_void_p), C.c_uint32,
                                 C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_double), C.c_uint32,
                                 C.POINTER(C.c_uint32)]),
}

EST_MAX, EST_INDEPENDENCE, EST_LEARNED, EST_SAMPLE = 0, 1, 2, 3

_lib = None


def lib():
    """Loads libgpujoin.so (built in-tree by `make`); raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"native library {LIB_PATH} is missing: build it with `make -C contest-sigmod2025_amd`")
        l = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


# ---------------------------------------------------------------------------
# contest-style plan and tables
# ---------------------------------------------------------------------------
@dataclass
class Column:
    type: int
    pages: List[bytes] = field(default_factory=list)


@dataclass
class ColumnarTable:
    num_rows: int = 0
    columns: List[Column] = field(default_factory=list)


@dataclass
class PlanNode:
    kind: int
    output_attrs: List[Tuple[int, int]]
    base_table_id: int = 0
    build_left: bool = False
    left: int = 0
    right: int = 0
    left_attr: int = 0
    right_attr: int = 0


class Plan:
    """Mirror of the contest Plan (reference include/plan.h:112-149)."""

    def __init__(self):
        self.nodes: List[PlanNode] = []
        self.inputs: List[ColumnarTable] = []
        self.root = 0

    def new_scan_node(self, base_table_id, output_attrs):
        self.nodes.append(PlanNode(NODE_SCAN, list(output_attrs), base_table_id=base_table_id))
        return len(self.nodes) - 1

    def new_join_node(self, build_left, left, right, left_attr, right_attr, output_attrs):
        self.nodes.append(PlanNode(NODE_JOIN, list(output_attrs), build_left=bool(build_left), left=left,
                                   right=right, left_attr=left_attr, right_attr=right_attr))
        return len(self.nodes) - 1

    def new_input(self, table: ColumnarTable):
        self.inputs.append(table)
        return len(self.inputs) - 1


class _Marshal:
    """Keeps ctypes buffers alive while a call uses them."""

    def __init__(self, plan: Plan, tables: Optional[Sequence[ColumnarTable]] = None):
        self.keep = []
        n = len(plan.nodes)
        self.nodes = (Node * max(n, 1))()
        for i, pn in enumerate(plan.nodes):
            attrs = (Attr * max(len(pn.output_attrs), 1))()
            for k, (idx, t) in enumerate(pn.output_attrs):
                attrs[k].index = idx
                attrs[k].type = t
            self.keep.append(attrs)
            nd = self.nodes[i]
            nd.kind = pn.kind
            nd.build_left = 1 if pn.build_left else 0
            nd.base_table_id = pn.base_table_id
            nd.left, nd.right = pn.left, pn.right
            nd.left_attr, nd.right_attr = pn.left_attr, pn.right_attr
            nd.output_attrs = attrs
            nd.num_output_attrs = len(pn.output_attrs)
        self.plan = CPlan(self.nodes, n, plan.root)
        tables = plan.inputs if tables is None else tables
        self.tables = (HostTable * max(len(tables), 1))()
        for t, tab in enumerate(tables):
            cols = (HostColumn * max(len(tab.columns), 1))()
            for c, col in enumerate(tab.columns):
                # the library only reads input pages: a full-size bytes page is
                # passed by the address of its own buffer, anything else copied
                bufs = [C.c_char_p(p) if type(p) is bytes and len(p) == PAGE_SIZE
                        else C.create_string_buffer(bytes(p), PAGE_SIZE) for p in col.pages]
                ptrs = (C.c_void_p * max(len(bufs), 1))(*[C.cast(b, C.c_void_p) for b in bufs])
                self.keep.extend(bufs)
                self.keep.append(ptrs)
                cols[c].pages = C.cast(ptrs, C.POINTER(C.c_void_p))
                cols[c].num_pages = len(bufs)
                cols[c].type = col.type
            self.keep.append(cols)
            self.tables[t].num_rows = tab.num_rows
            self.tables[t].columns = cols
            self.tables[t].num_columns = len(tab.columns)
        self.ntables = len(tables)


def result_to_table(res: Result) -> ColumnarTable:
    out = ColumnarTable(num_rows=res.num_rows)
    for c in range(res.num_columns):
        rc = res.columns[c]
        col = Column(rc.type)
        for i in range(rc.num_pages):
            col.pages.append(C.string_at(rc.pages[i], PAGE_SIZE))
        out.columns.append(col)
    return out


def encode_column(type_: int, values, valid=None) -> Column:
    """Pages of a column written with the contest page writer (ColumnInserter)."""
    l = lib()
    n = len(values)
    vbuf = None
    chars = None
    if type_ == VARCHAR:
        enc = [v.encode() if isinstance(v, str) else bytes(v) for v in values]
        offs = np.zeros(n + 1, dtype=np.uint64)
        if n:
            offs[1:] = np.cumsum([len(e) for e in enc])
        chars = b"".join(enc)
        vbuf = offs
    else:
        dt = {INT32: np.int32, INT64: np.int64, FP64: np.float64}[type_]
        vbuf = np.ascontiguousarray(np.asarray(values, dtype=dt) if n else np.zeros(1, dtype=dt))
    vb = None
    if valid is not None:
        vb = np.ascontiguousarray(np.asarray(valid, dtype=np.uint8))
    out = ResultColumn()
    rc = l.cgj_encode_column(type_, vbuf.ctypes.data, chars if chars is not None else None,
                             vb.ctypes.data if vb is not None else None, n, C.byref(out))
    if rc != 0:
        raise RuntimeError("cgj_encode_column failed")
    col = Column(type_, [C.string_at(out.pages[i], PAGE_SIZE) for i in range(out.num_pages)])
    l.cgj_column_release(C.byref(out))
    return col


def make_table(columns: Sequence[Tuple[int, Sequence, Optional[Sequence]]], num_rows: Optional[int] = None):
    """Builds a ColumnarTable from (type, values, valid-or-None) per column."""
    cols = [encode_column(t, v, m) for (t, v, m) in columns]
    n = num_rows if num_rows is not None else (len(columns[0][1]) if columns else 0)
    return ColumnarTable(n, cols)


def host_generate(seed: int, first_row: int, num_rows: int, gen: GenColumn) -> Column:
    l = lib()
    out = ResultColumn()
    if l.cgj_host_generate(seed, first_row, num_rows, C.byref(gen), C.byref(out)) != 0:
        raise RuntimeError("cgj_host_generate failed")
    col = Column(out.type, [C.string_at(out.pages[i], PAGE_SIZE) for i in range(out.num_pages)])
    l.cgj_column_release(C.byref(out))
    return col


class DeviceTable:
    def __init__(self, ctx: "Context", handle):
        self.ctx = ctx
        self.handle = handle

    @property
    def num_rows(self):
        return lib().cgj_dtable_num_rows(self.handle)

    def num_pages(self, column):
        return lib().cgj_dtable_num_pages(self.handle, column)

    def download_pages(self, columns) -> HostPages:
        """Host copies of the given columns' pages (the pages a query scanned)."""
        n = lib().cgj_dtable_num_rows(self.handle)
        cols = []
        try:
            for c in columns:
                rc = ResultColumn()
                if lib().cgj_dtable_download(self.ctx.handle, self.handle, c, C.byref(rc)) != 0:
                    raise RuntimeError(self.ctx.error())
                cols.append(rc)
        except Exception:
            for rc in cols:
                lib().cgj_column_release(C.byref(rc))
            raise

        def rel():
            for rc in cols:
                lib().cgj_column_release(C.byref(rc))
        return HostPages(n, cols, rel)

    def prepare(self):
        """Rebuilds the row prefix and block index of every column (the
        reference's per-query search vectors)."""
        if lib().cgj_dtable_prepare(self.ctx.handle, self.handle) != 0:
            raise RuntimeError(self.ctx.error())

    def decode(self, column):
        """Decodes a column on the GPU into uint64 slots (NULL = all ones)."""
        out = np.zeros(max(self.num_rows, 1), dtype=np.uint64)
        if lib().cgj_scan_decode(self.ctx.handle, self.handle, column, out.ctypes.data) != 0:
            raise RuntimeError(self.ctx.error())
        return out[: self.num_rows]

    def free(self):
        if self.handle:
            lib().cgj_dtable_free(self.ctx.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class HostPages:
    """Library-owned host pages of a table (a downloaded result or device
    table), kept as page pointers without per-page Python objects: a checker
    reads them through host_table(). Call release() when done."""

    def __init__(self, num_rows, columns, release):
        self.num_rows = num_rows
        self._cols = columns          # ResultColumn structures
        self._release = release
        self._view = None

    def host_table(self) -> HostTable:
        if self._view is None:
            cols = (HostColumn * max(len(self._cols), 1))()
            for i, c in enumerate(self._cols):
                cols[i] = HostColumn(c.pages, c.num_pages, c.type)
            self._view = (cols, HostTable(self.num_rows, cols, len(self._cols)))
        return self._view[1]

    def release(self):
        if self._release:
            self._release()
            self._release = None
            self._cols = []
            self._view = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class DeviceResult:
    def __init__(self, ctx: "Context", handle):
        self.ctx = ctx
        self.handle = handle

    @property
    def num_rows(self):
        return lib().cgj_dresult_num_rows(self.handle)

    @property
    def num_pages(self):
        return lib().cgj_dresult_num_pages(self.handle)

    def download(self) -> ColumnarTable:
        res = Result()
        if lib().cgj_dresult_download(self.ctx.handle, self.handle, C.byref(res)) != 0:
            raise RuntimeError(self.ctx.error())
        try:
            return result_to_table(res)
        finally:
            lib().cgj_result_release(C.byref(res))

    def download_pages(self) -> HostPages:
        """The result's pages in host memory as the contest receives them,
        without Python copies."""
        res = Result()
        if lib().cgj_dresult_download(self.ctx.handle, self.handle, C.byref(res)) != 0:
            raise RuntimeError(self.ctx.error())
        cols = [res.columns[i] for i in range(res.num_columns)]
        return HostPages(res.num_rows, cols, lambda: lib().cgj_result_release(C.byref(res)))

    def free(self):
        if self.handle:
            lib().cgj_dresult_free(self.ctx.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Batch:
    """Independent two-table join plans over device tables, run together by one
    persistent kernel launch (cgj_batch_*)."""

    def __init__(self, ctx: "Context", plans: Sequence[Plan], tables: Sequence[Sequence[DeviceTable]]):
        self.ctx = ctx
        self._marshals = [_Marshal(p, []) for p in plans]
        cplans = (CPlan * max(len(plans), 1))(*[m.plan for m in self._marshals])
        flat = [t.handle for ts in tables for t in ts]
        arr = (C.c_void_p * max(len(flat), 1))(*flat)
        counts = (C.c_uint32 * max(len(plans), 1))(*[len(ts) for ts in tables])
        self._tables = tables
        self.handle = lib().cgj_batch_create(ctx.handle, cplans, len(plans), arr, counts)
        if not self.handle:
            raise RuntimeError(ctx.error())
        self.size = len(plans)

    def run(self):
        if lib().cgj_batch_run(self.ctx.handle, self.handle) != 0:
            raise RuntimeError(self.ctx.error())

    def autotune(self):
        if lib().cgj_batch_autotune(self.ctx.handle, self.handle) != 0:
            raise RuntimeError(self.ctx.error())

    def info(self) -> dict:
        s = BatchStats()
        lib().cgj_batch_info(self.handle, C.byref(s))
        return s.as_dict()

    def num_rows(self, j: int) -> int:
        return lib().cgj_batch_num_rows(self.handle, j)

    def download(self, j: int) -> ColumnarTable:
        res = Result()
        if lib().cgj_batch_download(self.ctx.handle, self.handle, j, C.byref(res)) != 0:
            raise RuntimeError(self.ctx.error())
        try:
            return result_to_table(res)
        finally:
            lib().cgj_result_release(C.byref(res))

    def download_pages(self, j: int) -> "HostPages":
        """Job j's result pages in host memory, without Python copies."""
        res = Result()
        if lib().cgj_batch_download(self.ctx.handle, self.handle, j, C.byref(res)) != 0:
            raise RuntimeError(self.ctx.error())
        cols = [res.columns[i] for i in range(res.num_columns)]
        return HostPages(res.num_rows, cols, lambda: lib().cgj_result_release(C.byref(res)))

    def free(self):
        if self.handle:
            lib().cgj_batch_free(self.ctx.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _f32(a, shape=None):
    import numpy as np
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None and a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a, a.ctypes.data_as(_F32P)


class Mlp:
    """The cardinality estimator's network on the device (cgj_mlp_*): three
    layers in -> 256 -> 256 -> 2, outputs (mu, log sigma)."""

    def __init__(self, ctx: "Context", in_features: int, seed: int = 1):
        self.ctx = ctx
        self.in_features = in_features
        self.handle = lib().cgj_mlp_create(ctx.handle, in_features, seed)
        if not self.handle:
            raise RuntimeError(ctx.error())
        self.num_params = lib().cgj_mlp_num_params(self.handle)

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(self.ctx.error())

    def set_params(self, params):
        a, p = _f32(params, (self.num_params,))
        self._check(lib().cgj_mlp_set_params(self.ctx.handle, self.handle, p))

    def get_params(self):
        import numpy as np
        out = np.empty(self.num_params, np.float32)
        self._check(lib().cgj_mlp_get_params(self.ctx.handle, self.handle, out.ctypes.data_as(_F32P)))
        return out

    def loss_grad(self, x, y):
        import numpy as np
        xa, xp = _f32(x)
        n = xa.shape[0]
        if xa.ndim != 2 or xa.shape[1] != self.in_features:
            raise ValueError("x must be n x in_features")
        ya, yp = _f32(y, (n,))
        loss = C.c_float()
        g = np.empty(self.num_params, np.float32)
        self._check(lib().cgj_mlp_loss_grad(self.ctx.handle, self.handle, xp, yp, n, C.byref(loss),
                                            g.ctypes.data_as(_F32P)))
        return loss.value, g

    def train(self, x, y, epochs: int, batch: int, lr: float, seed: int = 1, weight_decay: float = 0.0) -> float:
        xa, xp = _f32(x)
        n = xa.shape[0]
        if xa.ndim != 2 or xa.shape[1] != self.in_features:
            raise ValueError("x must be n x in_features")
        ya, yp = _f32(y, (n,))
        loss = C.c_float()
        self._check(lib().cgj_mlp_train(self.ctx.handle, self.handle, xp, yp, n, epochs, batch, lr, weight_decay,
                                        seed, C.byref(loss)))
        return loss.value

    def _predict(self, fn, x):
        import numpy as np
        xa, xp = _f32(x)
        if xa.ndim != 2 or xa.shape[1] != self.in_features:
            raise ValueError("x must be n x in_features")
        out = np.empty((xa.shape[0], 2), np.float32)
        self._check(fn(self.ctx.handle, self.handle, xp, xa.shape[0], out.ctypes.data_as(_F32P)))
        return out

    def predict(self, x):
        """BF16 MFMA inference, FP32 accumulation."""
        return self._predict(lib().cgj_mlp_predict, x)

    def predict_f32(self, x):
        """FP32 forward (the training path's)."""
        return self._predict(lib().cgj_mlp_predict_f32, x)

    def free(self):
        if self.handle:
            lib().cgj_mlp_free(self.ctx.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """Mirror of Contest::build_context / execute / destroy_context."""

    def __init__(self, device: int = 0):
        self.handle = lib().cgj_context_create(device)
        if not self.handle:
            raise RuntimeError("cgj_context_create failed (no GPU?)")

    def error(self):
        return lib().cgj_last_error(self.handle).decode()

    def execute(self, plan: Plan) -> ColumnarTable:
        m = _Marshal(plan)
        res = Result()
        if lib().cgj_execute(self.handle, C.byref(m.plan), m.tables, m.ntables, C.byref(res)) != 0:
            raise RuntimeError(self.error())
        try:
            return result_to_table(res)
        finally:
            lib().cgj_result_release(C.byref(res))

    def upload(self, table: ColumnarTable) -> DeviceTable:
        m = _Marshal(Plan(), [table])
        h = lib().cgj_dtable_upload(self.handle, C.byref(m.tables[0]))
        if not h:
            raise RuntimeError(self.error())
        return DeviceTable(self, h)

    def generate(self, seed: int, num_rows: int, gens: Sequence[GenColumn]) -> DeviceTable:
        arr = (GenColumn * len(gens))(*gens)
        h = lib().cgj_dtable_generate(self.handle, seed, num_rows, arr, len(gens))
        if not h:
            raise RuntimeError(self.error())
        return DeviceTable(self, h)

    def execute_device(self, plan: Plan, tables: Sequence[DeviceTable]) -> DeviceResult:
        m = _Marshal(plan, [])
        arr = (C.c_void_p * max(len(tables), 1))(*[t.handle for t in tables])
        out = C.c_void_p()
        if lib().cgj_execute_device(self.handle, C.byref(m.plan), arr, len(tables), C.byref(out)) != 0:
            raise RuntimeError(self.error())
        return DeviceResult(self, out.value)

    def execute_many(self, plans: Sequence[Plan], tables: Sequence[Sequence[DeviceTable]],
                     workers: int = 4) -> List[DeviceResult]:
        """Runs independent plans concurrently on `workers` streams; one result per plan."""
        marshals = [_Marshal(p, []) for p in plans]
        cplans = (CPlan * max(len(plans), 1))(*[m.plan for m in marshals])
        flat = [t.handle for ts in tables for t in ts]
        arr = (C.c_void_p * max(len(flat), 1))(*flat)
        counts = (C.c_uint32 * max(len(plans), 1))(*[len(ts) for ts in tables])
        out = (C.c_void_p * max(len(plans), 1))()
        if lib().cgj_execute_many(self.handle, cplans, len(plans), arr, counts, workers, out) != 0:
            raise RuntimeError(self.error())
        return [DeviceResult(self, out[i]) for i in range(len(plans))]

    def count_device(self, plan: Plan, tables: Sequence[DeviceTable]) -> int:
        """Result cardinality of the plan over device tables (no pages written)."""
        m = _Marshal(plan, [])
        arr = (C.c_void_p * max(len(tables), 1))(*[t.handle for t in tables])
        out = C.c_uint64()
        if lib().cgj_count_device(self.handle, C.byref(m.plan), arr, len(tables), C.byref(out)) != 0:
            raise RuntimeError(self.error())
        return out.value

    def batch(self, plans: Sequence[Plan], tables: Sequence[Sequence[DeviceTable]]) -> "Batch":
        """Prepares plans (each one join of two scans) for batched execution."""
        return Batch(self, plans, tables)

    def mlp(self, in_features: int, seed: int = 1) -> "Mlp":
        return Mlp(self, in_features, seed)

    # -- cardinality estimation for join ordering --------------------------
    def set_estimator(self, mode: int, mlp: Optional["Mlp"] = None, max_sigma: float = 0.5):
        """EST_SAMPLE (the reference's planner: the max-of-inputs rule over
        sampled cardinalities, the default), EST_MAX (the rule over exact
        sizes), EST_INDEPENDENCE, or EST_LEARNED with a network of
        num_features() inputs (low-confidence sets take the sampling estimate)."""
        h = mlp.handle if mlp is not None else None
        if lib().cgj_set_estimator(self.handle, mode, h, max_sigma) != 0:
            raise RuntimeError(self.error())

    def clear_statistics(self):
        """Empties the catalog of pairwise join counts the independence and
        learned estimators read (they are counted again on demand)."""
        if lib().cgj_clear_statistics(self.handle) != 0:
            raise RuntimeError(self.error())

    @staticmethod
    def num_features() -> int:
        return lib().cgj_num_features()

    def _plan_args(self, plan: Plan, tables: Sequence[DeviceTable]):
        m = _Marshal(plan, [])
        arr = (C.c_void_p * max(len(tables), 1))(*[t.handle for t in tables])
        return m, arr

    def sample_cardinalities(self, plan: Plan, tables: Sequence[DeviceTable]):
        """The sampling estimate of each relation of the plan (numbered as for
        plan_features); a set's estimate is the largest of its relations'."""
        import numpy as np
        m, arr = self._plan_args(plan, tables)
        cards = np.zeros(64, np.float64)
        n = C.c_uint32()
        if lib().cgj_sample_cardinalities(self.handle, C.byref(m.plan), arr, len(tables),
                                          cards.ctypes.data_as(C.POINTER(C.c_double)), 64, C.byref(n)) != 0:
            raise RuntimeError(self.error())
        return cards[: n.value]

    def plan_features(self, plan: Plan, tables: Sequence[DeviceTable]):
        """(sets, features) of every connected relation set of two or more
        relations; relation r is the plan's r-th scan in depth-first order."""
        import numpy as np
        m, arr = self._plan_args(plan, tables)
        n = lib().cgj_plan_features(self.handle, C.byref(m.plan), arr, len(tables), None, None, 0)
        if n < 0:
            raise RuntimeError(self.error())
        sets = np.zeros(max(n, 1), np.uint64)
        feats = np.zeros((max(n, 1), self.num_features()), np.float32)
        n2 = lib().cgj_plan_features(self.handle, C.byref(m.plan), arr, len(tables),
                                     sets.ctypes.data_as(C.POINTER(C.c_uint64)), feats.ctypes.data_as(_F32P), n)
        if n2 != n:
            raise RuntimeError(self.error())
        return sets[:n], feats[:n]

    def count_subset(self, plan: Plan, tables: Sequence[DeviceTable], relset: int) -> int:
        """Exact result rows of the join of a connected relation set of the plan."""
        m, arr = self._plan_args(plan, tables)
        out = C.c_uint64()
        if lib().cgj_count_subset(self.handle, C.byref(m.plan), arr, len(tables), relset, C.byref(out)) != 0:
            raise RuntimeError(self.error())
        return out.value

    def plan_joins(self, plan: Plan, tables: Sequence[DeviceTable]):
        """Join order under the current estimator: [(build set, probe set, estimated rows)], children first."""
        m, arr = self._plan_args(plan, tables)
        cap = 64
        b = (C.c_uint64 * cap)()
        pr = (C.c_uint64 * cap)()
        cards = (C.c_double * cap)()
        n = C.c_uint32()
        if lib().cgj_plan_joins(self.handle, C.byref(m.plan), arr, len(tables), b, pr, cards, cap, C.byref(n)) != 0:
            raise RuntimeError(self.error())
        return [(b[i], pr[i], cards[i]) for i in range(min(n.value, cap))]

    def copy_probe(self, nbytes: int, reps: int = 5) -> float:
        """Mean time (ms) of a plain HBM-to-HBM copy of `nbytes` bytes."""
        ms = lib().cgj_copy_probe(self.handle, nbytes, reps)
        if ms < 0:
            raise RuntimeError(self.error())
        return ms

    def set_option(self, name: str, value: float):
        if lib().cgj_set_option(self.handle, name.encode(), float(value)) != 0:
            raise ValueError(f"unknown option {name}")

    def stats(self) -> dict:
        s = Stats()
        lib().cgj_last_stats(self.handle, C.byref(s))
        return s.as_dict()

    def close(self):
        if self.handle:
            lib().cgj_context_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def table_rows(t: ColumnarTable):
    """Decodes a ColumnarTable into a list of row tuples (None = NULL)."""
    cols = [decode_column(c.pages, c.type) for c in t.columns]
    for vals, valid in cols:
        if len(valid) != t.num_rows:
            raise ValueError(f"column holds {len(valid)} rows, table says {t.num_rows}")
    rows = []
    for r in range(t.num_rows):
        rows.append(tuple((v[r] if not isinstance(v, np.ndarray) else v[r].item()) if m[r] else None
                          for (v, m) in cols))
    return rows
