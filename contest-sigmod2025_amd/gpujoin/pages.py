"""Contest page format: decoding and validation in numpy.

Page layout (reference include/plan.h:54-58, README.md:131-147,
engine/query/DataSource.hpp:14-79): u16 row count, u16 non-NULL count;
INT32 values at +4, INT64/FP64 at +8; VARCHAR end offsets (u16) at +4 and the
characters right after them; an LSB-first validity bitmap in the last
ceil(rows/8) bytes; long strings over pages tagged 0xffff (first) and 0xfffe
(continuations), each holding a u16 byte count at +2 and the bytes from +4.
"""
from __future__ import annotations

import struct
from typing import List, Sequence, Tuple

import numpy as np

PAGE_SIZE = 8192
INT32, INT64, FP64, VARCHAR = 0, 1, 2, 3
_NAMES = {INT32: "INT32", INT64: "INT64", FP64: "FP64", VARCHAR: "VARCHAR"}
_DT = {INT32: np.int32, INT64: np.int64, FP64: np.float64}


def type_name(t: int) -> str:
    return _NAMES[t]


def page_header(p: bytes) -> Tuple[int, int]:
    return struct.unpack_from("<HH", p, 0)


def _bitmap(p: bytes, nr: int) -> np.ndarray:
    nb = (nr + 7) // 8
    bm = np.frombuffer(p, dtype=np.uint8, count=nb, offset=PAGE_SIZE - nb)
    return np.unpackbits(bm, bitorder="little")[:nr].astype(bool)


def decode_column(pages: Sequence[bytes], type_: int):
    """Returns (values, valid): a numpy array (zeros at NULLs) or, for VARCHAR,
    a list of bytes (None at NULLs); `valid` is a bool array."""
    if type_ == VARCHAR:
        return _decode_varchar(pages)
    dt = _DT[type_]
    off = 4 if type_ == INT32 else 8
    vals, masks = [], []
    for p in pages:
        nr, nv = page_header(p)
        if nr == 0:
            continue
        valid = np.ones(nr, dtype=bool) if nv == nr else _bitmap(p, nr)
        v = np.zeros(nr, dtype=dt)
        v[valid] = np.frombuffer(p, dtype=dt, count=nv, offset=off)
        vals.append(v)
        masks.append(valid)
    if not vals:
        return np.zeros(0, dtype=dt), np.zeros(0, dtype=bool)
    return np.concatenate(vals), np.concatenate(masks)


def _decode_varchar(pages: Sequence[bytes]):
    out: List = []
    valid: List[bool] = []
    i = 0
    while i < len(pages):
        p = pages[i]
        nr, nv = page_header(p)
        if nr == 0xFFFF:
            s = bytearray()
            j = i
            while True:
                ln = struct.unpack_from("<H", pages[j], 2)[0]
                s += pages[j][4:4 + ln]
                j += 1
                if j >= len(pages) or page_header(pages[j])[0] != 0xFFFE:
                    break
            out.append(bytes(s))
            valid.append(True)
            i = j
            continue
        if nr == 0xFFFE:
            raise ValueError("long-string continuation page without a first page")
        bits = np.ones(nr, dtype=bool) if nv == nr else _bitmap(p, nr)
        offs = np.frombuffer(p, dtype=np.uint16, count=nv, offset=4)
        base = 4 + 2 * nv
        k = 0
        for r in range(nr):
            if bits[r]:
                st = int(offs[k - 1]) if k else 0
                out.append(bytes(p[base + st: base + int(offs[k])]))
                valid.append(True)
                k += 1
            else:
                out.append(None)
                valid.append(False)
        i += 1
    return out, np.array(valid, dtype=bool)


def validate_column(pages: Sequence[bytes], type_: int) -> int:
    """Checks every page against the format rules; returns the row count."""
    rows = 0
    i = 0
    while i < len(pages):
        p = pages[i]
        if len(p) != PAGE_SIZE:
            raise ValueError("page of wrong size")
        nr, nv = page_header(p)
        if type_ == VARCHAR and nr in (0xFFFF, 0xFFFE):
            if nr == 0xFFFE:
                raise ValueError(f"page {i}: stray long-string continuation")
            ln = struct.unpack_from("<H", p, 2)[0]
            if ln > PAGE_SIZE - 4:
                raise ValueError(f"page {i}: long-string chunk too large")
            rows += 1
            i += 1
            while i < len(pages) and page_header(pages[i])[0] == 0xFFFE:
                i += 1
            continue
        if nv > nr:
            raise ValueError(f"page {i}: {nv} values > {nr} rows")
        nb = (nr + 7) // 8
        if nr and nv != nr:
            bits = _bitmap(p, nr)
            if int(bits.sum()) != nv:
                raise ValueError(f"page {i}: bitmap has {int(bits.sum())} set bits, header says {nv}")
        if type_ == VARCHAR:
            offs = np.frombuffer(p, dtype=np.uint16, count=nv, offset=4).astype(np.int64)
            if nv and (np.any(np.diff(offs) < 0) or offs[0] < 0):
                raise ValueError(f"page {i}: string offsets decrease")
            end = 4 + 2 * nv + (int(offs[-1]) if nv else 0)
        else:
            w = 4 if type_ == INT32 else 8
            end = (4 if type_ == INT32 else 8) + w * nv
        if end + nb > PAGE_SIZE:
            raise ValueError(f"page {i}: data ({end} B) overlaps the bitmap ({nb} B)")
        rows += nr
        i += 1
    return rows
