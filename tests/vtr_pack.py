"""An independent writer and reader of the .vtr layout (src/vtr1.c:39-135
writes it, :164-300 reads it), in Python with struct, for the format tests:
files of versions 1, 2 and 3 (the C writer only writes version 3), string
columns, empty row groups, and byte-level comparison with the C writer.

Columns are (name, kind, values, valid, annotation) with kind one of
"i64", "f64", "bool", "str"; valid is a bool array or None (all valid)."""
from __future__ import annotations

import math
import struct

import numpy as np

TYPE_CODE = {"i64": 0, "f64": 1, "bool": 2, "str": 3}   # src/types.h:7-12


def _bitmap(valid: np.ndarray, n: int) -> bytes:
    return np.packbits(np.asarray(valid, bool), bitorder="little").tobytes() if n else b""


def rowgroup_stats(kind: str, vals, valid) -> tuple[int, bytes, bytes]:
    """(has_stats, min bytes, max bytes) of one column of a row group, as the
    format defines them: the smallest and largest valid value as a scan in
    row order meets them with < and > (a NaN first value therefore stays),
    booleans as (some valid FALSE, some valid TRUE), nothing for strings,
    empty groups and all-NA columns."""
    n = len(vals)
    z = struct.pack("<q", 0)
    if kind == "str" or n == 0:
        return 0, z, z
    idx = [i for i in range(n) if valid is None or valid[i]]
    if kind == "bool":
        has_f = any(not vals[i] for i in idx)
        has_t = any(bool(vals[i]) for i in idx)
        return 1, struct.pack("<Q", int(has_f)), struct.pack("<Q", int(has_t))
    if not idx:
        return 0, z, z
    lo = hi = vals[idx[0]]
    for i in idx[1:]:
        v = vals[i]
        if v < lo:
            lo = v
        if v > hi:
            hi = v
    if kind == "i64":
        return 1, struct.pack("<q", int(lo)), struct.pack("<q", int(hi))
    return 1, struct.pack("<d", float(lo)), struct.pack("<d", float(hi))


def pack(cols: list[tuple], rg_rows: int, version: int = 3, empty_groups: tuple[int, ...] = ()) -> bytes:
    """The file's bytes; row groups of rg_rows rows, plus an empty group
    inserted before each index in empty_groups."""
    n = len(cols[0][2]) if cols else 0
    bounds = [(r0, min(n, r0 + rg_rows)) for r0 in range(0, n, rg_rows)] if n else []
    for g in sorted(empty_groups, reverse=True):
        bounds.insert(g, (0, 0))
    out = [b"VTR1", struct.pack("<HH", version, len(cols))]
    for name, kind, _, _, ann in cols:
        nb = name.encode()
        out.append(struct.pack("<H", len(nb)) + nb + struct.pack("<B", TYPE_CODE[kind]))
        if version >= 2:
            ab = (ann or "").encode()
            out.append(struct.pack("<H", len(ab)) + ab)
    out.append(struct.pack("<I", len(bounds)))
    for r0, r1 in bounds:
        m = r1 - r0
        out.append(struct.pack("<Q", m))
        if version >= 3:
            for _, kind, vals, valid, _ in cols:
                v = vals[r0:r1]
                ok = None if valid is None else valid[r0:r1]
                has, lo, hi = rowgroup_stats(kind, v, ok)
                out.append(struct.pack("<B", has) + lo + hi)
        for _, kind, vals, valid, _ in cols:
            ok = np.ones(m, bool) if valid is None else np.asarray(valid[r0:r1], bool)
            out.append(_bitmap(ok, m))
            v = vals[r0:r1]
            if kind == "i64":
                out.append(np.asarray(v, "<i8").tobytes())
            elif kind == "f64":
                out.append(np.asarray(v, "<f8").tobytes())
            elif kind == "bool":
                out.append(np.asarray(v, np.uint8).tobytes())
            else:
                enc = [s.encode() for s in v]
                offs = np.concatenate([[0], np.cumsum([len(e) for e in enc])]).astype("<i8")
                out.append(offs.tobytes() + struct.pack("<Q", int(offs[-1])) + b"".join(enc))
    return b"".join(out)


def header(data: bytes) -> dict:
    """Parses a file's header and row-group index (statistics included)."""
    pos = 0

    def take(fmt):
        nonlocal pos
        vals = struct.unpack_from(fmt, data, pos)
        pos += struct.calcsize(fmt)
        return vals

    assert data[:4] == b"VTR1"
    pos = 4
    version, n_cols = take("<HH")
    cols = []
    for _ in range(n_cols):
        (ln,) = take("<H")
        name = data[pos:pos + ln].decode()
        pos += ln
        (t,) = take("<B")
        ann = None
        if version >= 2:
            (al,) = take("<H")
            ann = data[pos:pos + al].decode() if al else None
            pos += al
        cols.append((name, t, ann))
    (n_rg,) = take("<I")
    groups = []
    for _ in range(n_rg):
        (m,) = take("<Q")
        stats = []
        if version >= 3:
            for _ in cols:
                stats.append(take("<BQQ"))
        for _, t, _ in cols:
            pos += (m + 7) // 8
            if t == 3:
                pos += 8 * (m + 1)
                (ln,) = take("<Q")
                pos += ln
            else:
                pos += m * (1 if t == 2 else 8)
        groups.append({"n_rows": m, "stats": stats})
    assert pos == len(data)
    return {"version": version, "cols": cols, "groups": groups}


def f64_bits(x: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def is_nan_bits(b: int) -> bool:
    return math.isnan(struct.unpack("<d", struct.pack("<Q", b))[0])
