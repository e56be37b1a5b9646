"""Test helpers: tables, plans, the CPU oracle and the HIP engine via ctypes.

The oracle (oracle/libcj_oracle.so) is the checker; the engine
(calamarijoins_amd/libcalamarijoins_amd.so) is what is tested.  Both take
the same cj_api.h descriptors, built here from Python rows or from the
engine's synthetic-table generator.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import struct
import subprocess
import sys

import numpy as np

from calamarijoins_amd import abi

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(REPO, "oracle", "libcj_oracle.so")
PAGE = abi.PAGE_SIZE
T = abi.TYPE_NAMES  # name -> id

sys.path.insert(0, os.path.join(REPO, "oracle"))
from cj_oracle import OracleResult, OracleRun  # noqa: E402,F401
from cj_oracle import lib as oracle  # noqa: E402,F401


# --------------------------------------------------------------- columns
from calamarijoins_amd.host import Column, Table, Plan, Engine, EngineResult  # noqa: E402,F401


def encode_fixed(type_id: int, values, valid) -> np.ndarray:
    n = len(valid)
    w = 4 if type_id == abi.CJ_INT32 else 8
    dt = {abi.CJ_INT32: np.int32, abi.CJ_INT64: np.int64, abi.CJ_FP64: np.float64}[type_id]
    vals = np.ascontiguousarray(values, dtype=dt)
    val = np.ascontiguousarray(valid, dtype=np.uint8)
    per = 1984 if w == 4 else 1007
    maxp = n // per + n // 60000 + 4
    out = np.zeros((maxp, PAGE), dtype=np.uint8)
    got = oracle().cjo_encode_fixed(type_id, n, vals.ctypes.data, val.ctypes.data, out.ctypes.data, maxp)
    assert got != C.c_size_t(-1).value
    return out[:got].copy()


def encode_varchar(strings) -> np.ndarray:
    n = len(strings)
    valid = np.array([s is not None for s in strings], dtype=np.uint8)
    blobs = [(s.encode() if isinstance(s, str) else s) if s is not None else b"" for s in strings]
    offs = np.zeros(n + 1, dtype=np.uint64)
    if n:
        offs[1:] = np.cumsum([len(b) for b in blobs])
    data = b"".join(blobs) + b"\0"
    buf = C.create_string_buffer(data, len(data))
    maxp = n + int(offs[-1]) // (PAGE - 8) + 4
    out = np.zeros((maxp, PAGE), dtype=np.uint8)
    got = oracle().cjo_encode_varchar(n, offs.ctypes.data, C.addressof(buf), valid.ctypes.data, out.ctypes.data, maxp)
    assert got != C.c_size_t(-1).value
    return out[:got].copy()


def column_from_values(type_name: str, values: list) -> Column:
    t = T[type_name]
    if t == abi.CJ_VARCHAR:
        return Column(t, encode_varchar(values))
    valid = [v is not None for v in values]
    vals = [v if v is not None else 0 for v in values]
    return Column(t, encode_fixed(t, vals, valid))


def table_from_rows(types: list[str], rows: list[list]) -> Table:
    cols = [column_from_values(tn, [r[i] for r in rows]) for i, tn in enumerate(types)]
    return Table(len(rows), cols)


# ------------------------------------------------------------ decoding
def decode_pages(type_id: int, page_ptrs, num_pages: int) -> list:
    """Decode a column (list of page pointers) into Python values (None = NULL)."""
    lib = oracle()
    arr = (C.c_void_p * max(num_pages, 1))(*[page_ptrs[i] for i in range(num_pages)])
    if type_id == abi.CJ_VARCHAR:
        total = C.c_uint64(0)
        n = lib.cjo_decode_varchar(arr, num_pages, None, None, 0, None, C.byref(total))
        assert n >= 0, "malformed VARCHAR column"
        offs = np.zeros(n + 1, dtype=np.uint64)
        data = np.zeros(int(total.value) + 1, dtype=np.uint8)
        valid = np.zeros(max(n, 1), dtype=np.uint8)
        lib.cjo_decode_varchar(arr, num_pages, offs.ctypes.data, data.ctypes.data, total.value, valid.ctypes.data, None)
        raw = data.tobytes()
        return [raw[int(offs[i]):int(offs[i + 1])] if valid[i] else None for i in range(n)]
    n = lib.cjo_decode_fixed(type_id, arr, num_pages, None, None)
    assert n >= 0
    dt = {abi.CJ_INT32: np.int32, abi.CJ_INT64: np.int64, abi.CJ_FP64: np.float64}[type_id]
    vals = np.zeros(max(n, 1), dtype=dt)
    valid = np.zeros(max(n, 1), dtype=np.uint8)
    lib.cjo_decode_fixed(type_id, arr, num_pages, vals.ctypes.data, valid.ctypes.data)
    out = vals[:n].tolist()
    return [out[i] if valid[i] else None for i in range(n)]


def decode_column_arrays(type_id: int, page_ptrs, num_pages: int):
    """Fixed-width column -> (values ndarray, valid ndarray)."""
    lib = oracle()
    arr = (C.c_void_p * max(num_pages, 1))(*[page_ptrs[i] for i in range(num_pages)])
    n = lib.cjo_decode_fixed(type_id, arr, num_pages, None, None)
    dt = {abi.CJ_INT32: np.int32, abi.CJ_INT64: np.int64, abi.CJ_FP64: np.float64}[type_id]
    vals = np.zeros(max(n, 1), dtype=dt)
    valid = np.zeros(max(n, 1), dtype=np.uint8)
    lib.cjo_decode_fixed(type_id, arr, num_pages, vals.ctypes.data, valid.ctypes.data)
    return vals[:n], valid[:n].astype(bool)


def canon_key(row):
    out = []
    for v in row:
        if v is None:
            out.append((1, 0))
        elif isinstance(v, float):
            out.append((0, struct.unpack("<q", struct.pack("<d", v))[0]))
        elif isinstance(v, str):
            out.append((0, v.encode()))
        else:
            out.append((0, v))
    return tuple(out)


def canon_rows(rows):
    return sorted([tuple(r) for r in rows], key=canon_key)


# -------------------------------------------------------------- results
# A "result" is anything with num_rows, types and page_lists (per-column
# lists of host page addresses): an oracle run or an EngineResult.
def result_columns(r):
    return [decode_pages(t, pl, len(pl)) for t, pl in zip(r.types, r.page_lists)]


def result_rows(r):
    cols = result_columns(r)
    for c in cols:
        assert len(c) == r.num_rows, (len(c), r.num_rows)
    return [tuple(c[i] for c in cols) for i in range(r.num_rows)]


def result_fingerprint(r):
    """Order-independent fingerprint of a result's row multiset: row count,
    and the wrapping sums of the row hashes and of their squares (for very
    large results, where sorting every row hash would dominate the test)."""
    validate_result(r)
    keep = []
    desc = result_table_desc(r, keep)
    out = np.zeros(max(r.num_rows, 1), dtype=np.uint64)
    rc = oracle().cjo_row_hashes(C.byref(desc), out.ctypes.data)
    assert rc == 0, oracle().cjo_last_error()
    h = out[: r.num_rows]
    with np.errstate(over="ignore"):
        return (int(r.num_rows), int(h.sum(dtype=np.uint64)), int((h * h).sum(dtype=np.uint64)))


def result_table_desc(r, keep: list):
    cols = (abi.ColumnDesc * max(len(r.types), 1))()
    for i, (t, pl) in enumerate(zip(r.types, r.page_lists)):
        arr = (C.c_void_p * max(len(pl), 1))(*pl)
        keep.append(arr)
        cols[i] = abi.ColumnDesc(t, len(pl), C.cast(arr, C.POINTER(C.c_void_p)))
    keep.append(cols)
    return abi.TableDesc(r.num_rows, len(r.types), cols)


_POP8 = np.array([bin(i).count("1") for i in range(256)], dtype=np.int32)


def _column_pages(pl) -> np.ndarray:
    """A column's pages (list of host addresses) as an (n, 8192) uint8 array
    (a view when they are contiguous)."""
    n = len(pl)
    if n == 0:
        return np.zeros((0, PAGE), dtype=np.uint8)
    if all(pl[i + 1] - pl[i] == PAGE for i in range(n - 1)):
        return np.ctypeslib.as_array((C.c_uint8 * (n * PAGE)).from_address(pl[0])).reshape(n, PAGE)
    out = np.empty((n, PAGE), dtype=np.uint8)
    for i, a in enumerate(pl):
        out[i] = np.ctypeslib.as_array((C.c_uint8 * PAGE).from_address(a))
    return out


def validate_pages(type_id: int, pages: np.ndarray, where: str = "") -> int:
    """Structural check of a column's pages against the page format
    (reference README.md:65-104): a value-page holds 1..65535 rows;
    n_vals equals the number of set bits among the page's n_rows bitmap bits;
    values (or VARCHAR offsets and chars) plus the bitmap fit in 8192 bytes;
    VARCHAR end offsets never decrease; 0xffff / 0xfffe long-string pages
    only in VARCHAR columns, 0xfffe only after 0xffff or 0xfffe, each with
    1..8188 chars.  Returns the number of rows the pages describe."""
    n = pages[:, 0].astype(np.int64) | (pages[:, 1].astype(np.int64) << 8)
    nv = pages[:, 2].astype(np.int64) | (pages[:, 3].astype(np.int64) << 8)
    long_pg = n >= 0xfffe
    if type_id != abi.CJ_VARCHAR:
        assert not long_pg.any(), f"{where}: long-string page in a fixed-width column"
    if long_pg.any():
        idx = np.nonzero(n == 0xfffe)[0]
        assert (idx > 0).all() and long_pg[idx - 1].all(), f"{where}: 0xfffe page without a 0xffff/0xfffe before it"
        assert ((nv[long_pg] >= 1) & (nv[long_pg] <= PAGE - 4)).all(), f"{where}: long-string page char count"
    vp = ~long_pg
    assert (n[vp] >= 1).all(), f"{where}: page without rows"
    assert (nv[vp] <= n[vp]).all(), f"{where}: more values than rows"
    nb = (n + 7) // 8
    w = {abi.CJ_INT32: 4, abi.CJ_INT64: 8, abi.CJ_FP64: 8}.get(type_id, 0)
    if type_id != abi.CJ_VARCHAR:
        head = 4 if w == 4 else 8
        assert (head + w * nv[vp] + nb[vp] <= PAGE).all(), f"{where}: values and bitmap overflow the page"
    # n_vals == popcount of the first n_rows bitmap bits (the bitmap is the
    # page's last ceil(n_rows / 8) bytes, row j at byte PAGE - nb + j / 8)
    for a in range(0, len(n), 2048):
        sl = slice(a, min(a + 2048, len(n)))
        pg, nn, bb, ok = pages[sl], n[sl], nb[sl], vp[sl]
        # only the page tails that hold bitmaps (the widest in this chunk)
        span = int(bb[ok].max()) if ok.any() else 1
        tail = pg[:, PAGE - span:]
        inside = np.arange(PAGE - span, PAGE)[None, :] >= (PAGE - bb)[:, None]
        cnt = (_POP8[tail] * inside).sum(axis=1)
        r = nn % 8
        extra = np.where(r > 0, _POP8[(pg[:, PAGE - 1].astype(np.int64) >> r) & 0xFF], 0)
        cnt = cnt - extra
        bad = ok & (cnt != nv[sl])
        assert not bad.any(), f"{where}: n_vals != bitmap popcount on page {a + int(np.argmax(bad))}"
    if type_id == abi.CJ_VARCHAR:
        for i in np.nonzero(vp)[0]:
            k = int(nv[i])
            offs = pages[i, 4:4 + 2 * k].view(np.uint16).astype(np.int64) if k else np.zeros(0, np.int64)
            assert (np.diff(offs) >= 0).all() if k > 1 else True, f"{where}: VARCHAR offsets decrease on page {i}"
            end = 4 + 2 * k + (int(offs[-1]) if k else 0)
            assert end + int(nb[i]) <= PAGE, f"{where}: VARCHAR chars and bitmap overflow page {i}"
    return int(n[vp].sum() + (n == 0xffff).sum())


def validate_result(r) -> None:
    """validate_pages on every column of a result; each column describes
    exactly num_rows rows."""
    for i, (t, pl) in enumerate(zip(r.types, r.page_lists)):
        rows = validate_pages(t, _column_pages(pl), where=f"column {i}")
        assert rows == r.num_rows, (i, rows, r.num_rows)


def result_row_hashes(r) -> np.ndarray:
    """Sorted per-row hashes: equal arrays <=> equal row multisets (w.h.p.).
    The pages are checked structurally first (validate_result)."""
    validate_result(r)
    keep = []
    desc = result_table_desc(r, keep)
    out = np.zeros(max(r.num_rows, 1), dtype=np.uint64)
    rc = oracle().cjo_row_hashes(C.byref(desc), out.ctypes.data)
    assert rc == 0, oracle().cjo_last_error()
    return np.sort(out[: r.num_rows])


# ---------------------------------------------------------- KAT loading
def kat_tables(case) -> list[Table]:
    out = []
    for t in case["tables"]:
        types = t["types"]
        if "rows" in t:
            rows = t["rows"]
        elif "seq" in t:
            a, b, s = t["seq"]
            rows = [[i] for i in range(a, b, s)]
        else:
            m = t["mod"]
            rows = []
            for i in range(m["n"]):
                if m["null_every"] and i % m["null_every"] == 0:
                    rows.append([None])
                else:
                    rows.append([(i % m["mod"]) + m["add"]])
        out.append(table_from_rows(types, rows))
    return out


def kat_plan(case) -> Plan:
    p = Plan()
    for t in kat_tables(case):
        p.new_input(t)
    for n in case["nodes"]:
        attrs = [(a, t) for a, t in n["out"]]
        if "scan" in n:
            p.new_scan_node(n["scan"], attrs)
        else:
            bl, l, r, la, ra = n["join"]
            p.new_join_node(bl, l, r, la, ra, attrs)
    p.root = case["root"]
    return p


def check_kat(case, result):
    """Assert a result against a KAT's expectations."""
    exp = case["expect"]
    assert [abi.TYPE_IDS[t] for t in result.types] == exp["types"], result.types
    validate_result(result)
    if "num_rows" in exp:
        assert result.num_rows == exp["num_rows"]
    if exp.get("num_rows_positive"):
        assert result.num_rows > 0
    if result.num_rows == 0:
        return
    rows = result_rows(result)
    norm = [tuple(v.decode() if isinstance(v, bytes) else v for v in r) for r in rows]
    if "rows" in exp:
        assert canon_rows(norm) == canon_rows([tuple(r) for r in exp["rows"]])
    if "seq_rows" in exp:
        a, b, k = exp["seq_rows"]
        assert canon_rows(norm) == [tuple([v] * k) for v in range(a, b)]
    if exp.get("no_nulls"):
        assert all(v is not None for r in norm for v in r)
    if exp.get("cols_equal"):
        assert all(len(set(r)) == 1 for r in norm)
    if "first_col_mod" in exp:
        m, rr = exp["first_col_mod"]
        assert all(r[0] % m == rr for r in norm)
