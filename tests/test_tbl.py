"""The .tbl table cache (cj_tbl_write / cj_tbl_open / cj_tbl_cache_name)
against the oracle restatement of the reference's DumpTable / from_cache
(oracle/tbl_format.py).  Host only: no GPU."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

import cjtest
from calamarijoins_amd import abi, host

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import tbl_format  # noqa: E402


def sample_table():
    rows = []
    for i in range(5000):
        rows.append([
            None if i % 7 == 0 else i * 3 - 2500,
            None if i % 11 == 0 else (i << 33) - 5,
            None if i % 13 == 0 else i / 7.0,
            None if i % 17 == 0 else ("x" * (i % 40) if i != 4321 else "L" * 20000),
        ])
    return cjtest.table_from_rows(["INT32", "INT64", "FP64", "VARCHAR"], rows)


def test_header_layout_matches_the_c_struct():
    class TableMeta(C.Structure):  # DataType is an enum class with int underlying type
        _fields_ = [("num_rows", C.c_uint64), ("num_cols", C.c_uint64), ("types", C.c_int * 16),
                    ("num_pages", C.c_uint64 * 16)]
    assert C.sizeof(TableMeta) == tbl_format.HEADER.size == 208
    assert TableMeta.num_pages.offset == 80


def test_fnv1a64_published_vectors():
    assert tbl_format.fnv1a64(b"") == 0xCBF29CE484222325
    assert tbl_format.fnv1a64(b"a") == 0xAF63DC4C8601EC8C
    assert tbl_format.fnv1a64(b"foobar") == 0x85944171F73967E8


def test_cache_name_matches_the_harness_rule():
    assert host.tbl_cache_name("title") == tbl_format.cache_name("title", None)
    f = "(t.production_year > 2000 AND t.kind_id = 1)"
    assert host.tbl_cache_name("title", f) == tbl_format.cache_name("title", f)
    assert host.tbl_cache_name("title", f).startswith("title_") and host.tbl_cache_name("title", f).endswith(".tbl")


def test_product_writer_bytes_equal_the_oracle_dump(tmp_path):
    t = sample_table()
    path = tmp_path / "t.tbl"
    host.write_tbl(str(path), t)
    data = path.read_bytes()
    exp = tbl_format.dump(t.num_rows, [(c.type, c.pages) for c in t.columns])
    assert data == exp
    assert len(data) == 8192 * (1 + sum(c.num_pages for c in t.columns))
    assert data[208:8192] == bytes(8192 - 208)


def test_product_reader_maps_the_oracle_dump(tmp_path):
    t = sample_table()
    path = tmp_path / "o.tbl"
    path.write_bytes(tbl_format.dump(t.num_rows, [(c.type, c.pages) for c in t.columns]))
    r = host.read_tbl(str(path))
    try:
        assert r.num_rows == t.num_rows and len(r.columns) == len(t.columns)
        for a, b in zip(r.columns, t.columns):
            assert a.type == b.type and a.num_pages == b.num_pages
            assert np.array_equal(a.pages, b.pages)
        # the decoded values are the rows that were encoded
        for c in range(4):
            got = cjtest.decode_pages(r.columns[c].type, r.columns[c].page_ptrs(), r.columns[c].num_pages)
            exp = cjtest.decode_pages(t.columns[c].type, t.columns[c].page_ptrs(), t.columns[c].num_pages)
            assert got == exp
    finally:
        r.close()


def test_round_trip_and_oracle_parse(tmp_path):
    t = sample_table()
    path = tmp_path / "rt.tbl"
    host.write_tbl(str(path), t)
    n, cols = tbl_format.parse(path.read_bytes())
    assert n == t.num_rows
    for (ty, pg), c in zip(cols, t.columns):
        assert ty == c.type and np.array_equal(pg, c.pages)
    r = host.read_tbl(str(path))
    host.write_tbl(str(tmp_path / "rt2.tbl"), r)  # from the mapping itself
    r.close()
    assert (tmp_path / "rt2.tbl").read_bytes() == path.read_bytes()


def test_empty_table_and_zero_page_columns(tmp_path):
    t = cjtest.table_from_rows(["INT32", "VARCHAR"], [])
    path = tmp_path / "e.tbl"
    host.write_tbl(str(path), t)
    assert path.stat().st_size == 8192 * (1 + sum(c.num_pages for c in t.columns))
    r = host.read_tbl(str(path))
    assert r.num_rows == 0 and [c.type for c in r.columns] == [abi.CJ_INT32, abi.CJ_VARCHAR]
    r.close()


@pytest.mark.parametrize("case", ["missing", "empty", "short", "size", "cols", "type"])
def test_malformed_files_are_rejected(tmp_path, case):
    t = cjtest.table_from_rows(["INT32"], [[i] for i in range(3000)])
    good = tbl_format.dump(t.num_rows, [(c.type, c.pages) for c in t.columns])
    path = tmp_path / "bad.tbl"
    if case == "empty":
        path.write_bytes(b"")
    elif case == "short":
        path.write_bytes(good[:4000])
    elif case == "size":
        path.write_bytes(good[:-8192])
    elif case == "cols":
        path.write_bytes(good[:8] + (17).to_bytes(8, "little") + good[16:])
    elif case == "type":
        path.write_bytes(good[:16] + (9).to_bytes(4, "little") + good[20:])
    with pytest.raises(abi.CJError):
        host.read_tbl(str(path))


def test_writer_rejects_more_than_16_columns(tmp_path):
    t = cjtest.table_from_rows(["INT32"] * 17, [[1] * 17])
    with pytest.raises(abi.CJError):
        host.write_tbl(str(tmp_path / "w.tbl"), t)
    assert not any(p.name.startswith("w.tbl") for p in tmp_path.iterdir())
