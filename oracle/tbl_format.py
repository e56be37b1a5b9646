"""Oracle for the .tbl table cache format.  TEST INFRASTRUCTURE ONLY: used by
tests/, never by the product library.

A numpy restatement of the reference's cache writer and reader (citations
are to files of the reference repository):
  - header TableMeta {u64 num_rows; u64 num_cols; DataType types[16];
    u64 num_pages[16]} padded with zeros to one 8192-byte page
    .......................................... include/data_access/table.h:30-37
  - DumpTable::dump: header, filler, then each column's pages in order
    .......................................... include/data_access/table.h:208-221
  - Table::from_cache: header at offset 0, pages from offset 8192, column
    after column ............................. src/build_table.cpp:140-197
  - FNV-1a 64 of the filter text for cache names
    .......................................... include/data_access/table.h:15-27,
                                               tests/read_sql.cpp:1154-1158

Pinned in tests/test_tbl.py: the header layout against a ctypes mirror of
the C struct (DataType is a 4-byte enum: size 208, num_pages at byte 80),
FNV-1a 64 against the published test vectors, and round trips in both
directions against the product's reader and writer.
"""
from __future__ import annotations

import struct

import numpy as np

PAGE = 8192
HEADER = struct.Struct("<QQ16i16Q")  # TableMeta, little endian, no padding needed


def fnv1a64(data: bytes) -> int:
    h = 14695981039346656037
    for b in data:
        h ^= b
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def cache_name(table: str, filter_text: str | None) -> str:
    text = filter_text if filter_text is not None else "No filter"
    return f"{table}_{fnv1a64(text.encode())}.tbl"


def dump(num_rows: int, columns: list) -> bytes:
    """columns: list of (type_id, pages) with pages a (n, 8192) uint8 array."""
    assert len(columns) <= 16
    types = [t for t, _ in columns] + [0] * (16 - len(columns))
    npages = [int(p.shape[0]) for _, p in columns] + [0] * (16 - len(columns))
    head = HEADER.pack(num_rows, len(columns), *types, *npages)
    parts = [head, bytes(PAGE - len(head))]
    for _, p in columns:
        parts.append(np.ascontiguousarray(p, dtype=np.uint8).tobytes())
    return b"".join(parts)


def parse(data: bytes):
    """-> (num_rows, [(type_id, pages (n, 8192) uint8)])."""
    f = HEADER.unpack_from(data, 0)
    num_rows, num_cols = f[0], f[1]
    types, npages = f[2:18], f[18:34]
    out, off = [], PAGE
    for c in range(num_cols):
        n = npages[c]
        pages = np.frombuffer(data, dtype=np.uint8, count=n * PAGE, offset=off).reshape(n, PAGE)
        out.append((types[c], pages))
        off += n * PAGE
    return num_rows, out
