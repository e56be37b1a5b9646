"""CPU checks that pin the oracle (no GPU): the reference's known-answer
join scenarios, hashing primitives against published values, the page codec
against the format's capacity rules, and random plans against a pure-Python
brute force."""
import ctypes as C
import json
import os
import random

import numpy as np
import pytest

import bruteforce
import cjtest
from calamarijoins_amd import abi

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KATS = json.load(open(os.path.join(GOLDEN, "unit_kats.json")))["cases"]
HASHV = json.load(open(os.path.join(GOLDEN, "hash_vectors.json")))


@pytest.mark.parametrize("case", KATS, ids=[c["name"] for c in KATS])
def test_oracle_unit_kat(case):
    res = cjtest.OracleRun(cjtest.kat_plan(case), threads=4)
    cjtest.check_kat(case, res)


def test_crc32c_check_value():
    lib = cjtest.oracle()
    data = HASHV["crc32c_check"]["input"].encode()
    assert lib.cjo_crc32c_bytes(0xFFFFFFFF, data, len(data)) ^ 0xFFFFFFFF == HASHV["crc32c_check"]["value"]


def test_crc32c_matches_sse42_instruction():
    # The reference hashes with the SSE4.2 crc32 instruction (hashtable.h:146-154).
    bad = cjtest.oracle().cjo_crc32c_hw_mismatches(200000)
    if bad < 0:
        pytest.skip("no SSE4.2 on this host")
    assert bad == 0


def test_bloom_tags_closed_form():
    lib = cjtest.oracle()
    tags = [lib.cjo_bloom_tag(i) for i in range(2048)]
    assert tags[:9] == HASHV["bloom_first"]
    assert tags[-11:] == HASHV["bloom_last"]
    assert all(bin(t).count("1") == 5 for t in tags)
    assert all(a < b for a, b in zip(tags, tags[1:]))


# ------------------------------------------------------------------ codec
def test_fixed_page_capacity():
    n = 5000
    for t, per in ((abi.CJ_INT32, 1984), (abi.CJ_INT64, 1007), (abi.CJ_FP64, 1007)):
        pages = cjtest.encode_fixed(t, np.arange(n), np.ones(n, np.uint8))
        rows = [int(pages[i, :2].view(np.uint16)[0]) for i in range(pages.shape[0])]
        assert rows[:-1] == [per] * (len(rows) - 1)
        assert sum(rows) == n


def test_all_null_pages_stay_below_long_string_marker():
    n = 200000
    pages = cjtest.encode_fixed(abi.CJ_INT32, np.zeros(n), np.zeros(n, np.uint8))
    rows = [int(pages[i, :2].view(np.uint16)[0]) for i in range(pages.shape[0])]
    assert max(rows) < 0xfffe and sum(rows) == n


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_fixed_round_trip(seed):
    rng = np.random.default_rng(seed)
    n = 7000
    for t, dt in ((abi.CJ_INT32, np.int32), (abi.CJ_INT64, np.int64), (abi.CJ_FP64, np.float64)):
        vals = rng.integers(-2**31, 2**31, n).astype(dt) if t != abi.CJ_FP64 else rng.random(n) * 1e6
        valid = rng.random(n) > 0.3
        col = cjtest.Column(t, cjtest.encode_fixed(t, vals, valid))
        dec = cjtest.decode_pages(t, col.page_ptrs(), col.num_pages)
        exp = [v if ok else None for v, ok in zip(vals.tolist(), valid)]
        assert dec == exp


def test_varchar_round_trip_with_long_strings():
    rng = random.Random(7)
    strs = []
    for i in range(3000):
        r = rng.random()
        if r < 0.1:
            strs.append(None)
        elif r < 0.102:
            strs.append("L" * rng.choice([8185, 8186, 8188, 8189, 20000]))
        else:
            strs.append("".join(chr(rng.randrange(32, 127)) for _ in range(rng.randrange(0, 60))))
    col = cjtest.column_from_values("VARCHAR", strs)
    dec = cjtest.decode_pages(abi.CJ_VARCHAR, col.page_ptrs(), col.num_pages)
    assert dec == [s.encode() if s is not None else None for s in strs]
    # a string longer than PAGE_SIZE - 7 = 8185 bytes goes to 0xffff/0xfffe pages
    heads = [int(col.pages[i, :2].view(np.uint16)[0]) for i in range(col.num_pages)]
    assert 0xffff in heads and 0xfffe in heads


def test_long_string_page_split():
    col = cjtest.column_from_values("VARCHAR", ["x" * 20000])
    heads = [tuple(col.pages[i, :4].view(np.uint16)) for i in range(col.num_pages)]
    assert heads == [(0xffff, 8188), (0xfffe, 8188), (0xfffe, 20000 - 2 * 8188)]


# ------------------------------------------------------------ brute force
TYPE_POOL = ["INT32", "INT32", "INT64", "FP64", "VARCHAR"]


def rand_value(rng, t, key_domain):
    if rng.random() < 0.15:
        return None
    if t == "INT32":
        if rng.random() < 0.03:
            return bruteforce.I32_MIN
        return rng.randrange(key_domain)
    if t == "INT64":
        return rng.randrange(-2**63, 2**63)
    if t == "FP64":
        return rng.random() * 1e6
    if rng.random() < 0.03:
        # long strings: 0xffff / 0xfffe page runs above 8185 bytes
        return "".join(chr(rng.randrange(97, 123)) for _ in range(rng.randrange(8185, 20001)))
    return "".join(chr(rng.randrange(97, 123)) for _ in range(rng.randrange(0, 12)))


def rand_table(rng, nrows, key_domain):
    ncol = rng.randrange(1, 4)
    types = ["INT32"] + [rng.choice(TYPE_POOL) for _ in range(ncol)]
    rows = [[rand_value(rng, t, key_domain) for t in types] for _ in range(nrows)]
    return types, rows


def rand_plan(rng, njoins, tables):
    """Random left-deep plan; returns (cjtest.Plan, node dicts for bruteforce)."""
    p = cjtest.Plan()
    for types, rows in tables:
        p.new_input(cjtest.table_from_rows(types, rows))
    scans = []
    for ti, (types, rows) in enumerate(tables):
        k = rng.randrange(1, len(types) + 1)
        cols = sorted(rng.sample(range(len(types)), k))
        if 0 not in cols:
            cols = [0] + cols
        rng.shuffle(cols)
        scans.append(p.new_scan_node(ti, [(c, types[c]) for c in cols]))

    def types_of(node):
        return [t for _, t in p.nodes[node]["attrs"]]

    cur = scans[0]
    for j in range(njoins):
        right = scans[j + 1]
        lt, rt = types_of(cur), types_of(right)
        la = rng.choice([i for i, t in enumerate(lt) if t == "INT32"])
        ra = rng.choice([i for i, t in enumerate(rt) if t == "INT32"])
        allc = list(range(len(lt) + len(rt)))
        ts = lt + rt
        last = j == njoins - 1
        k = rng.randrange(1, len(allc) + 1)
        outs = rng.sample(allc, k)
        if not last and not any(ts[o] == "INT32" for o in outs):
            outs.append(la)
        swap = rng.random() < 0.5
        if swap:
            # right-deep step: intermediate on the right
            p_node = p.new_join_node(rng.random() < 0.5, right, cur, ra, la,
                                     [((o - len(lt) if o >= len(lt) else o + len(rt)), ts[o]) for o in outs])
        else:
            p_node = p.new_join_node(rng.random() < 0.5, cur, right, la, ra, [(o, ts[o]) for o in outs])
        cur = p_node
    p.root = cur
    return p


def norm_rows(rows):
    return cjtest.canon_rows([tuple(v.decode() if isinstance(v, bytes) else v for v in r) for r in rows])


@pytest.mark.parametrize("seed", range(40))
def test_oracle_matches_bruteforce(seed):
    rng = random.Random(1000 + seed)
    njoins = 1 + seed % 3
    tables = []
    for _ in range(njoins + 1):
        nrows = 0 if rng.random() < 0.04 else rng.choice([1, 3, 5, 7, 8, 9, 20, 60, 150])
        tables.append(rand_table(rng, nrows, key_domain=rng.choice([2, 4, 8, 16])))
    plan = rand_plan(rng, njoins, tables)
    exp = bruteforce.evaluate(plan.nodes, tables, plan.root)
    got = cjtest.OracleRun(plan, threads=3)
    assert got.num_rows == len(exp)
    assert norm_rows(cjtest.result_rows(got)) == norm_rows(exp)


def test_oracle_multithreaded_partitioned_build():
    # > 16 build pages and > 262,144 tuples: the parallel partition pass with
    # several radix partitions (hashtable.h:405-505).
    n = 600000
    keys = np.arange(n, dtype=np.int32)
    b = cjtest.Table(n, [cjtest.Column(abi.CJ_INT32, cjtest.encode_fixed(abi.CJ_INT32, keys, np.ones(n, np.uint8)))])
    pk = (np.arange(2 * n, dtype=np.int64) * 7919 % (n + n // 2)).astype(np.int32)
    pr = cjtest.Table(2 * n, [cjtest.Column(abi.CJ_INT32, cjtest.encode_fixed(abi.CJ_INT32, pk, np.ones(2 * n, np.uint8)))])
    p = cjtest.Plan()
    p.new_input(b)
    p.new_input(pr)
    p.new_scan_node(0, [(0, "INT32")])
    p.new_scan_node(1, [(0, "INT32")])
    p.root = p.new_join_node(True, 0, 1, 0, 0, [(0, "INT32"), (1, "INT32")])
    got = cjtest.OracleRun(p, threads=8)
    assert got.num_rows == int(np.sum(pk < n))
    cols = [cjtest.decode_column_arrays(abi.CJ_INT32, pl, len(pl)) for pl in got.page_lists]
    assert np.array_equal(cols[0][0], cols[1][0])
    assert np.array_equal(np.sort(cols[1][0]), np.sort(pk[pk < n]))


# ------------------------------------------------ the checkers themselves
def _hash_rows(types, rows):
    t = cjtest.table_from_rows(types, rows)

    class R:
        pass

    r = R()
    r.num_rows, r.types = t.num_rows, [c.type for c in t.columns]
    r.page_lists = [[c.pages.ctypes.data + i * cjtest.PAGE for i in range(c.num_pages)] for c in t.columns]
    r.keep = t
    return cjtest.result_row_hashes(r)


HASH_TYPES = ["INT32", "INT64", "FP64", "VARCHAR", "INT32"]
HASH_ROWS = [[i, -7 * i, i / 3.0, "s%d" % (i % 5) if i % 4 else None, None if i % 6 == 0 else 2 * i]
             for i in range(300)]


def test_row_hashes_ignore_row_order():
    base = _hash_rows(HASH_TYPES, HASH_ROWS)
    perm = list(HASH_ROWS)
    random.Random(5).shuffle(perm)
    assert np.array_equal(base, _hash_rows(HASH_TYPES, perm))


@pytest.mark.parametrize("change", ["value", "null", "byte", "swap_cols", "dup_row", "fp_sign", "string_to_null"])
def test_row_hashes_tell_rows_apart(change):
    """A changed value, NULL bit, string byte, a swap of two same-typed
    columns in one row, or a duplicated row changes the sorted hash array."""
    base = _hash_rows(HASH_TYPES, HASH_ROWS)
    rows = [list(r) for r in HASH_ROWS]
    if change == "value":
        rows[17][1] += 1
    elif change == "null":
        rows[19][4] = None
    elif change == "byte":
        rows[21][3] = "s2x"[:2] + "y"
    elif change == "swap_cols":
        rows[25][0], rows[25][4] = rows[25][4], rows[25][0]
    elif change == "dup_row":
        rows[0] = list(rows[1])
    elif change == "fp_sign":
        rows[0][2] = -0.0
    else:
        rows[22][3] = None
    assert change == "fp_sign" or rows != [list(r) for r in HASH_ROWS]  # -0.0 == 0.0 in Python
    assert not np.array_equal(base, _hash_rows(HASH_TYPES, rows))


def _pages_of(type_name, values):
    return cjtest.column_from_values(type_name, values).pages.copy()


def test_validator_accepts_codec_pages():
    rng = random.Random(9)
    for t in ("INT32", "INT64", "FP64"):
        vals = [None if rng.random() < 0.3 else rng.randrange(1000) for _ in range(7000)]
        pg = _pages_of(t, vals)
        assert cjtest.validate_pages(abi.TYPE_NAMES[t], pg) == len(vals)
    strs = [None if rng.random() < 0.2 else "x" * rng.randrange(0, 70) for _ in range(3000)] + ["y" * 20000, "z"]
    pg = _pages_of("VARCHAR", strs)
    assert cjtest.validate_pages(abi.CJ_VARCHAR, pg) == len(strs)


@pytest.mark.parametrize("mistake", ["n_vals", "overflow", "offsets", "lone_fffe", "long_in_fixed", "no_rows",
                                     "bitmap_bit"])
def test_validator_rejects_malformed_pages(mistake):
    """Mistakes planted in oracle-encoded pages (on the CPU) are caught."""
    if mistake in ("offsets", "lone_fffe"):
        pg = _pages_of("VARCHAR", ["abc", "de", None, "fghij"] * 50 + ["q" * 9000])
        t = abi.CJ_VARCHAR
        if mistake == "offsets":
            pg[0, 4:6] = np.array([200, 0], dtype=np.uint8)  # first end offset beyond the second
        else:
            pg[0, 0:2] = np.array([0xfe, 0xff], dtype=np.uint8)  # page 0 claims to continue a long string
    else:
        pg = _pages_of("INT32", [None if i % 3 == 0 else i for i in range(5000)])
        t = abi.CJ_INT32
        if mistake == "n_vals":
            pg[1, 2] ^= 1
        elif mistake == "overflow":
            pg[0, 0:2] = np.array([0x00, 0x20], dtype=np.uint8)  # 8192 rows: bitmap alone fills the page
            pg[0, 2:4] = np.array([0x00, 0x08], dtype=np.uint8)
        elif mistake == "long_in_fixed":
            pg[1, 0:2] = np.array([0xff, 0xff], dtype=np.uint8)
        elif mistake == "no_rows":
            pg[1, 0:4] = 0
        else:
            pg[0, cjtest.PAGE - 3] ^= 0x10  # a bitmap bit flipped: popcount no longer n_vals
    with pytest.raises(AssertionError):
        cjtest.validate_pages(t, pg)
