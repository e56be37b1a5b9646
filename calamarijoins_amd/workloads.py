"""Synthetic workloads (include/cj_workload.h) as host Tables and Plans.

Configurations (decimal M; seed = 1000 + config index):
  C1  100k build x 1M probe, INT32 keys U[0, 2^17) with 5 % NULL, one full-range
      INT32 payload per side; output [build key, build payload, probe payload].
  C2  4M build keys = seeded permutation of [0, 4M); 64M probe keys =
      perm(rank), rank ~ Zipf(1.0), 30 % misses in [4M, 2^31); payloads INT32 +
      INT64 per side and FP64 (U[0, 1e6)) on the probe side, 2 % NULL;
      output [b.i32, b.i64, p.i32, p.i64, p.f64].
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .host import PAGE, Column, Plan, Table

GEN_PERM, GEN_UNIFORM, GEN_ZIPF_FK, GEN_I32_FULL, GEN_I64_FULL, GEN_F64_SCALED, GEN_VARCHAR, GEN_UNIFORM_FK, \
    GEN_VARCHAR_VOCAB = range(9)


class ColGen(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("type", C.c_int32), ("stream", C.c_uint64), ("null_frac", C.c_double),
        ("domain", C.c_uint64), ("perm_stream", C.c_uint64), ("zipf_s", C.c_double), ("miss_frac", C.c_double),
        ("lo", C.c_int64), ("hi", C.c_int64), ("miss_lo", C.c_int64), ("miss_hi", C.c_int64),
        ("scale", C.c_double), ("long_frac", C.c_double), ("long_lo", C.c_int64), ("long_hi", C.c_int64),
        ("vocab", C.c_char_p), ("vocab_bytes", C.c_uint64),
    ]


class GenTable(C.Structure):
    _fields_ = [
        ("num_rows", C.c_uint64), ("num_columns", C.c_uint32), ("types", C.POINTER(C.c_int32)),
        ("num_pages", C.POINTER(C.c_uint32)), ("pages", C.POINTER(C.POINTER(C.c_uint8))), ("opaque", C.c_void_p),
    ]


def _lib():
    lib = abi.load_library()
    if not getattr(lib, "_gen_sigs", False):
        lib.cj_gen_table_create.restype = C.c_int
        lib.cj_gen_table_create.argtypes = [C.c_uint64, C.c_double, C.c_uint64, C.c_uint32, C.POINTER(ColGen),
                                            C.c_int32, C.POINTER(GenTable)]
        lib.cj_gen_table_free.restype = None
        lib.cj_gen_table_free.argtypes = [C.POINTER(GenTable)]
        lib.cj_pages_encode_fixed.restype = C.c_int
        lib.cj_pages_encode_fixed.argtypes = [C.c_int32, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64,
                                              C.POINTER(C.c_uint64)]
        lib.cj_pages_encode_varchar.restype = C.c_int
        lib.cj_pages_encode_varchar.argtypes = [C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                                C.c_uint64, C.POINTER(C.c_uint64)]
        lib._gen_sigs = True
    return lib


class _Owned:
    """Keeps a generated table's library-owned pages alive."""

    def __init__(self, gt: GenTable):
        self.gt = gt

    def __del__(self):
        try:
            _lib().cj_gen_table_free(C.byref(self.gt))
        except Exception:
            pass


def col(kind, type_name, stream, null_frac=0.0, **kw) -> ColGen:
    """A column generator; for GEN_VARCHAR_VOCAB pass words=[str, ...]."""
    g = ColGen()
    g.kind = kind
    g.type = abi.TYPE_NAMES[type_name]
    g.stream = stream
    g.null_frac = null_frac
    words = kw.pop("words", None)
    if words is not None:
        blob = b"".join(w.encode("utf-8") + b"\0" for w in words)
        g._blob = blob  # kept alive with the generator
        g.vocab, g.vocab_bytes = blob, len(blob)
    for k, v in kw.items():
        setattr(g, k, v)
    return g


def generate(candidate_rows: int, seed: int, cols: list, keep_frac: float = 1.0, threads: int = 0) -> Table:
    arr = (ColGen * len(cols))(*cols)
    gt = GenTable()
    rc = _lib().cj_gen_table_create(candidate_rows, keep_frac, seed, len(cols), arr, threads, C.byref(gt))
    if rc != 0:
        msg = _lib().cj_last_error()
        raise RuntimeError(f"table generation failed: {msg.decode() if msg else rc}")
    owner = _Owned(gt)
    columns = []
    for i in range(gt.num_columns):
        np_ = gt.num_pages[i]
        buf = np.ctypeslib.as_array(gt.pages[i], shape=(max(np_, 1) * PAGE,))[: np_ * PAGE]
        columns.append(Column(gt.types[i], buf.reshape(np_, PAGE), owner=owner))
    return Table(gt.num_rows, columns, owner=owner)


def encode_fixed(type_id: int, values: np.ndarray, valid: np.ndarray) -> np.ndarray:
    """Product page writer (cj_pages_encode_fixed)."""
    n = len(valid)
    per = 1984 if type_id == abi.CJ_INT32 else 1007
    maxp = n // per + n // 60000 + 4
    out = np.zeros((maxp, PAGE), dtype=np.uint8)
    got = C.c_uint64()
    v = np.ascontiguousarray(values)
    ok = np.ascontiguousarray(valid, dtype=np.uint8)
    rc = _lib().cj_pages_encode_fixed(type_id, n, v.ctypes.data, ok.ctypes.data, out.ctypes.data, maxp, C.byref(got))
    if rc != 0:
        raise RuntimeError("encode failed")
    return out[: got.value].copy()


# ------------------------------------------------------------------ configs
def c1_tables(seed: int = 1001, build_rows: int = 100_000, probe_rows: int = 1_000_000):
    b = generate(build_rows, seed, [
        col(GEN_UNIFORM, "INT32", 1, 0.05, lo=0, hi=1 << 17),
        col(GEN_I32_FULL, "INT32", 2),
    ])
    p = generate(probe_rows, seed, [
        col(GEN_UNIFORM, "INT32", 11, 0.05, lo=0, hi=1 << 17),
        col(GEN_I32_FULL, "INT32", 12),
    ])
    return b, p


def c1_plan(build: Table, probe: Table) -> Plan:
    p = Plan()
    p.new_input(build)
    p.new_input(probe)
    p.new_scan_node(0, [(0, "INT32"), (1, "INT32")])
    p.new_scan_node(1, [(0, "INT32"), (1, "INT32")])
    p.root = p.new_join_node(True, 0, 1, 0, 0, [(0, "INT32"), (1, "INT32"), (3, "INT32")])
    return p


C2_BUILD, C2_PROBE = 4_000_000, 64_000_000


def c2_tables(seed: int = 1002, build_rows: int = C2_BUILD, probe_rows: int = C2_PROBE, threads: int = 0):
    perm_stream = 7
    b = generate(build_rows, seed, [
        col(GEN_PERM, "INT32", 1, 0.0, domain=build_rows, perm_stream=perm_stream),
        col(GEN_I32_FULL, "INT32", 2, 0.02),
        col(GEN_I64_FULL, "INT64", 3, 0.02),
    ], threads=threads)
    p = generate(probe_rows, seed, [
        col(GEN_ZIPF_FK, "INT32", 11, 0.0, domain=build_rows, perm_stream=perm_stream, zipf_s=1.0, miss_frac=0.30,
            miss_lo=build_rows, miss_hi=1 << 31),
        col(GEN_I32_FULL, "INT32", 12, 0.02),
        col(GEN_I64_FULL, "INT64", 13, 0.02),
        col(GEN_F64_SCALED, "FP64", 14, 0.02, scale=1e6),
    ], threads=threads)
    return b, p


C3_FACT = 36_000_000
C3_DIMS = (2_500_000, 4_000_000, 130_000, 1_200_000)


def _unit(seed: int, stream: int) -> float:
    """A seeded U[0, 1) draw (SplitMix64 of seed and stream)."""
    def mix(z):
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9 & 0xFFFFFFFFFFFFFFFF
        z = (z ^ (z >> 27)) * 0x94D049BB133111EB & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)
    phi = 0x9E3779B97F4A7C15
    return (mix((seed ^ (stream * phi & 0xFFFFFFFFFFFFFFFF)) + phi & 0xFFFFFFFFFFFFFFFF) >> 11) / float(1 << 53)


def c3_tables(seed: int = 1003, fact_rows: int = C3_FACT, dims=C3_DIMS, threads: int = 0):
    """Star schema: a fact table with one INT32 foreign key per dimension
    (uniform over the dimension's unfiltered key domain) and four dimensions
    with unique shuffled INT32 keys, pre-filtered to a seeded 1-40 % of their
    rows, each with a VARCHAR payload (length U[5, 60] printable ASCII; one
    string in 10^5 is 9-20 KiB and goes to long-string pages)."""
    fact = generate(fact_rows, seed, [
        col(GEN_UNIFORM, "INT32", 20 + j, 0.0, lo=0, hi=d) for j, d in enumerate(dims)
    ], threads=threads)
    out = []
    keep = []
    for j, d in enumerate(dims):
        p = 0.01 + 0.39 * _unit(seed, 900 + j)
        keep.append(p)
        out.append(generate(d, seed + 100 * (j + 1), [
            col(GEN_PERM, "INT32", 1, 0.0, domain=d, perm_stream=50 + j),
            col(GEN_VARCHAR, "VARCHAR", 2, 0.0, lo=5, hi=60, long_frac=1e-5, long_lo=9216, long_hi=20480),
        ], keep_frac=p, threads=threads))
    return fact, out, keep


def c3_plan(fact: Table, dims) -> Plan:
    """(((F x D1) x D2) x D3) x D4, build_left = false, output
    [D1.s, D2.s, D3.s, F.fk1, F.fk4]."""
    p = Plan()
    p.new_input(fact)
    for d in dims:
        p.new_input(d)
    f = p.new_scan_node(0, [(0, "INT32"), (1, "INT32"), (2, "INT32"), (3, "INT32")])
    ds = [p.new_scan_node(1 + j, [(0, "INT32"), (1, "VARCHAR")]) for j in range(4)]
    j1 = p.new_join_node(False, f, ds[0], 0, 0,
                         [(0, "INT32"), (1, "INT32"), (2, "INT32"), (3, "INT32"), (5, "VARCHAR")])
    j2 = p.new_join_node(False, j1, ds[1], 1, 0,
                         [(0, "INT32"), (2, "INT32"), (3, "INT32"), (4, "VARCHAR"), (6, "VARCHAR")])
    j3 = p.new_join_node(False, j2, ds[2], 1, 0,
                         [(0, "INT32"), (2, "INT32"), (3, "VARCHAR"), (4, "VARCHAR"), (6, "VARCHAR")])
    p.root = p.new_join_node(False, j3, ds[3], 1, 0,
                             [(2, "VARCHAR"), (3, "VARCHAR"), (4, "VARCHAR"), (0, "INT32"), (1, "INT32")])
    return p


def c2_plan(build: Table, probe: Table) -> Plan:
    p = Plan()
    p.new_input(build)
    p.new_input(probe)
    p.new_scan_node(0, [(0, "INT32"), (1, "INT32"), (2, "INT64")])
    p.new_scan_node(1, [(0, "INT32"), (1, "INT32"), (2, "INT64"), (3, "FP64")])
    # output: b.i32, b.i64, p.i32, p.i64, p.f64
    p.root = p.new_join_node(True, 0, 1, 0, 0,
                             [(1, "INT32"), (2, "INT64"), (4, "INT32"), (5, "INT64"), (6, "FP64")])
    return p


# ---------------------------------------------------------------- C5 stream
C5_TABLES = 12


def c5_tables(seed: int = 1005, scale: float = 1.0, threads: int = 0, keep_lo: float = 0.01, keep_hi: float = 0.40):
    """Twelve tables with candidate sizes log-uniform in [0.1M, 36M] (times
    `scale`), each pre-filtered to a seeded keep_lo-keep_hi (1-40 %) of its rows.  Columns:
    0 primary key (unique shuffled INT32 over the candidate domain), 1-2
    foreign keys into two other tables (Zipf(1.0) over the target's key
    ranks, so the join keeps the target's filter fraction), 3 INT64, 4 FP64,
    5 VARCHAR (5-60 printable ASCII).  Returns (tables, meta) where meta[t] =
    {"rows", "domain", "keep", "refs": (target of fk1, target of fk2)}."""
    sizes, keeps, refs = [], [], []
    for t in range(C5_TABLES):
        u = _unit(seed, 1000 + t)
        sizes.append(max(1000, int(round(1e5 * (360.0 ** u) * scale))))
        keeps.append(keep_lo + (keep_hi - keep_lo) * _unit(seed, 2000 + t))
        a = int(_unit(seed, 3000 + t) * (C5_TABLES - 1))
        b = int(_unit(seed, 4000 + t) * (C5_TABLES - 2))
        ra = (t + 1 + a) % C5_TABLES
        others = [x for x in range(C5_TABLES) if x not in (t, ra)]
        refs.append((ra, others[b % len(others)]))
    tables, meta = [], []
    for t in range(C5_TABLES):
        cols = [col(GEN_PERM, "INT32", 1, 0.0, domain=sizes[t], perm_stream=100 + t)]
        for k, r in enumerate(refs[t]):
            cols.append(col(GEN_ZIPF_FK, "INT32", 2 + k, 0.0, domain=sizes[r], perm_stream=100 + r, zipf_s=1.0,
                            miss_frac=0.0, miss_lo=0, miss_hi=1))
        cols += [col(GEN_I64_FULL, "INT64", 5, 0.02), col(GEN_F64_SCALED, "FP64", 6, 0.02, scale=1e6),
                 col(GEN_VARCHAR, "VARCHAR", 7, 0.02, lo=5, hi=60)]
        tb = generate(sizes[t], seed + 17 * t, cols, keep_frac=keeps[t], threads=threads)
        tables.append(tb)
        meta.append({"rows": tb.num_rows, "domain": sizes[t], "keep": keeps[t], "refs": refs[t]})
    return tables, meta


def c5_plan(tables, meta, plan_seed: int) -> Plan:
    """A random left-deep plan of 3-8 key / foreign-key joins over the C5
    tables: start at a random table, then repeatedly join a table that a
    foreign key already in the intermediate references (not yet joined).
    The intermediate keeps every key column; the root outputs up to six
    payload and key columns of the joined tables."""
    njoins = 3 + int(_unit(plan_seed, 1) * 6)
    start = int(_unit(plan_seed, 2) * C5_TABLES)
    p = Plan()
    for tb in tables:
        p.new_input(tb)
    key_attrs = [(0, "INT32"), (1, "INT32"), (2, "INT32")]
    pay = [(3, "INT64"), (4, "FP64"), (5, "VARCHAR")]
    # intermediate layout: per joined table, its three keys then one payload
    def scan(t):
        pc = pay[int(_unit(plan_seed, 100 + t) * 3)]
        return p.new_scan_node(t, key_attrs + [pc]), pc[1]

    node, ptype = scan(start)
    layout = [(start, "pk", "INT32"), (start, "fk1", "INT32"), (start, "fk2", "INT32"), (start, "pay", ptype)]
    joined = [start]
    for j in range(njoins):
        cands = []
        for i, (t, what, _) in enumerate(layout):
            if what in ("fk1", "fk2"):
                target = meta[t]["refs"][0 if what == "fk1" else 1]
                if target not in joined:
                    cands.append((i, target))
        if not cands:
            break
        i, target = cands[int(_unit(plan_seed, 200 + j) * len(cands))]
        right, rtype = scan(target)
        rlay = [(target, "pk", "INT32"), (target, "fk1", "INT32"), (target, "fk2", "INT32"), (target, "pay", rtype)]
        both = layout + rlay
        last = j == njoins - 1 or len(joined) + 1 == C5_TABLES
        if last:
            pays = [k for k, x in enumerate(both) if x[1] == "pay"][-5:]
            outs = [(k, both[k][2]) for k in pays] + [(i, "INT32")]
            p.root = p.new_join_node(False, node, right, i, 0, outs)
            return p
        node = p.new_join_node(False, node, right, i, 0, [(k, x[2]) for k, x in enumerate(both)])
        layout = both
        joined.append(target)
    # ran out of joinable tables: close with the last join's columns
    pays = [k for k, x in enumerate(layout) if x[1] == "pay"][-5:]
    if p.nodes[-1]["kind"] == abi.CJ_NODE_JOIN:
        p.nodes[-1]["attrs"] = [(k, layout[k][2]) for k in pays] + [(0, "INT32")]
        p.root = len(p.nodes) - 1
        return p
    raise ValueError("plan with no join")


# ------------------------------------------------------------- C4 batch
def c4_plan_tables(k: int, scale: float = 1.0, threads: int = 0):
    """Plan k of the batch (seed 2000 + k): build rows log-uniform in
    [1e3, 2e6], probe rows log-uniform in [1e4, 2e7] (times scale), key domain
    [0, build rows).  Build keys are uniform over the domain (repeats are
    allowed, so matches fan out); probe keys are uniform or Zipf(0.8) over
    the domain by a per-plan coin.  Each side has a NULL rate U[0, 0.10] on
    its key and one full-range INT32 payload.  (Zipf on both sides would make
    a hot key's output quadratic in its frequency.)"""
    seed = 2000 + k
    nb = max(8, int(round(1e3 * (2000.0 ** _unit(seed, 1)) * scale)))
    npr = max(8, int(round(1e4 * (2000.0 ** _unit(seed, 2)) * scale)))
    zipf = _unit(seed, 3) < 0.5
    nb_null, np_null = 0.10 * _unit(seed, 4), 0.10 * _unit(seed, 5)
    b = generate(nb, seed, [col(GEN_UNIFORM, "INT32", 1, nb_null, lo=0, hi=nb), col(GEN_I32_FULL, "INT32", 2)],
                 threads=threads)
    if zipf:
        pk = col(GEN_ZIPF_FK, "INT32", 11, np_null, domain=nb, perm_stream=60, zipf_s=0.8, miss_frac=0.0,
                 miss_lo=0, miss_hi=1)
    else:
        pk = col(GEN_UNIFORM, "INT32", 11, np_null, lo=0, hi=nb)
    p = generate(npr, seed + 1, [pk, col(GEN_I32_FULL, "INT32", 12)], threads=threads)
    return b, p


def c4_plan(build: Table, probe: Table) -> Plan:
    """[build key, build payload, probe payload]."""
    p = Plan()
    p.new_input(build)
    p.new_input(probe)
    p.new_scan_node(0, [(0, "INT32"), (1, "INT32")])
    p.new_scan_node(1, [(0, "INT32"), (1, "INT32")])
    p.root = p.new_join_node(True, 0, 1, 0, 0, [(0, "INT32"), (1, "INT32"), (3, "INT32")])
    return p
