"""Python mirror of the contest API objects over the C ABI.

Column / Table / Plan follow the reference's Plan builder
(include/data_model/plan.h:149-185) and produce the cj_api.h descriptors;
Engine wraps a cj_ctx.  Used by the benchmark harness and the tests.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi

PAGE = abi.PAGE_SIZE
T = abi.TYPE_NAMES


class Column:
    """A column: its type id and its pages (uint8 array, n_pages x 8192)."""

    def __init__(self, type_id: int, pages: np.ndarray, owner=None):
        self.type = type_id
        self.pages = pages.reshape(-1, PAGE) if pages.size else pages.reshape(0, PAGE)
        self.owner = owner
        self._ptrs = None

    @property
    def num_pages(self) -> int:
        return self.pages.shape[0]

    def page_ptrs(self):
        if self._ptrs is None:
            base = self.pages.ctypes.data
            n = self.num_pages
            self._ptrs = (C.c_void_p * max(n, 1))(*[base + i * PAGE for i in range(n)])
        return self._ptrs


class Table:
    def __init__(self, num_rows: int, columns: list, owner=None):
        self.num_rows = num_rows
        self.columns = columns
        self.owner = owner
        self._desc = None

    def desc(self) -> abi.TableDesc:
        if self._desc is None:
            cols = (abi.ColumnDesc * max(len(self.columns), 1))()
            for i, c in enumerate(self.columns):
                cols[i] = abi.ColumnDesc(c.type, c.num_pages, C.cast(c.page_ptrs(), C.POINTER(C.c_void_p)))
            self._cols = cols
            self._desc = abi.TableDesc(self.num_rows, len(self.columns), cols)
        return self._desc

    def input_bytes(self, col_indices) -> int:
        return sum(self.columns[i].num_pages * PAGE for i in col_indices)


class TblTable(Table):
    """A .tbl cache file mapped read-only (cj_tbl_open, like the harness's
    Table::from_cache) or also streamed into HBM as a resident table
    (cj_tbl_load).  Column pages are numpy views of the mapping; close()
    releases it (unregister a loaded table first)."""

    def __init__(self, handle: C.c_void_p, desc: abi.TableDesc, table_id=None, engine=None):
        cols = []
        for c in range(desc.num_columns):
            cd = desc.columns[c]
            if cd.num_pages:
                buf = (C.c_uint8 * (cd.num_pages * PAGE)).from_address(cd.pages[0])
                arr = np.ctypeslib.as_array(buf)
            else:
                arr = np.zeros(0, dtype=np.uint8)
            cols.append(Column(cd.type, arr, owner=self))
        super().__init__(desc.num_rows, cols)
        self.handle = handle
        self.table_id = table_id
        self.engine = engine

    def close(self):
        if self.handle:
            if self.table_id is not None and self.engine is not None and self.engine.ctx:
                self.engine.unregister(self.table_id)
                self.table_id = None
            abi.check(abi.load_library().cj_tbl_close(self.handle))
            self.handle = None


def write_tbl(path: str, table: Table) -> None:
    """DumpTable::dump equivalent (cj_tbl_write)."""
    abi.check(abi.load_library().cj_tbl_write(str(path).encode(), C.byref(table.desc())))


def read_tbl(path: str) -> TblTable:
    """Table::from_cache equivalent (cj_tbl_open): pages stay in the file mapping."""
    h = C.c_void_p()
    d = abi.TableDesc()
    abi.check(abi.load_library().cj_tbl_open(str(path).encode(), C.byref(h), C.byref(d)))
    return TblTable(h, d)


def tbl_cache_name(table: str, filter_text=None) -> str:
    """The harness's cache file name of a scan: <table>_<FNV-1a 64 of the filter text>.tbl."""
    out = C.create_string_buffer(512)
    abi.check(abi.load_library().cj_tbl_cache_name(table.encode(), filter_text.encode() if filter_text else None,
                                                    out, 512))
    return out.value.decode()


class Plan:
    def __init__(self):
        self.nodes = []
        self.inputs: list = []
        self.root = 0

    def new_scan_node(self, table_id, attrs):
        self.nodes.append(dict(kind=abi.CJ_NODE_SCAN, table=table_id, attrs=list(attrs)))
        return len(self.nodes) - 1

    def new_join_node(self, build_left, left, right, left_attr, right_attr, attrs):
        self.nodes.append(dict(kind=abi.CJ_NODE_JOIN, build_left=build_left, left=left, right=right,
                               left_attr=left_attr, right_attr=right_attr, attrs=list(attrs)))
        return len(self.nodes) - 1

    def new_input(self, table):
        self.inputs.append(table)
        return len(self.inputs) - 1

    def desc(self) -> abi.PlanDesc:
        keep = []
        nodes = (abi.NodeDesc * max(len(self.nodes), 1))()
        for i, n in enumerate(self.nodes):
            attrs = (abi.Attr * max(len(n["attrs"]), 1))()
            for j, (idx, t) in enumerate(n["attrs"]):
                attrs[j] = abi.Attr(idx, T[t] if isinstance(t, str) else t)
            keep.append(attrs)
            d = abi.NodeDesc()
            d.kind = n["kind"]
            if n["kind"] == abi.CJ_NODE_SCAN:
                d.base_table_id = n["table"]
            else:
                d.build_left = int(n["build_left"])
                d.left, d.right = n["left"], n["right"]
                d.left_attr, d.right_attr = n["left_attr"], n["right_attr"]
            d.num_output_attrs = len(n["attrs"])
            d.output_attrs = attrs
            nodes[i] = d
        tabs = (abi.TableDesc * max(len(self.inputs), 1))()
        for i, t in enumerate(self.inputs):
            tabs[i] = t.desc()
        pd = abi.PlanDesc(len(self.nodes), nodes, len(self.inputs), tabs, self.root)
        self._keep = (keep, nodes, tabs)
        return pd


class Engine:
    """A cj_ctx (one GPU)."""

    def __init__(self, **opts):
        self.lib = abi.load_library()
        o = abi.CtxOptions()
        for k, v in opts.items():
            setattr(o, k, v)
        self.ctx = C.c_void_p()
        abi.check(self.lib.cj_ctx_create(C.byref(o), C.byref(self.ctx)))

    def close(self):
        if self.ctx:
            self.lib.cj_ctx_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def device_info(self):
        name = C.create_string_buffer(256)
        cus = C.c_int32()
        abi.check(self.lib.cj_ctx_device_info(self.ctx, name, 256, C.byref(cus)))
        return name.value.decode(), cus.value

    def register(self, table: Table) -> int:
        tid = C.c_uint64()
        abi.check(self.lib.cj_table_register(self.ctx, C.byref(table.desc()), C.byref(tid)))
        return tid.value

    def unregister(self, tid: int):
        abi.check(self.lib.cj_table_unregister(self.ctx, tid))

    def filter(self, table: Table, program, flags: int = 0, timing: bool = False):
        """Rows of `table` that pass `program` (a list of abi.FilterNode in
        postfix order, e.g. from job.filter_program), every column, in row
        order (cj_filter)."""
        n = len(program)
        arr = (abi.FilterNode * max(n, 1))(*program)
        res = abi.Result()
        tm = abi.Timing()
        abi.check(self.lib.cj_filter(self.ctx, C.byref(table.desc()), arr, n, flags, C.byref(res),
                                     C.byref(tm) if timing else None))
        return EngineResult(self, res, tm if timing else None)

    def load_tbl(self, path: str) -> TblTable:
        """Stream a .tbl file into HBM as a resident table (cj_tbl_load)."""
        h = C.c_void_p()
        d = abi.TableDesc()
        tid = C.c_uint64()
        abi.check(self.lib.cj_tbl_load(self.ctx, str(path).encode(), C.byref(h), C.byref(d), C.byref(tid)))
        return TblTable(h, d, tid.value, self)

    def execute(self, plan: Plan, flags: int = 0, timing: bool = False):
        res = abi.Result()
        tm = abi.Timing()
        abi.check(self.lib.cj_execute(self.ctx, C.byref(plan.desc()), flags, C.byref(res),
                                      C.byref(tm) if timing else None))
        return EngineResult(self, res, tm if timing else None)

    # ---- serving loop (cj_serve_start / cj_submit / cj_wait / cj_serve_stop)
    def serve_start(self, workers: int, admit_bytes: int):
        abi.check(self.lib.cj_serve_start(self.ctx, C.byref(abi.ServeOptions(workers, admit_bytes))))
        self._inflight = {}

    def submit(self, plan: Plan, flags: int = 0) -> int:
        t = C.c_uint64()
        abi.check(self.lib.cj_submit(self.ctx, C.byref(plan.desc()), flags, C.byref(t)))
        self._inflight[t.value] = plan  # the plan's pages must outlive the request
        return t.value

    def wait(self, ticket: int, timeout_ms: int = -1):
        """The result of a submitted plan (EngineResult with timing), or None
        when timeout_ms passes first."""
        res = abi.Result()
        tm = abi.Timing()
        rc = self.lib.cj_wait(self.ctx, ticket, timeout_ms, C.byref(res), C.byref(tm))
        if rc == abi.CJ_ERR_BUSY and timeout_ms >= 0:
            return None
        self._inflight.pop(ticket, None)
        abi.check(rc)
        return EngineResult(self, res, tm)

    def serve_stats(self) -> dict:
        st = abi.ServeStats()
        abi.check(self.lib.cj_serve_stats_get(self.ctx, C.byref(st)))
        return {k: getattr(st, k) for k, _ in abi.ServeStats._fields_}

    def memory(self, reset_peak: bool = False) -> dict:
        v = [C.c_uint64() for _ in range(4)]
        abi.check(self.lib.cj_ctx_memory(self.ctx, int(reset_peak), *[C.byref(x) for x in v]))
        return dict(zip(("reserved", "in_use", "peak_in_use", "budget"), (x.value for x in v)))

    def serve_stop(self):
        abi.check(self.lib.cj_serve_stop(self.ctx))
        self._inflight = {}

    def batch_config(self, load_factor=None, log_part=12, chunk=8192, join_queue=0, grid=0):
        """Pin the batch configuration (cj_batch_config_set); with no
        load_factor, unpin it (autotuner / default again)."""
        if load_factor is None:
            abi.check(self.lib.cj_batch_config_set(self.ctx, None))
            return
        cfg = abi.BatchConfig(load_factor, log_part, chunk, join_queue, grid)
        abi.check(self.lib.cj_batch_config_set(self.ctx, C.byref(cfg)))

    def execute_batch(self, plans, flags: int = 0):
        """Run many plans through cj_execute_batch; returns ([EngineResult],
        abi.BatchStats)."""
        n = len(plans)
        descs = (abi.PlanDesc * max(n, 1))()
        for i, p in enumerate(plans):
            descs[i] = p.desc()
        res = (abi.Result * max(n, 1))()
        st = abi.BatchStats()
        abi.check(self.lib.cj_execute_batch(self.ctx, descs, n, flags, res, C.byref(st)))
        return [EngineResult(self, res[i], None) for i in range(n)], st

    @staticmethod
    def footprint(plan: Plan, engine: "Engine" = None) -> int:
        """Footprint the serving loop charges for `plan` (with an engine: its
        resident tables counted as such)."""
        b = C.c_uint64()
        ctx = engine.ctx if engine is not None else None
        abi.check(abi.load_library().cj_plan_footprint(ctx, C.byref(plan.desc()), C.byref(b)))
        return b.value

    def kernel_stats(self):
        arr = (abi.KernelStat * 128)()
        n = C.c_uint32()
        abi.check(self.lib.cj_kernel_stats(self.ctx, arr, 128, C.byref(n)))
        return {arr[i].name.decode(): (arr[i].launches, arr[i].total_ms, arr[i].bytes)
                for i in range(min(n.value, 128))}

    def kernel_stats_reset(self):
        abi.check(self.lib.cj_kernel_stats_reset(self.ctx))


class EngineResult:
    def __init__(self, eng: Engine, res: abi.Result, timing):
        self.eng = eng
        self.res = res
        self.timing = timing
        self.num_rows = res.num_rows
        self.types = [res.columns[i].type for i in range(res.num_columns)]
        self._refresh()

    def _refresh(self):
        res = self.res
        self.page_lists = []
        for i in range(res.num_columns):
            c = res.columns[i]
            if c.pages and c.num_pages:
                # the page pointer array read at once (an index per page costs ~1 us in ctypes)
                addr = C.cast(c.pages, C.c_void_p).value
                arr = np.ctypeslib.as_array((C.c_uint64 * c.num_pages).from_address(addr))
                self.page_lists.append(arr.tolist())
            else:
                self.page_lists.append([])

    def to_host(self):
        abi.check(self.eng.lib.cj_result_to_host(self.eng.ctx, C.byref(self.res)))
        self._refresh()

    def num_pages(self):
        return [self.res.columns[i].num_pages for i in range(self.res.num_columns)]

    def release(self):
        if self.res.opaque:
            self.eng.lib.cj_result_release(self.eng.ctx, C.byref(self.res))

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


# ------------------------------------------------------ result fingerprints
_MIX = np.uint64(0x9E3779B97F4A7C15)


def _mix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser on a uint64 array."""
    x = x.copy()
    with np.errstate(over="ignore"):
        x ^= x >> np.uint64(30)
        x *= np.uint64(0xBF58476D1CE4E5B9)
        x ^= x >> np.uint64(27)
        x *= np.uint64(0x94D049BB133111EB)
        x ^= x >> np.uint64(31)
    return x


def _column_row_words(type_id: int, page_ptrs) -> np.ndarray:
    """One uint64 per row of a result column (host pages): the value's bits
    (8-byte values), the zero-extended INT32, a hash of the VARCHAR bytes;
    NULL rows get a marker distinct from every value word."""
    words = []
    null_word = np.uint64(0xD1B54A32D192ED03)
    i = 0
    while i < len(page_ptrs):
        p = np.ctypeslib.as_array((C.c_uint8 * PAGE).from_address(page_ptrs[i]))
        n = int(p[0]) | (int(p[1]) << 8)
        nv = int(p[2]) | (int(p[3]) << 8)
        if type_id == abi.CJ_VARCHAR and n == 0xFFFF:  # a long string: this page and its 0xfffe continuations
            chunks = [p[4:4 + nv].tobytes()]
            i += 1
            while i < len(page_ptrs):
                q = np.ctypeslib.as_array((C.c_uint8 * PAGE).from_address(page_ptrs[i]))
                if (int(q[0]) | (int(q[1]) << 8)) != 0xFFFE:
                    break
                chunks.append(q[4:4 + (int(q[2]) | (int(q[3]) << 8))].tobytes())
                i += 1
            words.append(np.array([hash(b"".join(chunks)) & 0xFFFFFFFFFFFFFFFF], dtype=np.uint64))
            continue
        bits = np.unpackbits(p[PAGE - (n + 7) // 8:], bitorder="little")[:n].astype(bool)
        w = np.full(n, null_word, dtype=np.uint64)
        if type_id == abi.CJ_VARCHAR:
            offs = p[4:4 + 2 * nv].view(np.uint16).astype(np.int64)
            chars = p[4 + 2 * nv:]
            start = np.concatenate([[0], offs[:-1]]) if nv else offs
            vals = np.array([hash(chars[a:b].tobytes()) & 0xFFFFFFFFFFFFFFFF for a, b in zip(start, offs)],
                            dtype=np.uint64)
        elif type_id == abi.CJ_INT32:
            vals = p[4:4 + 4 * nv].view(np.uint32).astype(np.uint64)
        else:
            vals = p[8:8 + 8 * nv].view(np.uint64)
        w[bits] = vals[:int(bits.sum())]
        words.append(w)
        i += 1
    return np.concatenate(words) if words else np.zeros(0, dtype=np.uint64)


def result_fingerprint(res: "EngineResult") -> np.ndarray:
    """The sorted per-row hashes of a result held in host pages: equal
    arrays mean equal row multisets (with overwhelming probability).  Used
    to show that a plan gives the same rows alone and in a served mix."""
    h = np.zeros(res.num_rows, dtype=np.uint64)
    for k, (t, pl) in enumerate(zip(res.types, res.page_lists)):
        w = _column_row_words(t, pl)
        if len(w) != res.num_rows:
            raise ValueError(f"column {k} holds {len(w)} rows, the result {res.num_rows}")
        with np.errstate(over="ignore"):
            h = _mix64(h * np.uint64(0x100000001B3) ^ (w + np.uint64(k + 1) * _MIX))
    return np.sort(h)
