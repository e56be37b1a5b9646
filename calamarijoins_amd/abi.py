"""ctypes mirror of include/cj_api.h and the loader of libcalamarijoins_amd.so.

The structs below follow the C declarations field by field.  Python code in
this package (the benchmark harness and the tests) reaches the HIP engine
only through these declarations.  Loading fails loudly when the library has
not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

PAGE_SIZE = 8192

CJ_OK = 0
CJ_ERR_INVALID, CJ_ERR_UNSUPPORTED, CJ_ERR_DEVICE, CJ_ERR_OOM, CJ_ERR_BUSY = -1, -2, -3, -4, -5
CJ_INT32, CJ_INT64, CJ_FP64, CJ_VARCHAR = 0, 1, 2, 3
CJ_NODE_JOIN, CJ_NODE_SCAN = 0, 1
CJ_EXEC_OUTPUT_DEVICE = 1
CJ_EXEC_NO_FAST_PATH = 2
CJ_EXEC_NO_DIRECT = 4
CJ_EXEC_NO_ALIGNED = 8
CJ_NOALLOC_SCRATCH_BYTES = 256 << 20
# cj_timing.phase_ms indices (the TimingStats phases)
CJ_PHASE_HASHTABLE_BUILD, CJ_PHASE_HASH_JOIN_PROBE, CJ_PHASE_NESTED_LOOP_JOIN, CJ_PHASE_MATERIALIZE, \
    CJ_PHASE_SETUP, CJ_PHASE_INTERMEDIATE = range(6)

TYPE_NAMES = {"INT32": CJ_INT32, "INT64": CJ_INT64, "FP64": CJ_FP64, "VARCHAR": CJ_VARCHAR}
TYPE_IDS = {v: k for k, v in TYPE_NAMES.items()}


class ColumnDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("num_pages", C.c_uint32), ("pages", C.POINTER(C.c_void_p))]


class TableDesc(C.Structure):
    _fields_ = [("num_rows", C.c_uint64), ("num_columns", C.c_uint32), ("columns", C.POINTER(ColumnDesc))]


class Attr(C.Structure):
    _fields_ = [("index", C.c_uint64), ("type", C.c_int32)]


class NodeDesc(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("build_left", C.c_int32),
        ("base_table_id", C.c_uint64),
        ("left", C.c_uint64),
        ("right", C.c_uint64),
        ("left_attr", C.c_uint64),
        ("right_attr", C.c_uint64),
        ("num_output_attrs", C.c_uint32),
        ("output_attrs", C.POINTER(Attr)),
    ]


class PlanDesc(C.Structure):
    _fields_ = [
        ("num_nodes", C.c_uint32),
        ("nodes", C.POINTER(NodeDesc)),
        ("num_inputs", C.c_uint32),
        ("inputs", C.POINTER(TableDesc)),
        ("root", C.c_uint64),
    ]


class Timing(C.Structure):
    _fields_ = [
        ("hashtable_build_ms", C.c_int64),
        ("hash_join_probe_ms", C.c_int64),
        ("nested_loop_join_ms", C.c_int64),
        ("materialize_ms", C.c_int64),
        ("setup_ms", C.c_int64),
        ("total_execution_ms", C.c_int64),
        ("intermediate_ms", C.c_int64),
        ("device_ms", C.c_double),
        ("h2d_ms", C.c_double),
        ("d2h_ms", C.c_double),
        ("bytes_alg", C.c_uint64),
        ("queue_ms", C.c_double),
        ("service_ms", C.c_double),
        ("phase_ms", C.c_double * 6),
        ("served_alone", C.c_int32),
    ]


class ServeStats(C.Structure):
    _fields_ = [("admitted", C.c_uint64), ("budget_waits", C.c_uint64), ("retried_alone", C.c_uint64),
                ("peak_admitted_bytes", C.c_uint64), ("peak_running", C.c_int32)]


class ServeOptions(C.Structure):
    _fields_ = [("workers", C.c_int32), ("admit_bytes", C.c_uint64)]


CJ_BATCH_AUTOTUNE = 16

CJ_F_CMP, CJ_F_AND, CJ_F_OR, CJ_F_NOT = range(4)
(CJ_OP_EQ, CJ_OP_NEQ, CJ_OP_LT, CJ_OP_GT, CJ_OP_LEQ, CJ_OP_GEQ, CJ_OP_LIKE, CJ_OP_NOT_LIKE, CJ_OP_IS_NULL,
 CJ_OP_IS_NOT_NULL) = range(10)
CJ_LIT_NONE, CJ_LIT_INT, CJ_LIT_FLOAT, CJ_LIT_STR = range(4)


class FilterNode(C.Structure):
    _fields_ = [("kind", C.c_int32), ("op", C.c_int32), ("column", C.c_uint32), ("lit_kind", C.c_int32),
                ("ival", C.c_int64), ("fval", C.c_double), ("sval", C.c_char_p), ("slen", C.c_uint32)]


class BatchConfig(C.Structure):
    _fields_ = [("load_factor", C.c_uint32), ("log_part", C.c_uint32), ("chunk", C.c_uint32),
                ("join_queue", C.c_uint32), ("grid", C.c_uint32)]


class BatchStats(C.Structure):
    _fields_ = [
        ("plans_batched", C.c_uint32), ("plans_single", C.c_uint32), ("device_ms", C.c_double),
        ("total_ms", C.c_double), ("output_rows", C.c_uint64), ("load_factor", C.c_uint32), ("grid", C.c_uint32),
        ("tuned", C.c_int32), ("graph_replayed", C.c_int32),
        ("log_part", C.c_uint32),
        ("chunk", C.c_uint32),
        ("join_queue", C.c_uint32),
    ]


class ResultColumn(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("num_pages", C.c_uint32),
        ("pages", C.POINTER(C.c_void_p)),
        ("device_pages", C.c_void_p),
    ]


class Result(C.Structure):
    _fields_ = [
        ("num_rows", C.c_uint64),
        ("num_columns", C.c_uint32),
        ("columns", C.POINTER(ResultColumn)),
        ("opaque", C.c_void_p),
    ]


class CtxOptions(C.Structure):
    _fields_ = [
        ("device", C.c_int32),
        ("device_budget", C.c_uint64),
        ("num_streams", C.c_int32),
        ("cache_decoded", C.c_int32),
        ("kernel_timing", C.c_int32),
        ("host_threads", C.c_int32),
    ]


class KernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("launches", C.c_uint64), ("total_ms", C.c_double), ("bytes", C.c_uint64)]


PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libcalamarijoins_amd.so")

# Every function include/cj_api.h declares (checked by the CPU test suite).
EXPORTS = [
    "cj_ctx_create", "cj_ctx_destroy", "cj_last_error", "cj_ctx_device_info",
    "cj_table_register", "cj_table_unregister",
    "cj_execute", "cj_result_release", "cj_result_to_host",
    "cj_kernel_stats", "cj_kernel_stats_reset",
    "cj_serve_start", "cj_submit", "cj_wait", "cj_serve_stop", "cj_plan_footprint", "cj_execute_batch", "cj_batch_config_set",
    "cj_tbl_write", "cj_tbl_open", "cj_tbl_load", "cj_tbl_close", "cj_tbl_cache_name", "cj_filter",
]

_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load the HIP engine.  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build the HIP engine first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(path)
    P = C.POINTER
    sig = {
        "cj_ctx_create": (C.c_int, [P(CtxOptions), P(C.c_void_p)]),
        "cj_ctx_destroy": (C.c_int, [C.c_void_p]),
        "cj_last_error": (C.c_char_p, []),
        "cj_ctx_device_info": (C.c_int, [C.c_void_p, C.c_char_p, C.c_size_t, P(C.c_int32)]),
        "cj_ctx_memory": (C.c_int, [C.c_void_p, C.c_int32, P(C.c_uint64), P(C.c_uint64), P(C.c_uint64),
                                    P(C.c_uint64)]),
        "cj_table_register": (C.c_int, [C.c_void_p, P(TableDesc), P(C.c_uint64)]),
        "cj_table_unregister": (C.c_int, [C.c_void_p, C.c_uint64]),
        "cj_execute": (C.c_int, [C.c_void_p, P(PlanDesc), C.c_uint32, P(Result), P(Timing)]),
        "cj_result_release": (C.c_int, [C.c_void_p, P(Result)]),
        "cj_execute_noalloc": (C.c_int, [C.c_void_p, P(PlanDesc), C.c_uint32, P(C.c_void_p), P(C.c_uint32),
                                         C.c_uint32, P(C.c_uint64), P(C.c_uint32)]),
        "cj_result_to_host": (C.c_int, [C.c_void_p, P(Result)]),
        "cj_kernel_stats": (C.c_int, [C.c_void_p, P(KernelStat), C.c_uint32, P(C.c_uint32)]),
        "cj_kernel_stats_reset": (C.c_int, [C.c_void_p]),
        "cj_serve_start": (C.c_int, [C.c_void_p, P(ServeOptions)]),
        "cj_submit": (C.c_int, [C.c_void_p, P(PlanDesc), C.c_uint32, P(C.c_uint64)]),
        "cj_wait": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int64, P(Result), P(Timing)]),
        "cj_serve_stop": (C.c_int, [C.c_void_p]),
        "cj_serve_stats_get": (C.c_int, [C.c_void_p, P(ServeStats)]),
        "cj_plan_footprint": (C.c_int, [C.c_void_p, P(PlanDesc), P(C.c_uint64)]),
        "cj_execute_batch": (C.c_int, [C.c_void_p, P(PlanDesc), C.c_uint32, C.c_uint32, P(Result), P(BatchStats)]),
        "cj_batch_config_set": (C.c_int, [C.c_void_p, P(BatchConfig)]),
        "cj_tbl_write": (C.c_int, [C.c_char_p, P(TableDesc)]),
        "cj_filter": (C.c_int, [C.c_void_p, P(TableDesc), P(FilterNode), C.c_uint32, C.c_uint32, P(Result),
                                P(Timing)]),
        "cj_tbl_open": (C.c_int, [C.c_char_p, P(C.c_void_p), P(TableDesc)]),
        "cj_tbl_load": (C.c_int, [C.c_void_p, C.c_char_p, P(C.c_void_p), P(TableDesc), P(C.c_uint64)]),
        "cj_tbl_close": (C.c_int, [C.c_void_p]),
        "cj_tbl_cache_name": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_size_t]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class CJError(RuntimeError):
    pass


def check(rc: int) -> None:
    if rc != CJ_OK:
        msg = load_library().cj_last_error()
        raise CJError(f"cj status {rc}: {msg.decode() if msg else ''}")
