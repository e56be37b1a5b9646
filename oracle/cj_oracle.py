"""ctypes binding of the CPU oracle (oracle/libcj_oracle.so).

TEST INFRASTRUCTURE: imported by the test suite, the smoke check and the
CPU-baseline leg of bench.py only; the product never loads the oracle.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libcj_oracle.so")

_lib = None


class OracleResult(C.Structure):
    _fields_ = [
        ("num_rows", C.c_uint64),
        ("num_columns", C.c_uint32),
        ("types", C.POINTER(C.c_int32)),
        ("num_pages", C.POINTER(C.c_uint32)),
        ("pages", C.POINTER(C.POINTER(C.c_void_p))),
    ]


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-C", HERE], check=True, capture_output=True)
    from calamarijoins_amd import abi

    L = C.CDLL(ORACLE_SO)
    P = C.POINTER
    L.cjo_encode_fixed.restype = C.c_size_t
    L.cjo_encode_fixed.argtypes = [C.c_int32, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
    L.cjo_encode_varchar.restype = C.c_size_t
    L.cjo_encode_varchar.argtypes = [C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
    L.cjo_decode_fixed.restype = C.c_int64
    L.cjo_decode_fixed.argtypes = [C.c_int32, P(C.c_void_p), C.c_uint32, C.c_void_p, C.c_void_p]
    L.cjo_decode_varchar.restype = C.c_int64
    L.cjo_decode_varchar.argtypes = [P(C.c_void_p), C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p,
                                     P(C.c_uint64)]
    L.cjo_row_hashes.restype = C.c_int
    L.cjo_row_hashes.argtypes = [P(abi.TableDesc), C.c_void_p]
    L.cjo_execute.restype = C.c_int
    L.cjo_execute.argtypes = [P(abi.PlanDesc), C.c_int, P(OracleResult)]
    L.cjo_result_free.restype = None
    L.cjo_result_free.argtypes = [P(OracleResult)]
    L.cjo_last_error.restype = C.c_char_p
    L.cjo_crc32c_u32.restype = C.c_uint32
    L.cjo_crc32c_u32.argtypes = [C.c_uint32, C.c_uint32]
    L.cjo_crc32c_bytes.restype = C.c_uint32
    L.cjo_crc32c_bytes.argtypes = [C.c_uint32, C.c_char_p, C.c_size_t]
    L.cjo_bloom_tag.restype = C.c_uint16
    L.cjo_bloom_tag.argtypes = [C.c_uint32]
    L.cjo_crc32c_hw_mismatches.restype = C.c_int64
    L.cjo_crc32c_hw_mismatches.argtypes = [C.c_uint64]
    _lib = L
    return L


class OracleRun:
    """One execute() of a plan (calamarijoins_amd.host.Plan) on the CPU oracle."""

    def __init__(self, plan, threads: int = 0):
        L = lib()
        self.res = OracleResult()
        rc = L.cjo_execute(C.byref(plan.desc()), threads, C.byref(self.res))
        if rc != 0:
            raise RuntimeError(f"oracle status {rc}: {L.cjo_last_error().decode()}")
        self.num_rows = self.res.num_rows
        self.types = [self.res.types[i] for i in range(self.res.num_columns)]
        self.page_lists = [[self.res.pages[i][q] for q in range(self.res.num_pages[i])]
                           for i in range(self.res.num_columns)]

    def __del__(self):
        try:
            lib().cjo_result_free(C.byref(self.res))
        except Exception:
            pass
