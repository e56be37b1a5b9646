"""Memory microbenchmarks behind the roofline figures (include/cj_membench.h):

* a streaming device copy (16 bytes a lane) of 1 GiB: the achievable
  read+write ceiling, and the byte scale of the profiler's FETCH_SIZE /
  WRITE_SIZE counters (a copy of B bytes must read and write B bytes);
* random gathers of 16-byte records from a 64 MB table (the direct map of
  the 4M-row build side), once with uniform indices and once with the real
  lookup stream of the 4M x 64M join (the matched probe keys, Zipf(1.0) over
  a shuffled key order);
* random gathers of 4-byte values inside an L2-sized table (2 MB) and far
  beyond it (1 GB).

    python -m calamarijoins_amd.membench [--lookups N]

prints one JSON line: per case the best time, G lookups/s and the requested
bytes per second."""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys

import numpy as np

from . import abi, workloads


def _lib():
    lib = abi.load_library()
    lib.cj_membench_copy.restype = C.c_int
    lib.cj_membench_copy.argtypes = [C.c_int32, C.c_uint64, C.c_uint32, C.POINTER(C.c_double)]
    lib.cj_membench_gather.restype = C.c_int
    lib.cj_membench_gather.argtypes = [C.c_int32, C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32,
                                       C.POINTER(C.c_double)]
    return lib


def copy_ms(nbytes: int, reps: int = 5) -> float:
    ms = C.c_double()
    abi.check(_lib().cj_membench_copy(0, nbytes, reps, C.byref(ms)))
    return ms.value


def gather_ms(idx: np.ndarray, table_bytes: int, record_bytes: int, reps: int = 5) -> float:
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    ms = C.c_double()
    abi.check(_lib().cj_membench_gather(0, idx.ctypes.data, len(idx), table_bytes, record_bytes, reps, C.byref(ms)))
    return ms.value


def gather16_mode_ms(idx: np.ndarray, table_bytes: int, mode: int, reps: int = 5) -> float:
    """16-byte records: mode 0 plain loads, 1 non-temporal loads, 2 lookups
    routed by table eighth to the XCDs (cj_membench_gather16_mode)."""
    lib = _lib()
    lib.cj_membench_gather16_mode.restype = C.c_int
    lib.cj_membench_gather16_mode.argtypes = [C.c_int32, C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32,
                                              C.POINTER(C.c_double)]
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    ms = C.c_double()
    abi.check(lib.cj_membench_gather16_mode(0, idx.ctypes.data, len(idx), table_bytes, mode, reps, C.byref(ms)))
    return ms.value


def c2_lookup_slots(build_rows: int, probe_rows: int) -> np.ndarray:
    """Direct-map slots of the matched probe keys of the single large join,
    in probe order (its INT32 key pages hold no NULLs: values at +4)."""
    _, p = workloads.c2_tables(build_rows=build_rows, probe_rows=probe_rows)
    pages = p.columns[0].pages
    n = pages[:, 0].astype(np.int64) | (pages[:, 1].astype(np.int64) << 8)
    keys = np.concatenate([pages[i, 4:4 + 4 * int(n[i])].view(np.int32) for i in range(len(n))])
    return keys[keys < build_rows].astype(np.uint32)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--build-rows", type=int, default=workloads.C2_BUILD)
    ap.add_argument("--probe-rows", type=int, default=workloads.C2_PROBE)
    ap.add_argument("--reps", type=int, default=5)
    args = ap.parse_args()
    out = {}
    gib = 1 << 30
    ms = copy_ms(gib, args.reps)
    out["copy_1GiB"] = {"ms": ms, "GBps_read_plus_write": 2 * gib / (ms / 1e3) / 1e9}
    print("copy done", file=sys.stderr, flush=True)
    rng = np.random.default_rng(7)
    slots = c2_lookup_slots(args.build_rows, args.probe_rows)
    n = len(slots)
    table16 = args.build_rows * 16
    cases = [
        ("gather16_c2_zipf_64MB", slots, table16, 16),
        ("gather16_uniform_64MB", rng.integers(0, args.build_rows, n, dtype=np.uint32), table16, 16),
        ("gather16_uniform_2MB", rng.integers(0, (2 << 20) // 16, n, dtype=np.uint32), 2 << 20, 16),
        ("gather4_uniform_2MB", rng.integers(0, (2 << 20) // 4, n, dtype=np.uint32), 2 << 20, 4),
        ("gather4_uniform_1GB", rng.integers(0, gib // 4, n, dtype=np.uint32), gib, 4),
    ]
    for name, idx, tb, w in cases:
        ms = gather_ms(idx, tb, w, args.reps)
        out[name] = {"lookups": len(idx), "table_bytes": tb, "record_bytes": w, "ms": ms,
                     "G_lookups_per_s": len(idx) / (ms / 1e3) / 1e9,
                     "GBps_requested": len(idx) * w / (ms / 1e3) / 1e9}
        print(f"{name} done", file=sys.stderr, flush=True)
    # the C2 lookup stream with another cache policy, or routed by table
    # slice to the XCDs (an upper bound for XCD-aware routing of the gather)
    for name, mode in (("gather16_c2_zipf_64MB_nontemporal", 1), ("gather16_c2_zipf_64MB_xcd_sliced", 2)):
        ms = gather16_mode_ms(slots, table16, mode, args.reps)
        out[name] = {"lookups": n, "table_bytes": table16, "record_bytes": 16, "ms": ms,
                     "G_lookups_per_s": n / (ms / 1e3) / 1e9, "GBps_requested": n * 16 / (ms / 1e3) / 1e9}
        print(f"{name} done", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
