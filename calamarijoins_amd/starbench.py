"""Left-deep star pipeline (C3) at full scale: a 36M-row fact table with four
INT32 foreign keys joined to four dimensions of 2.5M, 4M, 0.13M and 1.2M rows
(unique shuffled keys, each pre-filtered to a seeded 1-40 % of its rows, a
VARCHAR payload of 5-60 ASCII characters with one string in 10^5 of 9-20 KiB
on long-string pages), output [D1.s, D2.s, D3.s, F.fk1, F.fk4].  All tables
are resident in HBM; every step is one synchronous `cj_execute` with the
result pages left in HBM (decode of the referenced columns included).

    python -m calamarijoins_amd.starbench [--scale S] [--steps K] [--warmup W]

prints one JSON line: output rows, wall and device ms per step, rows/s,
algorithmic bytes (input pages read + output pages written) per second, and
the per-kernel hipEvent averages of a separate profiled run.  Row-level
equality with the reference semantics is checked by the GPU test suite at
1/10 and 1/100 of these sizes.
"""
from __future__ import annotations

import argparse
import json
import statistics
import sys
import time

from . import abi, workloads
from .host import Engine


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--scale", type=float, default=1.0, help="table sizes x scale (1.0 = 36M fact rows)")
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    args = ap.parse_args(argv)

    t0 = time.time()
    fact_rows = int(workloads.C3_FACT * args.scale)
    dims = tuple(max(1, int(d * args.scale)) for d in workloads.C3_DIMS)
    fact, ds, keep = workloads.c3_tables(fact_rows=fact_rows, dims=dims)
    plan = workloads.c3_plan(fact, ds)
    print(f"generated C3 tables in {time.time() - t0:.1f}s: fact {fact.num_rows}, dims "
          f"{[d.num_rows for d in ds]}", file=sys.stderr, flush=True)
    flags = abi.CJ_EXEC_OUTPUT_DEVICE

    eng = Engine(cache_decoded=0)
    for t in [fact] + ds:
        eng.register(t)
    r = eng.execute(plan, flags=flags, timing=True)
    rows, bytes_alg = r.num_rows, r.timing.bytes_alg
    r.release()
    for _ in range(args.warmup):
        eng.execute(plan, flags=flags).release()
    # wall time of untimed steps (per-phase timing records events of its
    # own), then the device time of as many timed steps
    s = time.perf_counter()
    for _ in range(args.steps):
        eng.execute(plan, flags=flags).release()
    wall_ms = (time.perf_counter() - s) * 1e3 / args.steps
    dev = []
    for _ in range(args.steps):
        r = eng.execute(plan, flags=flags, timing=True)
        dev.append(r.timing.device_ms)
        r.release()
    eng.close()
    print(f"{rows} rows, {wall_ms:.2f} ms/step wall", file=sys.stderr, flush=True)

    # per-kernel hipEvent averages (a separate engine: the events serialize)
    eng = Engine(cache_decoded=0, kernel_timing=1)
    for t in [fact] + ds:
        eng.register(t)
    eng.execute(plan, flags=flags).release()
    eng.kernel_stats_reset()
    prof = 3
    for _ in range(prof):
        eng.execute(plan, flags=flags).release()
    stats = eng.kernel_stats()
    eng.close()
    kernels = {k: {"launches": v[0] // prof, "avg_ms": round(v[1] / v[0], 4),
                   "ms_per_step": round(v[1] / prof, 4)}
               for k, v in sorted(stats.items(), key=lambda kv: -kv[1][1])}

    dev_ms = statistics.median(dev)
    out = {
        "workload": "C3: left-deep 4-join star, fact {} rows x dims {} (kept fractions {}), output 3 VARCHAR + "
                    "2 INT32; tables resident in HBM, result pages in HBM".format(
                        fact.num_rows, [d.num_rows for d in ds], [round(k, 3) for k in keep]),
        "output_rows": rows,
        "steps": args.steps,
        "ms_per_step": wall_ms,
        "device_ms_per_step": dev_ms,
        "rows_per_s": rows / (wall_ms / 1e3),
        "bytes_alg_per_step": bytes_alg,
        "alg_GBps_device": bytes_alg / (dev_ms / 1e3) / 1e9,
        "kernels": kernels,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
