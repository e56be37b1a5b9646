"""Batched independent joins (C4): 256 single-join plans over HBM-resident
tables, run either one at a time through the drop-in `execute()` or in one
`cj_execute_batch` call (shared persistent-kernel launches over a device
work list, resident grids, phase graph replayed, autotuned load factor, radix bits, join tile and join scheduling).

    python -m calamarijoins_amd.batching [--scale S] [--plans K] [--reps R] [--no-tune] [--batch-only]

(--no-tune: the default batch configuration, no autotuning; --batch-only:
skip the timed one-at-a-time runs, e.g. under a counter profiler)

prints one JSON line with the wall time, device time, rows/s and effective
GB/s (algorithmic bytes: input pages read + output pages written, summed over
the plans) of both modes.  Per-plan row counts of the two modes must agree;
row-level equality is checked by the GPU test suite.
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
    ap.add_argument("--scale", type=float, default=1.0, help="plan sizes x scale (1.0 = build 1e3-2e6, probe 1e4-2e7)")
    ap.add_argument("--plans", type=int, default=256)
    ap.add_argument("--reps", type=int, default=5, help="timed repetitions of each mode")
    ap.add_argument("--no-tune", action="store_true", help="run the default batch configuration (no autotuner)")
    ap.add_argument("--batch-only", action="store_true", help="skip the timed one-at-a-time runs")
    args = ap.parse_args(argv)

    t0 = time.time()
    pairs = []
    for k in range(args.plans):
        pairs.append(workloads.c4_plan_tables(k, scale=args.scale))
        if (k + 1) % 32 == 0:
            print(f"generated {k + 1} plans ({time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
    plans = [workloads.c4_plan(b, p) for b, p in pairs]
    build_rows = sum(b.num_rows for b, _ in pairs)
    probe_rows = sum(p.num_rows for _, p in pairs)
    eng = Engine(cache_decoded=1)
    for b, p in pairs:
        eng.register(b)
        eng.register(p)
    flags = abi.CJ_EXEC_OUTPUT_DEVICE

    # one at a time (warm: decoded payload columns cached)
    rows_seq, bytes_alg = [], 0
    for pl in plans:
        r = eng.execute(pl, flags=flags, timing=True)
        rows_seq.append(r.num_rows)
        bytes_alg += r.timing.bytes_alg
        r.release()
    print(f"sequential warm-up done ({time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
    # wall time of untimed passes (per-phase timing records events of its
    # own), device time from as many timed passes
    seq_ms, seq_dev = [], []
    for _ in range(0 if args.batch_only else args.reps):
        s = time.perf_counter()
        for pl in plans:
            eng.execute(pl, flags=flags).release()
        seq_ms.append((time.perf_counter() - s) * 1e3)
    for _ in range(0 if args.batch_only else args.reps):
        dev = 0.0
        for pl in plans:
            r = eng.execute(pl, flags=flags, timing=True)
            dev += r.timing.device_ms
            r.release()
        seq_dev.append(dev)
    if seq_ms:
        print(f"sequential {statistics.median(seq_ms):.1f} ms ({time.time() - t0:.1f}s)", file=sys.stderr, flush=True)

    # batched: the first call tunes (and captures the phase graph)
    res, st = eng.execute_batch(plans, flags | (0 if args.no_tune else abi.CJ_BATCH_AUTOTUNE))
    rows_batch = [r.num_rows for r in res]
    for r in res:
        r.release()
    if rows_batch != rows_seq:
        bad = [k for k in range(len(plans)) if rows_batch[k] != rows_seq[k]]
        raise SystemExit(f"batched row counts differ from per-plan execute for plans {bad[:8]}")
    tuned = (st.load_factor, st.grid, st.log_part, st.chunk, st.join_queue)
    b_ms, b_dev, replayed = [], [], 0
    for _ in range(args.reps):
        s = time.perf_counter()
        res, st = eng.execute_batch(plans, flags)
        b_ms.append((time.perf_counter() - s) * 1e3)
        b_dev.append(st.device_ms)
        replayed += st.graph_replayed
        for r in res:
            r.release()
    print(f"batched {statistics.median(b_ms):.1f} ms ({time.time() - t0:.1f}s)", file=sys.stderr, flush=True)

    out_rows = sum(rows_seq)
    b_w, b_d = statistics.median(b_ms), statistics.median(b_dev)
    seq = None
    if seq_ms:
        seq_w, seq_d = statistics.median(seq_ms), statistics.median(seq_dev)
        seq = {"wall_ms": seq_w, "device_ms": seq_d, "rows_per_s": out_rows / (seq_w / 1e3),
               "gb_per_s": bytes_alg / (seq_w / 1e3) / 1e9}
    out = {
        "workload": f"C4: {len(plans)} independent joins (plan k seed 2000+k; build {scale_txt(args.scale, 1e3, 2e6)}, "
                    f"probe {scale_txt(args.scale, 1e4, 2e7)} rows log-uniform; key domain [0, build); "
                    "uniform or Zipf(0.8) probe keys; NULL U[0,0.10]; out [b.key, b.pay, p.pay])",
        "build_rows": build_rows, "probe_rows": probe_rows, "output_rows": out_rows,
        "bytes_alg": bytes_alg,
        "sequential": seq,
        "batched": {"wall_ms": b_w, "device_ms": b_d, "rows_per_s": out_rows / (b_w / 1e3),
                    "gb_per_s": bytes_alg / (b_w / 1e3) / 1e9, "device_gb_per_s": bytes_alg / (b_d / 1e3) / 1e9,
                    "load_factor": tuned[0], "join_write_grid": tuned[1], "log_part": tuned[2], "chunk": tuned[3], "join_queue": tuned[4],
                    "graph_replays": replayed,
                    "plans_batched": st.plans_batched},
        "speedup_wall": seq["wall_ms"] / b_w if seq else None,
        "reps": args.reps,
    }
    print(json.dumps(out), flush=True)
    eng.close()


def scale_txt(scale, lo, hi):
    return f"[{lo * scale:.3g}, {hi * scale:.3g}]"


if __name__ == "__main__":
    main()
