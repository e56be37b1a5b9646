"""Served plan stream (C5): open-loop Poisson arrivals of multi-join plans
over HBM-resident tables, run either one at a time (the plain drop-in
`execute()`, first come first served) or through the engine's serving loop
(cj_serve_start / cj_submit / cj_wait: admission by estimated footprint,
concurrent plans on separate streams).  Latency is measured from a plan's
arrival to its result being ready.

    python -m calamarijoins_amd.serving [--scale S] [--plans K] [--requests N] [--check]

prints one JSON line with p50 / p99 latency and rows/s of both modes, the
admission budget (by default twice the median plan footprint bounded from
the plans' shapes; --tight: 1.5 times the largest footprint the server
charges once the plans have run, their recorded peaks, so that it binds) and
the serving loop's counters, and the capacity of both modes (the first
--capacity requests issued at once: requests and rows per second).  --check
then serves the whole stream
once more and compares every result with its plan's result run alone (row
multiset fingerprints, calamarijoins_amd.host.result_fingerprint).
"""
from __future__ import annotations

import argparse
import json
import math
import random
import statistics
import sys
import threading
import time

from . import abi, workloads
from .host import Engine, result_fingerprint


def percentile(xs, q):
    if not xs:
        return None
    s = sorted(xs)
    k = (len(s) - 1) * q
    lo, hi = math.floor(k), math.ceil(k)
    return s[lo] + (s[hi] - s[lo]) * (k - lo)


def poisson_arrivals(n, qps, seed):
    rng = random.Random(seed)
    t, out = 0.0, []
    for _ in range(n):
        t += rng.expovariate(qps)
        out.append(t)
    return out


def run_plain(eng, plans, order, arrivals, flags):
    """One executor thread, first come first served (the drop-in path)."""
    lat, rows = [], 0
    t0 = time.perf_counter()
    for k, a in zip(order, arrivals):
        now = time.perf_counter() - t0
        if now < a:
            time.sleep(a - now)
        r = eng.execute(plans[k], flags=flags)
        done = time.perf_counter() - t0
        lat.append((done - a) * 1e3)
        rows += r.num_rows
        r.release()
    makespan = time.perf_counter() - t0
    return lat, rows, makespan


def run_served(eng, plans, order, arrivals, flags, collect=None):
    """Submissions at their arrival times; a collector thread waits on the
    tickets.  Latency = queue + service time measured by the server from
    submission (the arrival) to result ready."""
    tickets = []
    lock = threading.Condition()
    state = {"submitted": 0}

    def submitter():
        t0 = time.perf_counter()
        for k, a in zip(order, arrivals):
            now = time.perf_counter() - t0
            if now < a:
                time.sleep(a - now)
            t = eng.submit(plans[k], flags)
            with lock:
                tickets.append((k, t))
                state["submitted"] += 1
                lock.notify()

    lat, rows = [], 0
    t0 = time.perf_counter()
    th = threading.Thread(target=submitter)
    th.start()
    for i in range(len(order)):
        with lock:
            while state["submitted"] <= i:
                lock.wait()
            k, t = tickets[i]
        r = eng.wait(t)
        lat.append(r.timing.queue_ms + r.timing.service_ms)
        rows += r.num_rows
        if collect is not None:
            collect(k, r)
        r.release()
    th.join()
    makespan = time.perf_counter() - t0
    return lat, rows, makespan


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--scale", type=float, default=1.0, help="table sizes x scale (1.0 = 0.1M-36M rows)")
    ap.add_argument("--plans", type=int, default=32, help="distinct plans in the stream")
    ap.add_argument("--requests", type=int, default=2000, help="arrivals per mode")
    ap.add_argument("--workers", type=int, default=4)
    ap.add_argument("--admit-gib", type=float, default=0.0,
                    help="admission budget; 0 = twice the median plan footprint from the plans' shapes")
    ap.add_argument("--tight", action="store_true",
                    help="admission budget 1.5 x the largest footprint the server charges (learned after a run)")
    ap.add_argument("--check", action="store_true", help="serve once more and compare with each plan alone")
    ap.add_argument("--load", type=float, default=0.75, help="QPS = load / mean plain service time")
    ap.add_argument("--seed", type=int, default=1005)
    ap.add_argument("--capacity", type=int, default=1000,
                    help="requests served back to back in each mode for the capacity figure (0: skip)")
    args = ap.parse_args(argv)

    t0 = time.time()
    tables, meta = workloads.c5_tables(seed=args.seed, scale=args.scale)
    plans = [workloads.c5_plan(tables, meta, 5000 + k) for k in range(args.plans)]
    print(f"generated 12 tables ({sum(m['rows'] for m in meta)} rows) and {len(plans)} plans "
          f"in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    eng = Engine(cache_decoded=1, num_streams=max(4, args.workers))
    for tb in tables:
        eng.register(tb)
    flags = abi.CJ_EXEC_OUTPUT_DEVICE
    # footprints from the plans' shapes (before any has run), then the ones
    # the server charges once each plan has run: its recorded peak + 25 %
    foot_shape = [Engine.footprint(p, eng) for p in plans]
    # plain service time of each plan (warm: decoded columns cached)
    for p in plans:
        eng.execute(p, flags=flags).release()
    svc = []
    for p in plans:
        s = time.perf_counter()
        eng.execute(p, flags=flags).release()
        svc.append((time.perf_counter() - s) * 1e3)
    mean_svc = statistics.mean(svc)
    foot = [Engine.footprint(p, eng) for p in plans]
    if args.admit_gib > 0:
        admit = int(args.admit_gib * (1 << 30))
    elif args.tight:
        admit = int(1.5 * max(foot))  # binds: two of the largest plans never run together
    else:
        admit = 2 * int(statistics.median(foot_shape))
    qps = args.load / (mean_svc / 1e3)
    rng = random.Random(args.seed)
    order = [rng.randrange(len(plans)) for _ in range(args.requests)]
    arrivals = poisson_arrivals(args.requests, qps, args.seed + 1)
    print(f"mean plain service {mean_svc:.2f} ms -> {qps:.1f} QPS", file=sys.stderr, flush=True)

    lat_p, rows_p, span_p = run_plain(eng, plans, order, arrivals, flags)
    eng.serve_start(args.workers, admit)
    try:
        lat_s, rows_s, span_s = run_served(eng, plans, order, arrivals, flags)
        stats = eng.serve_stats()
    finally:
        eng.serve_stop()
    # capacity: the same requests all at once (no arrival waits), one at a
    # time against the serving loop, as requests per second
    cap = None
    if args.capacity:
        ncap = min(args.capacity, len(order))
        zero = [0.0] * ncap
        _, rows_pc, span_pc = run_plain(eng, plans, order[:ncap], zero, flags)
        eng.serve_start(args.workers, admit)
        try:
            _, rows_sc, span_sc = run_served(eng, plans, order[:ncap], zero, flags)
        finally:
            eng.serve_stop()
        cap = {"requests": ncap, "plain_qps": ncap / span_pc, "served_qps": ncap / span_sc,
               "plain_rows_per_s": rows_pc / span_pc, "served_rows_per_s": rows_sc / span_sc}
        print(f"capacity: {cap}", file=sys.stderr, flush=True)
    check = None
    if args.check:
        alone = {}
        for k in sorted(set(order)):
            r = eng.execute(plans[k], flags=0)
            alone[k] = (r.num_rows, result_fingerprint(r))
            r.release()
        bad = []

        def compare(k, r):
            r.to_host()
            if r.num_rows != alone[k][0] or not (result_fingerprint(r) == alone[k][1]).all():
                bad.append(k)

        eng.serve_start(args.workers, admit)
        try:
            run_served(eng, plans, order, arrivals, flags, collect=compare)
        finally:
            eng.serve_stop()
        check = {"requests": len(order), "distinct_plans": len(alone), "mismatches": len(bad)}
        print(f"check: {check}", file=sys.stderr, flush=True)
    out = {
        "workload": f"C5 served stream: 12 resident tables ({args.scale:g} x 0.1M-36M rows, 1-40% pre-filter), "
                    f"{len(plans)} plans of 3-8 joins, {args.requests} Poisson arrivals at {qps:.1f} QPS "
                    f"(load {args.load} of the mean plain service time {mean_svc:.2f} ms)",
        "plain": {"p50_ms": percentile(lat_p, 0.5), "p99_ms": percentile(lat_p, 0.99),
                  "rows_per_s": rows_p / span_p, "makespan_s": span_p},
        "served": {"p50_ms": percentile(lat_s, 0.5), "p99_ms": percentile(lat_s, 0.99),
                   "rows_per_s": rows_s / span_s, "makespan_s": span_s, "workers": args.workers,
                   "admit_bytes": admit, "stats": stats},
        "plan_footprint_bytes": {"min": min(foot), "median": statistics.median(foot), "max": max(foot)},
        "plan_footprint_bytes_from_shape": {"min": min(foot_shape), "median": statistics.median(foot_shape),
                                            "max": max(foot_shape)},
        "capacity": cap,
        "check_served_equals_alone": check,
    }
    print(json.dumps(out), flush=True)
    eng.close()


if __name__ == "__main__":
    main()
