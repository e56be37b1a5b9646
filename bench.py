#!/usr/bin/env python3
"""Benchmark: the single large join (4M-row build x 64M-row Zipf probe) on one
MI355X through the C ABI, base-table pages resident in HBM, result pages left
in HBM.

One step = one cj_execute() of the plan: key-column decode and the build
structure (direct map for the unique dense build keys, else radix-partitioned
hash table), probe count + pair write, and the page writers of the 5 payload
columns.  Prints one JSON line with the step throughput, the roofline of the
dominant kernel and a CPU baseline.
"""
import argparse
import json
import os
import platform
import statistics
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "join output rows/s + HBM GB/s (1 GPU); p50/p99 plan latency at fixed QPS"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def cpu_baseline(build, probe_rows, seed, threads, repeats):
    """The oracle's CPU restatement of execute() on a bounded sample: the full
    4M-row build side and the first `probe_rows` probe rows of the same
    generated workload (rows are a pure function of seed and row index)."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import cj_oracle
    from calamarijoins_amd import workloads

    _, p = workloads.c2_tables(seed=seed, build_rows=workloads.C2_BUILD, probe_rows=probe_rows)
    plan = workloads.c2_plan(build, p)
    cj_oracle.OracleRun(plan, threads=threads)  # cold run
    times, rows = [], 0
    for _ in range(repeats):
        t0 = time.perf_counter()
        r = cj_oracle.OracleRun(plan, threads=threads)
        times.append(time.perf_counter() - t0)
        rows = r.num_rows
        del r
    med = statistics.median(times)
    return {
        "value": rows / med,
        "unit": "rows/s",
        "cores": threads,
        "kind": "port",
        "sample": f"4M build x first {probe_rows // 1_000_000}M probe rows of the same workload; "
                  f"1 cold + {repeats} timed runs, median {med * 1e3:.0f} ms; {cpu_model()}",
    }


COPY_BYTES = 1 << 30


def copy_calibration():
    """Three device-to-device copies of COPY_BYTES: rocprofv3's FETCH_SIZE and
    WRITE_SIZE of this copy kernel should read COPY_BYTES / 1024 KiB each, and
    the ratio scales the same counters of the join kernels to bytes."""
    import torch
    src = torch.empty(COPY_BYTES, dtype=torch.uint8, device="cuda").fill_(1)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": COPY_BYTES, "kib": COPY_BYTES // 1024}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--build-rows", type=int, default=4_000_000)
    ap.add_argument("--probe-rows", type=int, default=64_000_000)
    ap.add_argument("--seed", type=int, default=1002)
    ap.add_argument("--cpu-sample-rows", type=int, default=16_000_000)
    ap.add_argument("--cpu-repeats", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of known size (for scaling rocprofv3 byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(one process drives one GPU; sharding across GPUs is not built)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return

    from calamarijoins_amd import abi, workloads
    from calamarijoins_amd.host import Engine

    t0 = time.time()
    build, probe = workloads.c2_tables(seed=args.seed, build_rows=args.build_rows, probe_rows=args.probe_rows)
    log(f"generated C2 tables in {time.time() - t0:.1f}s: build {build.num_rows} rows, probe {probe.num_rows} rows")
    plan = workloads.c2_plan(build, probe)

    eng = Engine(cache_decoded=0)
    name, cus = eng.device_info()
    log(f"device {name}, {cus} CUs")
    t0 = time.time()
    ids = [eng.register(build), eng.register(probe)]
    log(f"registered tables in HBM in {time.time() - t0:.1f}s")

    flags = abi.CJ_EXEC_OUTPUT_DEVICE
    first = eng.execute(plan, flags=flags, timing=True)
    out_rows = first.num_rows
    bytes_alg = first.timing.bytes_alg
    log(f"first run: {out_rows} rows, device {first.timing.device_ms:.3f} ms, bytes_alg {bytes_alg / 1e9:.3f} GB")
    first.release()
    for _ in range(args.warmup):
        eng.execute(plan, flags=flags).release()

    # timed region: K steps, synchronous execute() calls (each ends with the
    # result pages in HBM)
    dev_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = eng.execute(plan, flags=flags, timing=True)
        dev_ms.append(r.timing.device_ms)
        r.release()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    value = out_rows / (ms_per_step / 1e3)
    log(f"timed {args.steps} steps: {ms_per_step:.3f} ms/step wall, device median {statistics.median(dev_ms):.3f} ms")

    # per-kernel roofline from hipEvents on the execution stream
    eng.close()
    eng = Engine(cache_decoded=0, kernel_timing=1)
    ids = [eng.register(build), eng.register(probe)]
    eng.execute(plan, flags=flags).release()
    eng.kernel_stats_reset()
    prof_steps = 3
    for _ in range(prof_steps):
        eng.execute(plan, flags=flags).release()
    stats = eng.kernel_stats()
    eng.close()
    dominant = max(stats.items(), key=lambda kv: kv[1][1])
    kname, (launches, tot_ms, kbytes) = dominant
    avg_ms = tot_ms / launches
    per_launch_bytes = kbytes / launches
    achieved = per_launch_bytes / (avg_ms / 1e3) / 1e9
    kernels = {k: {"launches": v[0], "avg_ms": round(v[1] / v[0], 4),
                   "GBps": round(v[2] / (v[1] / 1e3) / 1e9, 1) if v[1] > 0 else None}
               for k, v in sorted(stats.items(), key=lambda kv: -kv[1][1])}
    log("kernels: " + json.dumps(kernels))

    cpu = None
    if not args.no_cpu_baseline:
        threads = min(os.cpu_count() or 1, 16)
        try:
            cpu = cpu_baseline(build, min(args.cpu_sample_rows, args.probe_rows), args.seed, threads, args.cpu_repeats)
        except Exception as e:  # the oracle is optional on the timing box
            log(f"cpu baseline unavailable: {e}")

    dev_med = statistics.median(dev_ms)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {
            "workload": "single INT32 equi-join, 4M-row build (unique shuffled keys) x 64M-row probe "
                        "(Zipf(1.0) foreign keys, 30% misses), 5 payload columns (INT32/INT64 per side, "
                        "FP64 probe), 2% NULL payloads, 8 KiB pages resident in HBM, result pages in HBM",
            "build_rows": build.num_rows,
            "probe_rows": probe.num_rows,
            "output_rows": out_rows,
            "seed": args.seed,
        },
        "device_ms_per_step": dev_med,
        "bytes_alg_per_step": bytes_alg,
        "hbm_GBps_alg": bytes_alg / (dev_med / 1e3) / 1e9,
        "hbm_frac_alg": bytes_alg / (dev_med / 1e3) / 1e9 / HBM_PEAK_GBS,
        "roofline": {
            "kernel": kname,
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "avg_launch_ms": avg_ms,
            "alg_bytes_per_launch": per_launch_bytes,
        },
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
