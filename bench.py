#!/usr/bin/env python3
"""Benchmark: the single large join (4M-row build x 64M-row Zipf probe) on one
MI355X through the C ABI, base-table pages resident in HBM, result pages left
in HBM.

One step = one cj_execute() of the plan: key-column decode and the build
structure (direct map for the unique dense build keys, else radix-partitioned
hash table), the counting pass, and the page writers of the 5 payload
columns.  Prints one JSON line with the step throughput, the roofline of the
dominant kernel, a measured device-copy ceiling and a CPU baseline.

Algorithmic bytes follow the per-unit definition used throughout: every
input page a step references is read once (8192 B), every output value and
bitmap byte is written once; build structures, match bitmaps and row-id lists
are not counted.
"""
import argparse
import ctypes as C
import json
import os
import platform
import re
import statistics
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

PHASES = ["hashtable_build", "hash_join_probe", "nested_loop_join", "materialize", "setup", "intermediate"]
METRIC = "join output rows/s + HBM GB/s (1 GPU); p50/p99 plan latency at fixed QPS"
HBM_PEAK_GBS = 8000.0
PAGE = 8192
TRAFFIC_FILE = os.path.join(REPO, "profiles", "c2_join_4Mx64M_pmc_traffic.md")
# timed scope of the engine -> HIP kernel it launches (for the PMC traffic table)
SCOPE_KERNEL = {
    "build_pages": "k_build_pages",
    "probe_count": "k_count_match_direct",
    "probe_filter_w4": "k_probe_filter_wave<4>",
    "probe_filter_w8": "k_probe_filter_wave<8>",
}


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


class Hip:
    """The few HIP runtime calls the copy ceiling needs (libamdhip64)."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        for n in ("hipMalloc", "hipFree", "hipMemcpy", "hipMemset", "hipEventCreate", "hipEventRecord",
                  "hipEventSynchronize", "hipEventElapsedTime", "hipEventDestroy", "hipDeviceSynchronize"):
            getattr(self.lib, n).restype = C.c_int
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        self.lib.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.lib.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]

    def ok(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed with HIP error {rc}")

    def copy_ms(self, nbytes, reps):
        """Best time of a device-to-device hipMemcpy of nbytes (a plain copy kernel)."""
        a, b = C.c_void_p(), C.c_void_p()
        self.ok(self.lib.hipMalloc(C.byref(a), C.c_size_t(nbytes)), "hipMalloc")
        self.ok(self.lib.hipMalloc(C.byref(b), C.c_size_t(nbytes)), "hipMalloc")
        e0, e1 = C.c_void_p(), C.c_void_p()
        self.ok(self.lib.hipEventCreate(C.byref(e0)), "hipEventCreate")
        self.ok(self.lib.hipEventCreate(C.byref(e1)), "hipEventCreate")
        try:
            self.ok(self.lib.hipMemset(a, 1, nbytes), "hipMemset")
            self.ok(self.lib.hipMemcpy(b, a, nbytes, 3), "hipMemcpy")  # warm
            best = float("inf")
            for _ in range(reps):
                self.ok(self.lib.hipEventRecord(e0, None), "hipEventRecord")
                self.ok(self.lib.hipMemcpy(b, a, nbytes, 3), "hipMemcpy")
                self.ok(self.lib.hipEventRecord(e1, None), "hipEventRecord")
                self.ok(self.lib.hipEventSynchronize(e1), "hipEventSynchronize")
                ms = C.c_float()
                self.ok(self.lib.hipEventElapsedTime(C.byref(ms), e0, e1), "hipEventElapsedTime")
                best = min(best, ms.value)
            return best
        finally:
            self.lib.hipEventDestroy(e0)
            self.lib.hipEventDestroy(e1)
            self.lib.hipFree(a)
            self.lib.hipFree(b)


COPY_BYTES = 1 << 30


def copy_calibration():
    """Three device-to-device copies of COPY_BYTES, a kernel of known traffic
    for checking how rocprofv3's counters convert to bytes: its read requests
    (TCC_EA0_RDREQ by size) must add up to COPY_BYTES read, its WRITE_SIZE to
    COPY_BYTES / 1024 KiB written."""
    Hip().copy_ms(COPY_BYTES, 3)
    print(json.dumps({"copy_calibration_bytes": COPY_BYTES, "kib": COPY_BYTES // 1024}), flush=True)


def page_payload_bytes(ptr, num_pages, type_id):
    """Non-NULL value bytes + bitmap bytes of a fixed-width result column
    whose pages are contiguous in host memory starting at ptr."""
    if not num_pages:
        return 0
    buf = np.ctypeslib.as_array((C.c_uint8 * (num_pages * PAGE)).from_address(ptr)).reshape(num_pages, PAGE)
    n = buf[:, 0].astype(np.int64) | (buf[:, 1].astype(np.int64) << 8)
    nv = buf[:, 2].astype(np.int64) | (buf[:, 3].astype(np.int64) << 8)
    w = 4 if type_id == 0 else 8
    return int((nv * w).sum() + ((n + 7) // 8).sum())


def pmc_traffic(kernel):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 --pmc
    summary: line-fill reads (TCC_EA0_RDREQ by request size) + WRITE_SIZE,
    both checked against a copy of known size."""
    try:
        text = open(TRAFFIC_FILE).read()
    except OSError:
        return None
    for line in text.splitlines():
        cells = [c.strip().strip("`") for c in line.strip().strip("|").split("|")]
        if len(cells) >= 4 and cells[0].startswith(kernel):
            try:
                return float(cells[-1])
            except ValueError:
                return None
    return None


def host_cpu_share():
    """Host threads this process may use: the affinity mask, capped by
    OMP_NUM_THREADS when the machine sets a per-job share."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    share = os.environ.get("OMP_NUM_THREADS")
    if share and share.isdigit() and int(share) > 0:
        n = min(n, int(share))
    return n


def cpu_baseline(plan, thread_counts, repeats):
    """The oracle's CPU restatement of execute() (radix-partitioned Bloom-tag
    hash join, per-page work stealing, per-thread page builders) on the same
    plan the GPU runs: 1 cold run, then `repeats` timed runs, median
    (as the reference's scripts/benchmark.py:76-127), for each thread count."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import cj_oracle

    runs = {}
    rows = 0
    for t in thread_counts:
        cj_oracle.OracleRun(plan, threads=t)  # cold run
        times = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            r = cj_oracle.OracleRun(plan, threads=t)
            times.append(time.perf_counter() - t0)
            rows = r.num_rows
            del r
        runs[t] = statistics.median(times)
        log(f"cpu baseline: {t} threads, median {runs[t] * 1e3:.0f} ms over {repeats} runs")
    best_t = min(runs, key=runs.get)
    return {
        "value": rows / runs[best_t],
        "unit": "rows/s",
        "cores": best_t,
        "kind": "port",
        "sample": f"CPU restatement of execute() (reported baseline only) on the full workload (4M build x 64M "
                  f"probe, {rows} output rows); 1 cold + {repeats} timed runs per thread count, median ms "
                  + ", ".join(f"{t} threads {runs[t] * 1e3:.0f}" for t in runs)
                  + f"; host: {cpu_model()}, {os.cpu_count()} hardware threads, {host_cpu_share()} in this job's share",
        "median_ms_by_threads": {str(t): runs[t] * 1e3 for t in runs},
    }


def thread_counts(spec):
    counts = []
    for t in spec.split(","):
        n = host_cpu_share() if t == "share" else int(t)
        if n not in counts:
            counts.append(n)
    return counts


def time_oracle(plans, threads, repeats):
    """The oracle over `plans` one after another (as execute() runs them),
    1 cold pass then `repeats` timed passes: (median seconds per pass, output
    rows of a pass, per-plan seconds of the last pass)."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import cj_oracle

    def one_pass():
        rows, per = 0, []
        for pl in plans:
            t0 = time.perf_counter()
            r = cj_oracle.OracleRun(pl, threads=threads)
            per.append(time.perf_counter() - t0)
            rows += r.num_rows
            del r
        return rows, per

    t0 = time.perf_counter()
    one_pass()
    log(f"  oracle cold pass, {threads} threads: {time.perf_counter() - t0:.1f} s")
    secs, rows, per = [], 0, []
    for _ in range(repeats):
        t0 = time.perf_counter()
        rows, per = one_pass()
        secs.append(time.perf_counter() - t0)
        log(f"  oracle timed pass, {threads} threads: {secs[-1]:.1f} s")
    return statistics.median(secs), rows, per


def other_workload(args):
    """--workload c3|c4|c5: the workload's GPU harness (run as a child
    process; its JSON line is kept whole under "gpu"), then the oracle on the
    same generated plans on this host's cores (a reported baseline, never the
    measured path), one JSON line."""
    import subprocess
    from calamarijoins_amd import workloads

    mod = {"c3": "starbench", "c4": "batching", "c5": "serving"}[args.workload]
    cmd = [sys.executable, "-m", f"calamarijoins_amd.{mod}"] + (["--check"] if args.workload == "c5" else [])
    log(f"{args.workload}: running {' '.join(cmd[1:])}")
    proc = subprocess.run(cmd, cwd=REPO, capture_output=True, text=True)
    sys.stderr.write(proc.stderr[-4000:])
    if proc.returncode != 0:
        print(proc.stdout[-2000:], file=sys.stderr)
        sys.exit(proc.returncode)
    gpu = json.loads([ln for ln in proc.stdout.splitlines() if ln.startswith("{")][-1])

    t0 = time.time()
    if args.workload == "c3":
        fact, ds, _ = workloads.c3_tables()
        plans = [workloads.c3_plan(fact, ds)]
        repeats = args.cpu_repeats
        gpu_value, unit = gpu["rows_per_s"], "rows/s"
    elif args.workload == "c4":
        pairs = [workloads.c4_plan_tables(k) for k in range(256)]
        plans = [workloads.c4_plan(b, p) for b, p in pairs]
        repeats = 1  # one pass is ~10 s of 16-thread CPU work
        gpu_value, unit = gpu["batched"]["rows_per_s"], "rows/s"
    else:
        tables, meta = workloads.c5_tables(seed=1005)
        plans = [workloads.c5_plan(tables, meta, 5000 + k) for k in range(32)]
        repeats = 1  # a pass is tens of seconds of CPU work
        gpu_value, unit = gpu["served"]["rows_per_s"], "rows/s"
    log(f"{args.workload}: generated {len(plans)} plans for the CPU baseline in {time.time() - t0:.1f}s")
    cpu = None
    if not args.no_cpu_baseline:
        runs = {}
        for t in thread_counts(args.cpu_threads):
            sec, rows, per = time_oracle(plans, t, repeats)
            runs[t] = (sec, rows, per)
            log(f"{args.workload} cpu baseline: {t} threads, {sec * 1e3:.0f} ms per pass of {len(plans)} plans")
        best = min(runs, key=lambda t: runs[t][0])
        sec, rows, per = runs[best]
        cpu = {"value": rows / sec, "unit": unit, "cores": best, "kind": "port",
               "sample": f"CPU restatement of execute() on the whole workload ({len(plans)} plan(s) one after "
                         f"another, {rows} output rows); 1 cold + {repeats} timed pass(es) per thread count, median "
                         "ms per pass " + ", ".join(f"{t} threads {runs[t][0] * 1e3:.0f}" for t in runs)
                         + f"; host: {cpu_model()}, {os.cpu_count()} hardware threads, {host_cpu_share()} in this "
                           "job's share",
               "median_ms_by_threads": {str(t): runs[t][0] * 1e3 for t in runs}}
        if args.workload == "c5":
            ms = sorted(x * 1e3 for x in per)
            cpu["plan_latency_ms"] = {"p50": ms[len(ms) // 2], "max": ms[-1]}
    line = {"workload": args.workload, "gpu": gpu, "gpu_value": gpu_value, "unit": unit, "cpu_baseline": cpu,
            "gpu_over_cpu": (gpu_value / cpu["value"]) if cpu else None}
    print(json.dumps(line), flush=True)


def raw_execute(eng, plan, flags):
    """cj_execute through the bare C ABI (no Python result wrapper):
    (wall ms of the call, cj_result, cj_timing)."""
    res = abi.Result()
    tm = abi.Timing()
    desc = plan.desc()
    t0 = time.perf_counter()
    abi.check(eng.lib.cj_execute(eng.ctx, C.byref(desc), flags, C.byref(res), C.byref(tm)))
    return (time.perf_counter() - t0) * 1e3, res, tm


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--build-rows", type=int, default=4_000_000)
    ap.add_argument("--probe-rows", type=int, default=64_000_000)
    ap.add_argument("--seed", type=int, default=1002)
    ap.add_argument("--cpu-repeats", type=int, default=10)
    ap.add_argument("--cpu-threads", default="share,32",
                    help="thread counts of the CPU baseline ('share' = the host threads of this job)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile", action="store_true",
                    help="only warmup + timed device-output steps (a short program for rocprofv3)")
    ap.add_argument("--workload", default="c2", choices=["c2", "c3", "c4", "c5"],
                    help="c2 (default): the single large join, the headline line; c3/c4/c5: that workload's "
                         "harness plus its CPU baseline")
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
    if args.workload != "c2":
        other_workload(args)
        return

    global abi
    from calamarijoins_amd import abi, workloads
    from calamarijoins_amd.host import Engine, EngineResult

    t0 = time.time()
    build, probe = workloads.c2_tables(seed=args.seed, build_rows=args.build_rows, probe_rows=args.probe_rows)
    log(f"generated C2 tables in {time.time() - t0:.1f}s: build {build.num_rows} rows, probe {probe.num_rows} rows")
    plan = workloads.c2_plan(build, probe)

    eng = Engine(cache_decoded=0)
    dev_name, cus = eng.device_info()
    log(f"device {dev_name}, {cus} CUs")
    t0 = time.time()
    ids = [eng.register(build), eng.register(probe)]
    log(f"registered tables in HBM in {time.time() - t0:.1f}s")
    flags = abi.CJ_EXEC_OUTPUT_DEVICE

    if args.profile:  # a short program for the profiler: warmup, then the timed steps
        for _ in range(args.warmup):
            eng.execute(plan, flags=flags).release()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            eng.execute(plan, flags=flags).release()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        print(json.dumps({"profile_steps": args.steps, "ms_per_step": ms}), flush=True)
        eng.close()
        return

    # runs with the result copied to host: exact output bytes per column and
    # the end-to-end time of the drop-in form (result pages in host memory),
    # reported, never `value`; the first run fills the pinned output buffers
    eng.execute(plan, flags=0).release()
    e2e = []
    for _ in range(3):
        call_ms, res, tm = raw_execute(eng, plan, 0)
        t0 = time.perf_counter()
        host = EngineResult(eng, res, tm)
        wrap_ms = (time.perf_counter() - t0) * 1e3
        e2e.append((call_ms, tm.device_ms, tm.d2h_ms, wrap_ms, list(tm.phase_ms)))
        if len(e2e) < 3:
            host.release()
    out_rows = host.num_rows
    out_bytes = [page_payload_bytes(host.page_lists[i][0] if host.page_lists[i] else 0, n, t)
                 for i, (n, t) in enumerate(zip(host.num_pages(), host.types))]
    host.release()
    calls = [{"ms": round(x[0], 3), "device_ms": round(x[1], 3), "d2h_ms": round(x[2], 3),
              "phase_ms": [round(v, 3) for v in x[4]]} for x in e2e]  # in call order
    e2e.sort()
    call_ms, e2e_dev, d2h_ms, wrap_ms, e2e_phases = e2e[1]  # the median call
    log(f"host-output calls: {[round(x[0], 2) for x in e2e]} ms (device {e2e_dev:.3f}, D2H {d2h_ms:.2f}, "
        f"python result wrapper {wrap_ms:.2f} ms)")
    first = eng.execute(plan, flags=flags, timing=True)
    bytes_alg = first.timing.bytes_alg
    log(f"first run: {out_rows} rows, device {first.timing.device_ms:.3f} ms, bytes_alg {bytes_alg / 1e9:.3f} GB, "
        f"output bytes per column {out_bytes}")
    first.release()
    for _ in range(args.warmup):
        eng.execute(plan, flags=flags).release()

    # timed region: K steps, synchronous execute() calls (each ends with the
    # result pages in HBM); no timing requested, so no phase events either
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.execute(plan, flags=flags).release()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    value = out_rows / (ms_per_step / 1e3)
    # device time of a step (first kernel .. last output page, hipEvents),
    # from separate steps that ask for timing
    dev_ms, phases = [], []
    for _ in range(max(5, args.steps // 4)):
        r = eng.execute(plan, flags=flags, timing=True)
        dev_ms.append(r.timing.device_ms)
        phases.append(list(r.timing.phase_ms))
        r.release()
    dev_med = statistics.median(dev_ms)
    phase_med = [statistics.median(p[k] for p in phases) for k in range(6)]
    log(f"timed {args.steps} steps: {ms_per_step:.3f} ms/step wall, device median {dev_med:.3f} ms")

    # per-kernel durations from hipEvents on the execution stream
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
    kernels = {k: {"launches": v[0] // prof_steps, "avg_ms": round(v[1] / v[0], 4),
                   "compulsory_GBps": round(v[2] / (v[1] / 1e3) / 1e9, 1) if v[1] > 0 else None}
               for k, v in sorted(stats.items(), key=lambda kv: -kv[1][1])}
    log("kernels: " + json.dumps(kernels))

    # roofline of the dominant kernel: algorithmic bytes of one launch
    kname, (launches, tot_ms, kbytes) = max(stats.items(), key=lambda kv: kv[1][1])
    avg_ms = tot_ms / launches
    in_pages = [c.num_pages * PAGE for c in probe.columns]  # key, i32, i64, f64
    per_launch = {
        "build_pages": in_pages[0] + out_bytes[0] + out_bytes[1],
        "probe_count": in_pages[0],
        "probe_filter_w4": in_pages[1] + out_bytes[2],
        "probe_filter_w8": (in_pages[2] + out_bytes[3] + in_pages[3] + out_bytes[4]) / 2,
    }
    alg = per_launch.get(kname, kbytes / launches)
    achieved = alg / (avg_ms / 1e3) / 1e9
    traffic = pmc_traffic(SCOPE_KERNEL[kname]) if kname in SCOPE_KERNEL else None

    copy_gbps = None
    try:
        cms = Hip().copy_ms(COPY_BYTES, 5)
        copy_gbps = 2 * COPY_BYTES / (cms / 1e3) / 1e9
    except Exception as e:  # reported only
        log(f"copy ceiling unavailable: {e}")

    cpu = None
    if not args.no_cpu_baseline:
        counts = thread_counts(args.cpu_threads)
        try:
            cpu = cpu_baseline(plan, counts, args.cpu_repeats)
        except Exception as e:  # the oracle is optional on the timing box
            log(f"cpu baseline unavailable: {e}")

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
        "copy_ceiling_GBps": copy_gbps,
        "end_to_end_host_output": {
            "ms": call_ms, "device_ms": e2e_dev, "d2h_ms": d2h_ms,
            "host_ms": call_ms - e2e_dev - d2h_ms, "python_result_wrapper_ms": wrap_ms,
            "rows_per_s": out_rows / (call_ms / 1e3),
            "phase_ms": dict(zip(PHASES, e2e_phases)),
            "calls": calls,
        },
        "device_phase_ms": dict(zip(PHASES, phase_med)),
        "device": {"name": dev_name, "cus": cus},
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": kname,
            "avg_launch_ms": avg_ms,
            "alg_bytes_per_launch": alg,
        },
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
