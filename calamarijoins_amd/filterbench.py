"""GPU filter throughput (cj_filter) on a cast_info-shaped table: 36M rows at
scale 1.0 of [id, person_id, movie_id, person_role_id, note, nr_order,
role_id] (the IMDB schema's cast_info), the note strings drawn from the
role notes the JOB queries filter on, resident in HBM.  Three predicates in
the style of the JOB filters on cast_info:

  like    : note LIKE '%(producer)%'
  in_eq   : note IN ('(voice)', '(voice: Japanese version)',
                     '(voice) (uncredited)', '(voice: English version)')
            AND role_id = 2
  mixed   : note NOT LIKE '%(uncredited)%' AND nr_order IS NOT NULL
            AND (person_role_id < 1000 OR person_role_id IS NULL)

    python -m calamarijoins_amd.filterbench [--scale S] [--reps R]

prints one JSON line: per predicate the device time per call, rows kept, and
the bytes read (pages of the referenced columns) plus written (result
pages) per second; and the same predicates evaluated row by row in Python
(job.Comparison.eval, one core) on a 100k-row sample as a CPU reference.
"""
from __future__ import annotations

import argparse
import json
import statistics
import sys
import time

from . import abi, job, workloads
from .host import Engine

NOTES = ["(producer)", "(executive producer)", "(co-producer)", "(voice)", "(voice: Japanese version)",
         "(voice) (uncredited)", "(voice: English version)", "(writer)", "(uncredited)", "(as Jack Smith)",
         "(archive footage)", "(segment \"Part 2\")", "(story)", "(novel)", "(head writer)", "(screenplay)",
         "(producer) (uncredited)", "(associate producer)"]

C, L = job.Comparison, job.Logical
PREDICATES = {
    "like": C(4, "LIKE", "%(producer)%"),
    "in_eq": L("AND", [L("OR", [L("OR", [L("OR", [C(4, "=", "(voice)"), C(4, "=", "(voice: Japanese version)")]),
                                         C(4, "=", "(voice) (uncredited)")]),
                                C(4, "=", "(voice: English version)")]),
                       C(6, "=", 2)]),
    "mixed": L("AND", [L("AND", [C(4, "NOT LIKE", "%(uncredited)%"), L("NOT", [C(5, "IS NULL", None)])]),
                       L("OR", [C(3, "<", 1000), C(3, "IS NULL", None)])]),
}


def make_table(rows: int):
    g = workloads
    return g.generate(rows, 77, [
        g.col(g.GEN_PERM, "INT32", 1, 0.0, domain=rows, perm_stream=5),
        g.col(g.GEN_UNIFORM, "INT32", 2, 0.0, lo=0, hi=4_167_491),
        g.col(g.GEN_UNIFORM, "INT32", 3, 0.0, lo=0, hi=2_528_312),
        g.col(g.GEN_UNIFORM, "INT32", 4, 0.3, lo=0, hi=3_140_339),
        g.col(g.GEN_VARCHAR_VOCAB, "VARCHAR", 5, 0.5, words=NOTES),
        g.col(g.GEN_UNIFORM, "INT32", 6, 0.5, lo=0, hi=100),
        g.col(g.GEN_UNIFORM, "INT32", 7, 0.0, lo=1, hi=12),
    ])


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--reps", type=int, default=5)
    args = ap.parse_args(argv)
    rows = int(36_244_344 * args.scale)
    t0 = time.time()
    table = make_table(rows)
    types = [c.type for c in table.columns]
    print(f"generated {rows} rows in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    eng = Engine(cache_decoded=1, kernel_timing=1)
    eng.register(table)
    out = {"workload": f"cast_info-shaped table, {rows} rows, resident in HBM, result pages in HBM", "predicates": {}}
    for name, st in PREDICATES.items():
        prog = job.filter_program(st, types)
        used = sorted({n.column for n in prog if n.kind == abi.CJ_F_CMP})
        in_bytes = sum(table.columns[c].num_pages * abi.PAGE_SIZE for c in used)
        r = eng.filter(table, prog, abi.CJ_EXEC_OUTPUT_DEVICE)  # warm: decoded columns cached
        kept = r.num_rows
        r.release()
        eng.kernel_stats_reset()
        dev, out_bytes = [], 0
        for _ in range(args.reps):
            r = eng.filter(table, prog, abi.CJ_EXEC_OUTPUT_DEVICE, timing=True)
            dev.append(r.timing.device_ms)
            out_bytes = sum(r.num_pages()) * abi.PAGE_SIZE
            r.release()
        ks = {k: round(v[1] / v[0], 4) for k, v in eng.kernel_stats().items()}
        ms = statistics.median(dev)
        # CPU reference: the same statement row by row on a sample
        sample = min(rows, 100_000)
        import numpy as np  # noqa: F401
        cols = []
        for c in table.columns:
            cols.append(_decode(c, sample))
        recs = list(zip(*cols))
        t1 = time.perf_counter()
        n_cpu = sum(1 for rec in recs if st.eval(rec, types))
        cpu_s = time.perf_counter() - t1
        out["predicates"][name] = {
            "rows_kept": kept, "device_ms": ms, "rows_per_s": rows / (ms / 1e3),
            "input_bytes": in_bytes, "output_bytes": out_bytes,
            "GBps": (in_bytes + out_bytes) / (ms / 1e3) / 1e9, "kernels_ms": ks,
            "cpu_python_rows_per_s": sample / cpu_s, "cpu_python_sample_kept": n_cpu, "cpu_sample_rows": sample,
        }
        print(f"{name}: {kept} rows, {ms:.3f} ms", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)
    eng.close()


def _decode(col, limit):
    """First `limit` rows of a column as Python values (host decode of the
    pages, for the CPU reference only)."""
    import struct
    out = []
    for p in range(col.num_pages):
        page = col.pages[p].tobytes()
        n, nv = struct.unpack_from("<HH", page, 0)
        if col.type == abi.CJ_VARCHAR and n >= 0xfffe:
            raise ValueError("long strings are not generated here")
        bm = page[8192 - (n + 7) // 8:]
        vi = 0
        offs = [struct.unpack_from("<H", page, 4 + 2 * i)[0] for i in range(nv)] if col.type == abi.CJ_VARCHAR else None
        for r in range(n):
            if not (bm[r >> 3] >> (r & 7)) & 1:
                out.append(None)
            elif col.type == abi.CJ_INT32:
                out.append(struct.unpack_from("<i", page, 4 + 4 * vi)[0])
                vi += 1
            elif col.type == abi.CJ_VARCHAR:
                start = offs[vi - 1] if vi else 0
                base = 4 + 2 * nv
                out.append(page[base + start: base + offs[vi]])
                vi += 1
            else:
                out.append(struct.unpack_from("<q", page, 8 + 8 * vi)[0])
                vi += 1
            if len(out) >= limit:
                return out
    return out


if __name__ == "__main__":
    main()
