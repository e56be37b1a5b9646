""".tbl cache loading: the C2 tables (4M build x 64M probe rows) written as
.tbl files, then brought into HBM two ways:

  mmap + register : cj_tbl_open (mmap with MAP_POPULATE, the reference's
                    Table::from_cache) then cj_table_register (pages copied
                    from the mapping through pinned buffers to HBM)
  stream          : cj_tbl_load (pages read from the file with parallel
                    preads straight into pinned buffers, DMA to HBM)

    python -m calamarijoins_amd.tblload [--dir D] [--reps R]

prints one JSON line with the bytes, times and GB/s of each path and the
C2 result row count over the loaded tables.  The files were just written,
so they are read from the page cache.
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import statistics
import sys
import tempfile
import time

from . import workloads
from .host import Engine, read_tbl, write_tbl


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--dir", default=None, help="directory for the .tbl files (default: a temporary one)")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--scale", type=float, default=1.0)
    args = ap.parse_args(argv)
    d = args.dir or tempfile.mkdtemp(prefix="cj_tbl_")
    os.makedirs(d, exist_ok=True)
    try:
        b, p = workloads.c2_tables(build_rows=int(workloads.C2_BUILD * args.scale),
                                   probe_rows=int(workloads.C2_PROBE * args.scale))
        paths = [os.path.join(d, "c2_build.tbl"), os.path.join(d, "c2_probe.tbl")]
        t0 = time.perf_counter()
        for t, path in zip((b, p), paths):
            write_tbl(path, t)
        write_s = time.perf_counter() - t0
        nbytes = sum(os.path.getsize(x) for x in paths)
        print(f"wrote {nbytes / 1e9:.2f} GB in {write_s:.2f} s", file=sys.stderr, flush=True)
        del b, p
        eng = Engine(cache_decoded=1)
        mm, st = [], []
        rows = None
        for _ in range(args.reps):
            t0 = time.perf_counter()
            tabs = [read_tbl(x) for x in paths]
            ids = [eng.register(t) for t in tabs]
            mm.append(time.perf_counter() - t0)
            for i in ids:
                eng.unregister(i)
            for t in tabs:
                t.close()
            t0 = time.perf_counter()
            loaded = [eng.load_tbl(x) for x in paths]
            st.append(time.perf_counter() - t0)
            if rows is None:
                r = eng.execute(workloads.c2_plan(*loaded), flags=1)
                rows = r.num_rows
                r.release()
            for t in loaded:
                t.close()
        out = {
            "workload": f"C2 tables ({workloads.C2_BUILD * args.scale:.0f} build + {workloads.C2_PROBE * args.scale:.0f} "
                        "probe rows) as .tbl files in the page cache",
            "bytes": nbytes, "write_s": write_s, "write_GBps": nbytes / write_s / 1e9,
            "mmap_register_s": statistics.median(mm), "mmap_register_GBps": nbytes / statistics.median(mm) / 1e9,
            "stream_load_s": statistics.median(st), "stream_load_GBps": nbytes / statistics.median(st) / 1e9,
            "c2_rows_over_loaded_tables": rows, "reps": args.reps,
        }
        print(json.dumps(out), flush=True)
        eng.close()
    finally:
        if not args.dir:
            shutil.rmtree(d, ignore_errors=True)


if __name__ == "__main__":
    main()
