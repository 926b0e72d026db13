"""Summaries of rocprofv3 CSV output for the committed profiles.

    python -m fastmulticolsort_amd.rocprof_summary stats DIR CMD > kernel_stats.md
    python -m fastmulticolsort_amd.rocprof_summary traffic FETCH_DIR WRITE_DIR CMD_F CMD_W > traffic.json

`stats` reads DIR/*kernel_stats.csv (--kernel-trace --stats), or the
rocpd database DIR/*.db that rocprofv3 writes by default, and prints a
markdown table of calls / total / average per kernel.  `traffic` reads the
per-dispatch FETCH_SIZE and WRITE_SIZE counters (KiB) of two --pmc runs and
scales each by what the same counter reads for the plain device copy the
run includes (bench.py --copy-probe: 256 MiB read + 256 MiB written by
hipMemcpy's copy kernel); the result is HBM bytes per dispatch by kernel.
"""
import csv
import glob
import json
import os
import re
import sqlite3
import sys
from collections import defaultdict

COPY_BYTES = 1 << 28
COPY_KERNEL = "__amd_rocclr_copyBuffer"


def short(name):
    name = re.sub(r"\(.*", "", name)
    return name.replace("void ", "").replace("fmcs::", "")


def _one(pattern):
    files = sorted(glob.glob(pattern))
    if not files:
        raise SystemExit(f"no file matches {pattern}")
    return files[0]


def _db(d):
    files = sorted(glob.glob(os.path.join(d, "*.db")))
    return sqlite3.connect(files[0]) if files else None


def kernel_rows(d):
    """(name, calls, total ns, average ns, percent) per kernel, CSV or rocpd."""
    csvs = sorted(glob.glob(os.path.join(d, "*kernel_stats.csv")))
    if csvs:
        return [(r["Name"], int(r["Calls"]), float(r["TotalDurationNs"]), float(r["AverageNs"]),
                 float(r["Percentage"])) for r in csv.DictReader(open(csvs[0]))]
    db = _db(d)
    if db is None:
        raise SystemExit(f"no kernel statistics under {d}")
    # the database's durations are nanoseconds per dispatch
    rows = db.execute("select name, count(*), sum(duration), avg(duration) from kernels group by name "
                      "order by sum(duration) desc").fetchall()
    tot = sum(r[2] for r in rows) or 1.0
    return [(n, c, float(t), float(a), 100.0 * t / tot) for n, c, t, a in rows]


def stats(d, cmd):
    out = [f"Command: `{cmd}`", "", "| kernel | calls | total ms | avg us | % of GPU time |", "|---|---|---|---|---|"]
    for name, calls, tot, avg, pct in kernel_rows(d):
        out.append(f"| {short(name)} | {calls} | {tot / 1e6:.3f} | {avg / 1e3:.1f} | {pct:.2f} |")
    return "\n".join(out)


def _pmc(d):
    per = defaultdict(list)
    csvs = sorted(glob.glob(os.path.join(d, "*counter_collection.csv")))
    if csvs:
        for r in csv.DictReader(open(csvs[0])):
            per[short(r["Kernel_Name"])].append((int(r["Grid_Size"]), float(r["Counter_Value"])))
        return per
    db = _db(d)
    if db is None:
        raise SystemExit(f"no counter collection under {d}")
    # one row per dispatch and counter; a counter summed over its instances
    for name, grid, val in db.execute("select kernel_name, grid_size, sum(value) from counters_collection "
                                      "group by dispatch_id, counter_name"):
        per[short(name)].append((int(grid), float(val)))
    return per


def traffic(fdir, wdir, cmd_f, cmd_w):
    f, w = _pmc(fdir), _pmc(wdir)
    # the copy probe is the largest plain copy dispatch of the run
    fc = max(v for g, v in f.get(COPY_KERNEL, [(0, 0.0)]))
    wc = max(v for g, v in w.get(COPY_KERNEL, [(0, 0.0)]))
    f_scale = COPY_BYTES / (fc * 1024) if fc else None
    w_scale = COPY_BYTES / (wc * 1024) if wc else None
    kern = {}
    for k in sorted(set(f) | set(w)):
        if k == COPY_KERNEL:
            continue
        fv = [v for _, v in f.get(k, [])]
        wv = [v for _, v in w.get(k, [])]
        fb = sum(fv) / len(fv) * 1024 * f_scale if fv and f_scale else None
        wb = sum(wv) / len(wv) * 1024 * w_scale if wv and w_scale else None
        kern[k] = {"dispatches": len(fv), "fetch_bytes": fb, "write_bytes": wb,
                   "hbm_bytes": (fb + wb) if fb is not None and wb is not None else None}
    return {"fetch_command": cmd_f, "write_command": cmd_w,
            "copy_probe": {"bytes_read": COPY_BYTES, "bytes_written": COPY_BYTES,
                           "fetch_kib_raw": fc, "write_kib_raw": wc,
                           "fetch_scale": f_scale, "write_scale": w_scale},
            "kernels": kern}


def main(argv):
    if argv[0] == "stats":
        print(stats(argv[1], argv[2]))
    elif argv[0] == "traffic":
        print(json.dumps(traffic(argv[1], argv[2], argv[3], argv[4]), indent=1))
    else:
        raise SystemExit(__doc__)


if __name__ == "__main__":
    main(sys.argv[1:])
