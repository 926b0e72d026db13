"""Summaries of rocprofv3 CSV output for the committed profiles.

    python -m fastmulticolsort_amd.rocprof_summary stats DIR CMD > kernel_stats.md
    python -m fastmulticolsort_amd.rocprof_summary traffic FETCH_DIR WRITE_DIR CMD_F CMD_W > traffic.json
    python -m fastmulticolsort_amd.rocprof_summary roofline TRACE_DIR FETCH_DIR WRITE_DIR BENCH_JSONL \
        CMD_T CMD_F CMD_W > kernels.md
    python -m fastmulticolsort_amd.rocprof_summary counters KERNELS DIR [DIR ...] > counters.md

`stats` reads DIR/*kernel_stats.csv (--kernel-trace --stats), or the
rocpd database DIR/*.db that rocprofv3 writes by default, and prints a
markdown table of calls / total / average per kernel.  `traffic` reads the
per-dispatch FETCH_SIZE and WRITE_SIZE counters (KiB) of two --pmc runs and
scales each by what the same counter reads for the plain device copy the
run includes (bench.py --copy-probe: 256 MiB read + 256 MiB written by
hipMemcpy's copy kernel); the result is HBM bytes per dispatch by kernel.
`roofline` joins a kernel trace and the two counter runs of a
`kernel_bench --copy-probe` command with the algorithmic bytes that the
command printed (`bytes_in_process` / `calls_in_process` of every kernel
row): per kernel, the traced average duration, algorithmic bytes per
launch, their ratio against the 8 TB/s HBM peak, and the counted HBM bytes
per dispatch.  There the probe is the process's first copy dispatch.
`counters` reads the --pmc databases of one or more runs (each counter
summed over its instances, averaged over a kernel's dispatches) and prints,
for the kernels whose names contain one of the comma-separated KERNELS, the
shares of wave time the SQ counters give: VALU / LDS instruction activity
and waits, per-wave VALU instructions, LDS bank-conflict cycles per LDS
instruction.
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


def _pmc(d, with_id=False):
    """Per kernel: (grid size, counter value) per dispatch, or (dispatch id,
    value) with `with_id`."""
    per = defaultdict(list)
    csvs = sorted(glob.glob(os.path.join(d, "*counter_collection.csv")))
    if csvs:
        for r in csv.DictReader(open(csvs[0])):
            first = int(r["Dispatch_Id"]) if with_id else int(r["Grid_Size"])
            per[short(r["Kernel_Name"])].append((first, float(r["Counter_Value"])))
        return per
    db = _db(d)
    if db is None:
        raise SystemExit(f"no counter collection under {d}")
    # one row per dispatch and counter; a counter summed over its instances
    col = "dispatch_id" if with_id else "grid_size"
    for name, first, val in db.execute(f"select kernel_name, {col}, sum(value) from counters_collection "
                                       "group by dispatch_id, counter_name"):
        per[short(name)].append((int(first), float(val)))
    return per


def base(name):
    return short(name).split("<")[0].strip()


# The library labels some launches by their role; the kernel each runs
LABEL_KERNEL = {"k_onesweep_msd": "k_onesweep", "k_onesweep_fuse": "k_onesweep", "k_tile_hist": "kd_seg_hist",
                "k_seg_hist": "kd_seg_hist", "k_tile_bucket_2t": "k_tile_bucket", "k_hash_count": "k_hash_local"}


def _algorithmic(jsonl):
    """Launches and algorithmic bytes per kernel over the process, from the
    kernel rows of a kernel_bench output."""
    calls, nbytes = defaultdict(int), defaultdict(int)

    def walk(o):
        if isinstance(o, dict):
            if "calls_in_process" in o and "bytes_in_process" in o:
                return True
            for k, v in o.items():
                if k == "kernels" and isinstance(v, dict):
                    for kn, row in v.items():
                        kn = LABEL_KERNEL.get(kn, kn)
                        calls[kn] += int(row["calls_in_process"])
                        nbytes[kn] += int(row["bytes_in_process"])
                else:
                    walk(v)
        elif isinstance(o, list):
            for v in o:
                walk(v)

    for line in open(jsonl):
        line = line.strip()
        if line.startswith("{"):
            walk(json.loads(line))
    return calls, nbytes


def roofline(tdir, fdir, wdir, jsonl, cmd_t, cmd_f, cmd_w, peak_gbs=8000.0):
    calls, nbytes = _algorithmic(jsonl)
    dur, cnt = defaultdict(float), defaultdict(int)
    for name, c, tot, _, _ in kernel_rows(tdir):
        dur[base(name)] += tot
        cnt[base(name)] += c
    f, w = _pmc(fdir, True), _pmc(wdir, True)
    fc = min(f.get(COPY_KERNEL, [(0, 0.0)]))[1]
    wc = min(w.get(COPY_KERNEL, [(0, 0.0)]))[1]
    f_scale = COPY_BYTES / (fc * 1024) if fc else None
    w_scale = COPY_BYTES / (wc * 1024) if wc else None
    fb, wb = defaultdict(list), defaultdict(list)
    for k, v in f.items():
        fb[base(k)] += [x for _, x in v]
    for k, v in w.items():
        wb[base(k)] += [x for _, x in v]
    out = [f"Kernel trace: `{cmd_t}`", "", f"FETCH_SIZE: `{cmd_f}`", "", f"WRITE_SIZE: `{cmd_w}`", "",
           f"Counter scale from the first plain copy dispatch (256 MiB read + 256 MiB written): FETCH_SIZE read "
           f"{fc:.0f} KiB (scale {f_scale or 0:.3f}), WRITE_SIZE read {wc:.0f} KiB (scale {w_scale or 0:.3f}).",
           "",
           "Per kernel (all template instances together): traced calls and average duration; algorithmic "
           "bytes per launch from the library's accounting (`calls` there must equal the traced calls); "
           "`frac` = algorithmic bytes / average duration / 8 TB/s; HBM = scaled FETCH_SIZE + WRITE_SIZE per "
           "dispatch; `HBM/alg` > 1 means bytes moved beyond the algorithm's.", "",
           "| kernel | traced calls | library calls | avg us | alg MB/launch | GB/s | frac | HBM MB/dispatch | HBM/alg |",
           "|---|---|---|---|---|---|---|---|---|"]
    for k in sorted(dur, key=lambda x: -dur[x]):
        if k == COPY_KERNEL and k not in calls:
            label = k + " (copies)"
        else:
            label = k
        avg = dur[k] / cnt[k]
        alg = nbytes[k] / calls[k] if calls.get(k) else None
        gbs = alg / avg if alg else None  # bytes per ns = GB/s
        hbm = None
        if fb.get(k) and wb.get(k) and f_scale and w_scale:
            hbm = (sum(fb[k]) / len(fb[k]) * f_scale + sum(wb[k]) / len(wb[k]) * w_scale) * 1024
        out.append("| {} | {} | {} | {:.1f} | {} | {} | {} | {} | {} |".format(
            label, cnt[k], calls.get(k, "—"), avg / 1e3,
            f"{alg / 1e6:.1f}" if alg else "—", f"{gbs:.0f}" if gbs else "—",
            f"{gbs / peak_gbs:.3f}" if gbs else "—", f"{hbm / 1e6:.1f}" if hbm else "—",
            f"{hbm / alg:.2f}" if hbm and alg else "—"))
    return "\n".join(out)


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


def counter_means(dirs):
    """{kernel: {counter: mean over dispatches of the instance sum}} over the
    rocpd databases of several --pmc runs."""
    out = defaultdict(dict)
    for d in dirs:
        db = _db(d)
        if db is None:
            raise SystemExit(f"no counter database under {d}")
        for name, cname, total, disp in db.execute(
                "select kernel_name, counter_name, sum(value), count(distinct dispatch_id) "
                "from counters_collection group by kernel_name, counter_name"):
            out[short(name)][cname] = float(total) / max(1, disp)
    return out


def counters(kernels, dirs):
    agg = counter_means(dirs)
    wanted = [k.strip() for k in kernels.split(",") if k.strip()]
    out = ["| kernel | waves | VALU instr / wave | VALU active % | LDS active % | waiting for LDS issue % | "
           "waiting (any) % | waiting for any issue % | LDS bank-conflict cycles / LDS instr |",
           "|---|---|---|---|---|---|---|---|---|"]
    for name, a in sorted(agg.items()):
        if not any(w in name for w in wanted) or "SQ_WAVE_CYCLES" not in a:
            continue
        wc = a["SQ_WAVE_CYCLES"] or 1.0
        pct = lambda c: 100.0 * a.get(c, 0.0) / wc  # noqa: E731
        waves = a.get("SQ_WAVES", 0.0)
        vpw = a.get("SQ_INSTS_VALU", 0.0) / waves if waves else float("nan")
        bc = a.get("SQ_LDS_BANK_CONFLICT", 0.0) / a["SQ_INSTS_LDS"] if a.get("SQ_INSTS_LDS") else float("nan")
        out.append(f"| {name} | {waves:.0f} | {vpw:.0f} | {pct('SQ_ACTIVE_INST_VALU'):.1f} | "
                   f"{pct('SQ_ACTIVE_INST_LDS'):.1f} | {pct('SQ_WAIT_INST_LDS'):.1f} | {pct('SQ_WAIT_ANY'):.1f} | "
                   f"{pct('SQ_WAIT_INST_ANY'):.1f} | {bc:.2f} |")
    return "\n".join(out)


def main(argv):
    if argv[0] == "counters":
        print(counters(argv[1], argv[2:]))
    elif argv[0] == "stats":
        print(stats(argv[1], argv[2]))
    elif argv[0] == "traffic":
        print(json.dumps(traffic(argv[1], argv[2], argv[3], argv[4]), indent=1))
    elif argv[0] == "roofline":
        print(roofline(*argv[1:8]))
    else:
        raise SystemExit(__doc__)


if __name__ == "__main__":
    main(sys.argv[1:])
