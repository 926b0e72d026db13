"""Benchmark: ORDER BY c1, c2 over 2^27 rows with 2^20 Zipf(1.1)-sized c1
groups (c1: 20-bit code in a u32 bank, c2: 40-bit uniform in a u64 bank).

One step = one full two-round sort on the GPU, inputs resident in HBM:
identity order, round-1 LSD radix sort of c1 with the oid payload (c2
travelling with it, packed beside c1 in u64 composite keys, so no oid
gather of c2 is needed), group boundaries, round-2 batched segmented sort of
c2 within the c1 groups (the reference's Composer::multiRoundSorting,
src/composer.cpp:48-135).  The timed steps replay the device-driven round
as one hipGraph (--mode graph, the default); per-kernel HIP-event times come
from a second pass over the same kernels launched eagerly.  The
host-synchronised batched form, the plain per-round form (one sort dispatch
per group, src/composer.cpp:1285-1323) and the single unbinned segmented
sort are timed beside it.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per dispatch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
# (scaled by a plain copy's readings), by kernel; written by
# fastmulticolsort_amd.rocprof_summary from the runs named inside it
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "config3_2p27_traffic.json")
# the bench's kernel timer names -> the kernel instance this workload runs
PROFILE_NAME = {"k_gather": "k_gather<unsigned long>", "k_tile_bucket": "k_tile_bucket<unsigned int, 8>",
                "k_tile_bucket_2t": "k_tile_bucket<unsigned int, 8>",
                "k_onesweep": "k_onesweep<unsigned long, 20, 8, NoPay, false, true, unsigned long>",
                "k_onesweep_fuse": "k_onesweep<unsigned int, 16, 8, unsigned long, true, true, unsigned int>",
                "k_onesweep_msd": "k_onesweep<unsigned long, 20, 8, NoPay, false, false, unsigned int>",
                "k_bwrite": "k_bwrite<unsigned int, 32>", "k_bcount": "k_bcount<unsigned int, 32>",
                "k_seg_hist": "kd_seg_hist<unsigned int, 20, 8>"}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter runs, or
    None when they do not cover it."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    k = t.get("kernels", {}).get(PROFILE_NAME.get(kernel, kernel))
    return k.get("hbm_bytes") if k else None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def check_sorted(c1, c2, order):
    """Size-independent properties: order is a permutation and the rows are
    in (c1, c2, oid) order."""
    n = len(order)
    seen = np.zeros(n, dtype=bool)
    seen[order] = True
    if not seen.all():
        return False
    a, b = c1[order].astype(np.int64), c2[order]
    o = order.astype(np.int64)
    da = np.diff(a)
    if (da < 0).any():
        return False
    eq = da == 0
    bb = b[1:] < b[:-1]
    if (eq & bb).any():
        return False
    tie = eq & (b[1:] == b[:-1])
    return not (tie & (o[1:] < o[:-1])).any()


def cpu_info():
    """CPU model and core counts of the host the baseline runs on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        share = len(os.sched_getaffinity(0))
    except Exception:
        share = None
    return {"cpu_model": model, "nproc": os.cpu_count(), "cores_in_affinity": share}


WARMUPS, RUNS = 5, 5  # the reference's practice (src/stitch_composer.cpp:4214-4223), median of 5


def cpu_baseline(c1, c2, sample_rows, full=True):
    """The reference's own AVX2 round loop (oracle/_ref, built from the
    reference sources), one pinned thread: 5 warm-ups and the median of 5
    runs on the first `sample_rows` rows, and one run on all the rows
    beside it.  Falls back to the C restatement (oracle) where the
    reference build is absent."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob
    kind = "reference" if ob.ref() is not None else "port"

    def run(a, b):
        t0 = time.perf_counter()
        if kind == "reference":
            ob.ref_sort_rounds([a, b], [20, 40])
        else:
            ob.multi_round_sort([a, b])
        return time.perf_counter() - t0

    s1 = np.ascontiguousarray(c1[:sample_rows])
    s2 = np.ascontiguousarray(c2[:sample_rows])
    try:
        old = os.sched_getaffinity(0)
        os.sched_setaffinity(0, {sorted(old)[0]})
    except Exception:
        old = None
    try:
        for _ in range(WARMUPS):
            run(s1, s2)
        t = float(np.median([run(s1, s2) for _ in range(RUNS)]))
        t_full = run(np.ascontiguousarray(c1), np.ascontiguousarray(c2)) if full and len(c1) > sample_rows else None
    finally:
        if old is not None:
            os.sched_setaffinity(0, old)
    out = {"value": sample_rows / t / 1e6, "unit": "Mrows/s", "cores": 1, "kind": kind,
           "sample": f"first {sample_rows} rows of the same c1/c2 columns, 2 rounds (c1 u32, c2 u64), "
                     f"{WARMUPS} warm-ups + median of {RUNS}, one pinned thread"}
    if t_full is not None:
        out["full_size"] = {"rows": len(c1), "value": len(c1) / t_full / 1e6, "unit": "Mrows/s",
                            "runs": 1, "seconds": round(t_full, 2)}
    out.update(cpu_info())
    return out


def cpu_baseline_all_cores(c1, c2, sample_rows, max_threads=16):
    """The same round loop on every core of this process's share (at most
    16): the reference's kernels under a parallel driver (oracle/_ref), the
    all-core comparator; None where the reference build is absent."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob
    if ob.ref() is None:
        return None
    try:
        cores = len(os.sched_getaffinity(0))
    except Exception:
        cores = os.cpu_count() or 1
    nt = max(1, min(max_threads, cores))
    s1 = np.ascontiguousarray(c1[:sample_rows])
    s2 = np.ascontiguousarray(c2[:sample_rows])
    times = []
    for it in range(WARMUPS + RUNS):
        t0 = time.perf_counter()
        ob.ref_sort_rounds_mt([s1, s2], [20, 40], nt)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times[WARMUPS:]))
    return {"value": sample_rows / t / 1e6, "unit": "Mrows/s", "cores": nt, "kind": "reference",
            "sample": f"first {sample_rows} rows of the same c1/c2 columns, 2 rounds, {nt} threads "
                      f"(radix partition + the reference's merge sort per part, groups over threads), "
                      f"{WARMUPS} warm-ups + median of {RUNS}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rows-log2", type=int, default=27)
    ap.add_argument("--groups-log2", type=int, default=20)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-comparators", action="store_true",
                    help="skip the host-synchronised batched and unbinned runs (a profile of the timed kernels alone)")
    ap.add_argument("--cpu-sample-log2", type=int, default=23)
    ap.add_argument("--no-cpu-full", action="store_true", help="skip the CPU baseline's full-size run")
    ap.add_argument("--autotune", action="store_true",
                    help="sweep the batched sort's bin thresholds and tile shapes first and save the best")
    ap.add_argument("--mode", choices=["graph", "device", "batched"], default="graph")
    ap.add_argument("--small-max", type=int, default=None,
                    help="tuning small_max (largest segment packed into a window tile; default: the tuning file)")
    ap.add_argument("--window", type=int, default=None, help="tuning window (default: the tuning file)")
    ap.add_argument("--med-max", type=int, default=None,
                    help="tuning med_max: 0 = level-synchronous tile sorts, 1024..16384 = the persistent "
                         "size-binned launch with single segments up to this many rows (default: the tuning file)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run one plain 256 MiB device copy (the traffic reference for PMC counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              f"(the multi-GPU path is design-only)", file=sys.stderr)
        sys.exit(2)

    import torch
    from fastmulticolsort_amd import datagen
    from fastmulticolsort_amd.engine import (Context, MODE_BATCHED, MODE_DEVICE, MODE_GRAPH, MODE_PLAIN,
                                             MODE_UNBINNED)

    n = 1 << args.rows_log2
    ng = 1 << args.groups_log2
    t0 = time.time()
    (c1, c2), bws = datagen.config3(n, ngroups=ng)
    log(f"generated {n} rows in {time.time() - t0:.1f}s")
    d1 = torch.from_numpy(c1).cuda()
    d2 = torch.from_numpy(c2).cuda()
    order = torch.empty(n, dtype=torch.uint32, device="cuda")
    ctx = Context(0)
    if args.copy_probe:
        a = torch.empty(1 << 28, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        log("copy probe: 268435456 bytes read + 268435456 bytes written")

    main_mode = {"graph": MODE_GRAPH, "device": MODE_DEVICE, "batched": MODE_BATCHED}[args.mode]
    prof_mode = MODE_DEVICE if main_mode == MODE_GRAPH else main_mode

    def step(mode=main_mode):
        ctx.multi_round_sort([d1, d2], bws, order=order, mode=mode, want_count=False)

    from fastmulticolsort_amd import autotune
    tuning = None
    if args.autotune:
        best, table = autotune.tune(ctx, step, log=log)
        autotune.save(best, table, f"config3 2^{args.rows_log2} rows")
        tuning = best
    else:
        tuning = autotune.load(ctx)
    if args.small_max is not None or args.window is not None or args.med_max is not None:
        t = ctx.tuning()
        sm = args.small_max if args.small_max is not None else t["small_max"]
        mm = args.med_max if args.med_max is not None else t["med_max"]
        if mm:
            sm = min(sm, 512)  # the persistent form's window tiles are 1024 rows
        w = min(args.window if args.window is not None else t["window"], sm + 1)
        ctx.set_tuning(sm, w, t["msd_bits"], med_max=mm)
        tuning = ctx.tuning()
    log(f"tuning: {tuning or ctx.tuning()}")

    for _ in range(args.warmup):
        step()
    ctx.sync()
    ok = check_sorted(c1, c2, order.cpu().numpy())
    log(f"full-size property check (permutation, (c1,c2,oid) order): {ok}")
    if not ok:
        raise SystemExit("result failed the sortedness check")

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ctx.sync()  # reports a device-detected input error, if any
    ms = (t1 - t0) * 1e3 / args.steps
    value = n / (ms / 1e3) / 1e6
    log(f"{args.mode}: {ms:.3f} ms/step, {value:.1f} Mrows/s")

    # per-kernel HIP-event times: the same kernels launched eagerly
    ctx.set_profiling(True)
    ctx.kernel_stats(reset=True)
    for _ in range(args.steps):
        step(prof_mode)
    stats = ctx.kernel_stats(reset=True)
    ctx.set_profiling(False)
    ok = check_sorted(c1, c2, order.cpu().numpy())
    if not ok:
        raise SystemExit("profiled pass failed the sortedness check")

    # dominant kernel and its roofline: algorithmic bytes per launch (for a
    # radix digit pass 2(w + 4) per row: key and oid read and written) over
    # the average launch; side outputs that later kernels of this design read
    # (a next-digit byte, the sorted c1 column) are reported apart
    dom = max(stats.items(), key=lambda kv: kv[1]["ms"] if kv[1]["bytes"] > 0 else 0.0)
    name, st = dom
    avg_ms = st["ms"] / st["launches"]
    bytes_per_launch = st["bytes"] / st["launches"]
    extra_per_launch = st.get("extra_bytes", 0.0) / st["launches"]
    achieved = bytes_per_launch / (avg_ms / 1e3) / 1e9
    kernels = {k: {"launches": v["launches"], "ms_per_step": v["ms"] / args.steps,
                   "GBs": (v["bytes"] / (v["ms"] / 1e3) / 1e9) if v["ms"] > 0 else None}
               for k, v in stats.items()}
    log("kernels: " + json.dumps(kernels))
    kernel_ms = sum(v["ms_per_step"] for v in kernels.values())

    # host-synchronised batched form (one readback per level)
    batched_ms = None
    if main_mode != MODE_BATCHED and not args.no_comparators:
        step(MODE_BATCHED)
        torch.cuda.synchronize()
        tb0 = time.perf_counter()
        for _ in range(3):
            step(MODE_BATCHED)
        torch.cuda.synchronize()
        batched_ms = (time.perf_counter() - tb0) * 1e3 / 3
        log(f"batched (host-synchronised): {batched_ms:.3f} ms")

    # comparators
    unb_ms = None
    if not args.no_comparators:
        ctx.multi_round_sort([d1, d2], bws, order=order, mode=MODE_UNBINNED)
        torch.cuda.synchronize()
        tu0 = time.perf_counter()
        ctx.multi_round_sort([d1, d2], bws, order=order, mode=MODE_UNBINNED)
        torch.cuda.synchronize()
        unb_ms = (time.perf_counter() - tu0) * 1e3
        log(f"unbinned: {unb_ms:.3f} ms")
    plain_ms = None
    if not args.no_plain and not args.no_comparators:
        tp0 = time.perf_counter()
        ctx.multi_round_sort([d1, d2], bws, order=order, mode=MODE_PLAIN)
        torch.cuda.synchronize()
        plain_ms = (time.perf_counter() - tp0) * 1e3
        log(f"plain per-round: {plain_ms:.1f} ms")

    cpu = cpu_all = None
    if not args.no_cpu:
        cpu = cpu_baseline(c1, c2, min(n, 1 << args.cpu_sample_log2), full=not args.no_cpu_full)
        log(f"cpu baseline: {cpu}")
        cpu_all = cpu_baseline_all_cores(c1, c2, min(n, 1 << args.cpu_sample_log2))
        log(f"cpu baseline, all cores: {cpu_all}")

    out = {
        "metric": "Mrows/s, batched segmented sort vs plain per-round launches; % of HBM peak",
        "value": round(value, 2),
        "unit": "Mrows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded SplitMix64 / fill_zipf counts)",
        "config": {"workload": f"ORDER BY c1,c2: 2^{args.rows_log2} rows, 2^{args.groups_log2} Zipf(1.1) c1 groups, "
                               f"c1 20-bit u32, c2 40-bit u64, batched segmented round 2",
                   "rows": n, "groups": ng},
        "roofline": {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(name),
                     "avg_launch_ms": round(avg_ms, 5), "bytes_per_launch": bytes_per_launch,
                     "extra_bytes_per_launch": extra_per_launch,
                     "frac_with_extra": round((bytes_per_launch + extra_per_launch) / (avg_ms / 1e3) / 1e9
                                              / HBM_PEAK_GBS, 4)},
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
        "comparators": {"plain_per_round_ms": plain_ms, "unbinned_ms": round(unb_ms, 3) if unb_ms else None,
                        "host_synchronised_batched_ms": (round(batched_ms, 3) if batched_ms else None),
                        "batched_vs_plain": (round(plain_ms / ms, 2) if plain_ms else None),
                        "batched_vs_unbinned": round(unb_ms / ms, 3) if unb_ms else None},
        "mode": args.mode,
        "kernel_ms_per_step": round(kernel_ms, 4),
        "tuning": tuning or ctx.tuning(),
        "kernels": kernels,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
