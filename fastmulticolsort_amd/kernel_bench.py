"""GPU measurements of the path's kernels on the other workloads, beside
bench.py's headline (the batched segmented sort, config 3).

    python -m fastmulticolsort_amd.kernel_bench [--configs 1,2,4] [--plans "16,26;42"|none]
        [--hash-types 1,0|none] [--copy-probe] > out.jsonl

Prints one JSON object per measurement:
  * config 1 (2^24 rows, c1 17-bit u32 over 8192 codes, c2 33-bit u64): every
    split of the two-round sweep (TwoRoundsExhaustive,
    src/stitch_composer.cpp:4197-4258) run end to end through the composer,
    the analytic model's pick and its regret, under the default constants
    and under constants calibrated on the device;
  * config 2 (2^26 rows, 12/20/30-bit columns stitched into one 62-bit u64
    key): the stitch, the LSD radix sort, the bitonic + merge-path sort and
    the payload gathers, with their HBM roofline fractions;
  * config 4 (2^25 rows, 2/1/12/27-bit columns): the seeded PlanEnumRRS
    plans plus the 1- and 2-round plans, end to end, and the hash GROUP BY
    with its exact group count.
Every result is checked (sorted order / group count) before it is printed.
Kernel times are HIP-event timers on the context's stream; `achieved` is
algorithmic bytes / kernel time against the 8 TB/s HBM peak.  Each kernel
row also carries `calls_in_process` and `bytes_in_process`: the launches
and algorithmic bytes of every call of the measured function in this
process (one warm-up, the timed calls, the profiled calls), which is what a
rocprofv3 trace of the same command sees (`rocprof_summary roofline` joins
the two).  `--copy-probe` starts with one plain 256 MiB device copy, the
first copy dispatch of the process, the reference for PMC traffic counters.
"""
import argparse
import json
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _emit(d):
    print(json.dumps(d), flush=True)


def _kernel_rows(stats, reps):
    out = {}
    calls = 1 + 2 * reps  # warm-up, timed and profiled calls of the function
    for k, v in stats.items():
        ms = v["ms"] / max(1, reps)
        gbs = v["bytes"] / (v["ms"] / 1e3) / 1e9 if v["ms"] > 0 and v["bytes"] > 0 else None
        out[k] = {"ms": round(ms, 4), "launches": v["launches"] // max(1, reps),
                  "GBs": round(gbs, 1) if gbs else None, "frac": round(gbs / HBM_PEAK_GBS, 4) if gbs else None,
                  "calls_in_process": v["launches"] // max(1, reps) * calls,
                  "bytes_in_process": int(v["bytes"] // max(1, reps) * calls)}
    return out


def _timed(ctx, fn, reps=3):
    """Wall time per call (ms, after a warm-up) and the per-kernel timers."""
    import torch
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    ctx.set_profiling(True)
    ctx.kernel_stats(reset=True)
    for _ in range(reps):
        fn()
    stats = ctx.kernel_stats(reset=True)
    ctx.set_profiling(False)
    return ms, _kernel_rows(stats, reps)


def config1(ctx, n=1 << 24):
    import torch
    from . import datagen, planner
    from .engine import MODE_DEVICE
    (c1, c2), bws = datagen.config1(n)
    cols = [_dev(c1), _dev(c2)]
    exp = np.lexsort((c2, c1)).astype(np.uint32)
    order = torch.empty(n, dtype=torch.uint32, device="cuda")
    rows = []
    for s in range(-bws[1], bws[0] + 1):
        split = [w for w in (bws[0] - s, bws[1] + s) if w > 0]
        ms, _ = _timed(ctx, lambda: planner.run_plan(ctx, cols, bws, [0, 1], split, mode=MODE_DEVICE, order=order),
                       reps=2)
        ok = bool(np.array_equal(order.cpu().numpy(), exp))
        rows.append({"split": split, "ms": round(ms, 3), "ok": ok})
        print(f"config 1 split {split}: {ms:.3f} ms ok={ok}", file=sys.stderr, flush=True)
    # the analytic pick under the default constants and under constants
    # measured on this device (fmcs_calibrate)
    card = [8192, datagen.exact_cardinality(c2)]
    _, pick, _, _ = planner.optimize(planner.default_cost_model(), n, bws, card)
    cal = planner.calibrate(ctx)
    _, cpick, _, _ = planner.optimize(cal, n, bws, card)
    best = min(rows, key=lambda r: r["ms"])
    picked = next((r for r in rows if r["split"] == pick), None)
    cpicked = next((r for r in rows if r["split"] == cpick), None)
    _emit({"config": 1, "what": "two-round sweep, every split end to end (composer, device-driven batched)",
           "rows": n, "plans": len(rows), "all_ok": all(r["ok"] for r in rows),
           "best": best, "worst": max(rows, key=lambda r: r["ms"]),
           "best_Mrows_s": round(n / best["ms"] / 1e3, 1),
           "analytic_pick": pick, "analytic_pick_ms": picked["ms"] if picked else None,
           "analytic_regret": round(picked["ms"] / best["ms"] - 1, 4) if picked else None,
           "calibrated_pick": cpick, "calibrated_pick_ms": cpicked["ms"] if cpicked else None,
           "calibrated_regret": round(cpicked["ms"] / best["ms"] - 1, 4) if cpicked else None,
           "calibrated_model": cal.to_dict(), "sweep": rows})


def config2(ctx, n=1 << 26):
    import torch
    from . import datagen
    from .engine import construct_entries
    cols, bws = datagen.config2(n)
    src = [_dev(c) for c in cols]
    ents = construct_entries(bws, [0, 1, 2], [62])
    key = torch.empty(n, dtype=torch.uint64, device="cuda")
    ms, kern = _timed(ctx, lambda: ctx.stitch(src, bws, ents, [key]))
    sbytes = n * (2 + 4 + 4 + 8)
    _emit({"config": 2, "what": "stitch 12|20|30 -> one 62-bit u64 key", "rows": n, "ms": round(ms, 3),
           "achieved_GBs": round(sbytes / (ms / 1e3) / 1e9, 1), "frac": round(sbytes / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
           "kernels": kern})
    k_host = ((cols[0].astype(np.uint64) << np.uint64(50)) | (cols[1].astype(np.uint64) << np.uint64(30))
              | cols[2].astype(np.uint64))
    assert np.array_equal(key.cpu().numpy(), k_host), "stitch mismatch"
    exp = np.argsort(k_host, kind="stable").astype(np.uint32)
    iota = _dev(np.arange(n, dtype=np.uint32))
    kk = torch.empty_like(key)
    oo = torch.empty_like(iota)

    def lsd():
        kk.copy_(key)
        oo.copy_(iota)
        ctx.sort_pairs(kk, oo, 62)

    def merge():
        kk.copy_(key)
        oo.copy_(iota)
        ctx.sort_pairs_merge(kk, oo)

    res = {}
    for name, fn in (("lsd_radix", lsd), ("bitonic_merge_path", merge)):
        ms, kern = _timed(ctx, fn)
        ok = bool(np.array_equal(oo.cpu().numpy(), exp))
        res[name] = {"ms": round(ms, 3), "Mrows_s": round(n / ms / 1e3, 1), "ok": ok, "kernels": kern}
    _emit({"config": 2, "what": "62-bit single-round sort of (key, oid): LSD radix vs bitonic tiles + merge path",
           "rows": n, **res})
    # oid payload gather of the three source columns in sorted order: one
    # column gather each, and all three through packed 16-byte row records
    outs = [torch.empty_like(c) for c in src]
    ms, kern = _timed(ctx, lambda: [ctx.gather(c, oo, out=o) for c, o in zip(src, outs)])
    ok = all(np.array_equal(o.cpu().numpy(), c[exp]) for o, c in zip(outs, cols))
    gbytes = n * (3 * 4 + 2 * (2 + 4 + 4))
    _emit({"config": 2, "what": "payload gather of the 3 source columns by the sorted oids, column by column",
           "rows": n, "ms": round(ms, 3), "ok": ok, "achieved_GBs": round(gbytes / (ms / 1e3) / 1e9, 1),
           "kernels": kern})
    outs = [torch.empty_like(c) for c in src]
    ms, kern = _timed(ctx, lambda: ctx.gather_columns(src, oo, outs=outs))
    ok = all(np.array_equal(o.cpu().numpy(), c[exp]) for o, c in zip(outs, cols))
    # pack: 10 B read + 16 B written a row; gather: 4 B order + 16 B record + 10 B written
    rbytes = n * (10 + 16 + 4 + 16 + 10)
    _emit({"config": 2, "what": "payload gather of the 3 source columns by the sorted oids, packed row records",
           "rows": n, "ms": round(ms, 3), "ok": ok, "achieved_GBs": round(rbytes / (ms / 1e3) / 1e9, 1),
           "kernels": kern})
    # the columns' declared widths (12 + 20 + 30 bits) in one 8-byte bit record
    outs = [torch.empty_like(c) for c in src]
    ms, kern = _timed(ctx, lambda: ctx.gather_columns(src, oo, outs=outs, bits=list(bws)))
    ok = all(np.array_equal(o.cpu().numpy(), c[exp]) for o, c in zip(outs, cols))
    # pack: 10 B read + 8 B written a row; gather: 4 B order + 8 B record + 10 B written
    rbytes = n * (10 + 8 + 4 + 8 + 10)
    _emit({"config": 2, "what": "payload gather of the 3 source columns by the sorted oids, 8-byte bit records",
           "rows": n, "ms": round(ms, 3), "ok": ok, "achieved_GBs": round(rbytes / (ms / 1e3) / 1e9, 1),
           "kernels": kern})


def config4(ctx, n=1 << 25, max_rrs=24, plans=None, hash_types=(1, 0)):
    import torch
    from . import datagen, planner
    from .engine import MODE_DEVICE
    cols, bws = datagen.config4(n)
    dcols = [_dev(c) for c in cols]
    stacked = [c.astype(np.uint64) for c in cols]
    exp = np.lexsort(tuple(reversed(stacked))).astype(np.uint32)
    model = planner.default_cost_model()
    card = [datagen.exact_cardinality(c) for c in cols]
    if plans is None:
        plans = [[42]] + planner.split_exhaustive(42, 2, 64) + \
            planner.plan_enum_rrs(model, n, bws, card, [0, 1, 2, 3], 6, 1)[:max_rrs]
    seen, uniq = set(), []
    for p in plans:
        t = tuple(int(w) for w in p)
        if t not in seen:
            seen.add(t)
            uniq.append(list(t))
    order = torch.empty(n, dtype=torch.uint32, device="cuda")
    rows = []
    for split in uniq:
        ms, kern = _timed(ctx, lambda: planner.run_plan(ctx, dcols, bws, [0, 1, 2, 3], split, mode=MODE_DEVICE,
                                                        order=order), reps=2)
        rows.append({"split": split, "ms": round(ms, 3), "ok": bool(np.array_equal(order.cpu().numpy(), exp))})
        if len(uniq) <= 4:
            rows[-1]["kernels"] = kern
        print(f"config 4 plan {split}: {ms:.3f} ms ok={rows[-1]['ok']}", file=sys.stderr, flush=True)
    if rows:
        best = {k: v for k, v in min(rows, key=lambda r: r["ms"]).items() if k != "kernels"}
        _emit({"config": 4, "what": "ORDER BY over 1-, 2-round and seeded PlanEnumRRS plans, end to end",
               "rows": n, "plans": len(rows), "all_ok": all(r["ok"] for r in rows), "best": best,
               "median_ms": float(np.median([r["ms"] for r in rows])),
               "best_Mrows_s": round(n / best["ms"] / 1e3, 1), "plans_run": rows})
    if not hash_types:
        return
    srt = [s[exp] for s in stacked]
    exact = 1 + int(np.count_nonzero(np.any(np.stack([np.diff(s) != 0 for s in srt]), axis=0)))
    res = {}
    for ht in hash_types:
        box = {}

        def run():
            box["r"] = ctx.multi_round_hash(dcols, hash_type=ht)

        ms, kern = _timed(ctx, run)
        res[f"hash_type_{ht}"] = {"ms": round(ms, 3), "ngroups": box["r"][2], "exact": box["r"][2] == exact,
                                  "kernels": kern}
    _emit({"config": 4, "what": "hash GROUP BY (flag, status, date, key)", "rows": n, "exact_groups": exact, **res})


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--configs", default="1,2,4")
    ap.add_argument("--plans", default=None,
                    help='config 4: the splits to run, "16,26;42" (default: 1- and 2-round and seeded RRS plans), or none')
    ap.add_argument("--hash-types", default="1,0", help="config 4: hash_type values to run, or none")
    ap.add_argument("--copy-probe", action="store_true",
                    help="start with one plain 256 MiB device copy (the traffic reference for PMC counters)")
    ap.add_argument("--handoff-probe", action="store_true",
                    help="time fmcs_probe_handoff (agent-scope releases between streaming tasks) and exit")
    args = ap.parse_args(argv)
    if args.handoff_probe:
        from .engine import Context
        ctx = Context(0)
        tasks = 65536
        bytes_moved = 2.0 * tasks * 4096 * 8
        for grid in (1024, 2048):
            ms = ctx.probe_handoff(tasks, grid)
            names = ["none", "every wave releases per task", "one release per task", "one release per 8 tasks"]
            for m, name in enumerate(names):
                releases = [0, tasks * 4, tasks, tasks // 8][m]
                extra_ns = (ms[m] - ms[0]) * 1e6 / releases if releases else 0.0
                _emit({"probe": "handoff", "grid": grid, "tasks": tasks, "task_bytes": 32768, "mode": m,
                       "what": name, "ms": round(ms[m], 4), "GBs": round(bytes_moved / ms[m] / 1e6, 1),
                       "releases": releases, "ns_per_release": round(extra_ns, 1)})
        return
    if args.copy_probe:
        import torch
        a = torch.empty(1 << 28, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
        print("copy probe: 268435456 bytes read + 268435456 bytes written", file=sys.stderr, flush=True)
    plans = None
    if args.plans is not None:
        plans = [] if args.plans == "none" else [[int(w) for w in p.split(",")] for p in args.plans.split(";")]
    hts = [] if args.hash_types == "none" else [int(x) for x in args.hash_types.split(",")]
    from .engine import Context
    ctx = Context(0)
    for c in [int(x) for x in args.configs.split(",") if x]:
        t0 = time.time()
        if c == 4:
            config4(ctx, plans=plans, hash_types=hts)
        else:
            {1: config1, 2: config2}[c](ctx)
        print(f"config {c} done in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)


if __name__ == "__main__":
    main()
