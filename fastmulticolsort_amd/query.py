"""Query runner over column files: the reference's experiments/main_tpch_*
drivers on the GPU path.

* `run` -- main_tpch_complete_scenario.cpp: read a column file (the TPC-H
  text layout of DataGenerator::ReadFromFile_tpch, src/generator.h:101-136,
  the plain .dat layout, or a binary column file), sort it with the chaining
  composer (`-p ch`: one round per column, ChainComposer) or the stitching
  composer (`-p st`: plan search, then the best plan; optimizeAndRun,
  src/stitch_composer.cpp:3839-4013), ORDER BY (`-o 1`) or GROUP BY
  (`-o 0`), and write the result (oid order, group ids, the sorted columns)
  as a binary column file.
* `exhaustive` -- main_tpch_exhaustive_actual.cpp: ExhaustiveSearch
  (src/stitch_composer.cpp:4471-4568): every split of every round count from
  1 up to ceil((2S - 2) / 16), per column order (all orders for GROUP BY),
  run and timed until more than `--max-plans` plans have run; with each
  plan's estimated cost (the cost-model accuracy of
  main_tpch_costmodel_accuracy.cpp).
* `gen` -- DataGenerator::GenAndSaveAsFile (src/generator.h:138-227): per
  column, values uniform over [0, card] repeated in groups of ceil(n / g)
  rows, shuffled; seeded here instead of time(0).

Options follow the reference's getopt names: -i file, -w widths "25B21",
-a asc/desc digits (1 ascending, anything else descending, one digit per
column from the left; parse_columnwidths), -p ch|st, -o 1|0, -n threads
(text parsing here).  Without -a every column ascends.  Prints one JSON
object per run (and per plan with `exhaustive --plans`).
"""
import argparse
import itertools
import json
import math
import os
import sys
import time

import numpy as np

from . import columnar


def parse_widths(s):
    """'25B21' -> [25, 21] (strtok on 'B', parse_columnwidths)."""
    ws = [int(t) for t in s.split("B") if t]
    if not ws or any(not 1 <= w <= 64 for w in ws):
        raise ValueError(f"bad widths {s!r}")
    return ws


def parse_asc_desc(a, k):
    """parse_columnwidths: the last digit is the last column's; digit 1
    ascends, any other (a missing leading digit reads 0) descends.  Returns
    descending flags."""
    a = int(a)
    desc = [True] * k
    for c in range(k):
        desc[k - 1 - c] = (a % 10) != 1
        a //= 10
    return desc


def load(path, fmt="auto", widths=None, nrows=None, nthreads=0):
    """Columns (numpy, in their banks), widths, cardinalities and names of a
    column file.  fmt: auto (binary if it starts with the magic, else tpch),
    tpch, dat, cols."""
    if fmt == "auto":
        with open(path, "rb") as f:
            fmt = "cols" if f.read(8) == b"FMCSCOLS" else "tpch"
    if fmt == "cols":
        info = columnar.columns_info(path)
        cols = columnar.read_columns(path)
        ws = widths or [c["width"] for c in info["columns"]]
        cards = [c["cardinality"] or None for c in info["columns"]]
        names = [c["name"] for c in info["columns"]]
        if any(c is None for c in cards):
            cards = None
        return cols, ws, cards, names
    if widths is None:
        raise ValueError("text files need their column widths (-w)")
    cols, cards = columnar.read_text(path, widths, nrows=nrows, tpch=(fmt == "tpch"), nthreads=nthreads)
    return cols, list(widths), cards, [f"c{i}" for i in range(len(widths))]


def _device_cols(cols):
    import torch
    out = []
    for c in cols:
        t = torch.from_numpy(c.view(np.int16 if c.dtype == np.uint16 else
                                    (np.int32 if c.dtype == np.uint32 else np.int64)))
        dt = torch.uint16 if c.dtype == np.uint16 else (torch.uint32 if c.dtype == np.uint32 else torch.uint64)
        out.append(t.pin_memory().to("cuda", non_blocking=True).view(dt))
    torch.cuda.synchronize()
    return out


def _host(t):
    import torch
    sdt = {torch.uint16: torch.int16, torch.uint32: torch.int32, torch.uint64: torch.int64}[t.dtype]
    ndt = {torch.uint16: np.uint16, torch.uint32: np.uint32, torch.uint64: np.uint64}[t.dtype]
    return t.view(sdt).cpu().numpy().view(ndt)


def reference_order(cols, desc):
    """The expected ORDER BY order: lexicographic by the columns (descending
    ones inverted), oid as the last key."""
    keys = []
    for c, d in zip(cols, desc):
        c = c.astype(np.uint64)
        keys.append(~c if d else c)
    n = len(cols[0])
    return np.lexsort([np.arange(n, dtype=np.uint64)] + keys[::-1]).astype(np.uint32)


def _model(args, ctx):
    from . import planner
    if getattr(args, "model", None):
        with open(args.model) as f:
            return planner.CostModel.from_dict(json.load(f))
    if getattr(args, "calibrate", False):
        return planner.calibrate(ctx)
    return planner.default_cost_model()


def run_query(cols, widths, desc, comptype="st", ordered=True, cards=None, model=None, ctx=None, repeat=1):
    """Sorts (or groups) device columns; returns (order, groups, info)."""
    import torch
    from . import planner
    from .engine import Context
    own = ctx is None
    if own:
        ctx = Context(0)
    try:
        n = cols[0].numel()
        k = len(cols)
        if cards is None:
            cards = [ctx.distinct_count(c) for c in cols]
        m = model if model is not None else planner.default_cost_model()
        t0 = time.perf_counter()
        if comptype == "ch":
            col_order, split = list(range(k)), list(widths)
        else:
            col_order, split, _, _ = planner.optimize(m, n, widths, cards, ordered=ordered)
        opt_ms = (time.perf_counter() - t0) * 1e3
        est, _ = planner.estimate_plan(m, n, widths, cards, col_order, split)
        order = torch.empty(n, dtype=torch.uint32, device=cols[0].device)
        groups = torch.empty(n, dtype=torch.uint32, device=cols[0].device)
        times = []
        for _ in range(max(1, repeat)):
            _, bd, tot, ng = planner.run_plan(ctx, cols, widths, col_order, split, order=order, groups=groups,
                                              descending=desc)
            times.append(tot)
        info = {"comptype": comptype, "ordered": bool(ordered), "rows": n, "widths": list(widths),
                "descending": list(desc), "cardinality": [int(c) for c in cards], "col_order": list(col_order),
                "split": list(split), "opt_ms": opt_ms, "est_ms": est, "gpu_ms": float(np.median(times)),
                "gpu_ms_all": times, "breakdown_ms": bd, "ngroups": ng}
        return order, groups, info
    finally:
        if own:
            ctx.close()


def cmd_run(args):
    import torch
    from . import engine
    widths = parse_widths(args.widths) if args.widths else None
    t0 = time.perf_counter()
    host, widths, cards, names = load(args.inputfile, args.format, widths, args.nrows, args.nthreads)
    load_ms = (time.perf_counter() - t0) * 1e3
    k = len(host)
    desc = parse_asc_desc(args.asc_desc, k) if args.asc_desc is not None else [False] * k
    t0 = time.perf_counter()
    cols = _device_cols(host)
    upload_ms = (time.perf_counter() - t0) * 1e3
    ctx = engine.Context(0)
    try:
        model = _model(args, ctx)
        order, groups, info = run_query(cols, widths, desc, args.comptype, bool(args.ordered), cards, model, ctx,
                                        args.repeat)
        info.update({"file": os.path.basename(args.inputfile), "load_ms": load_ms, "upload_ms": upload_ms})
        if args.check:
            o = _host(order)
            if args.ordered:
                info["check"] = bool(np.array_equal(o, reference_order(host, desc)))
            else:
                full = np.unique(np.stack([c.astype(np.uint64) for c in host]), axis=1).shape[1]
                info["check"] = info["ngroups"] == full
        if args.out:
            t0 = time.perf_counter()
            o = _host(order)
            g = _host(groups)
            out_cols = [o, g] + [h[o] for h in host]
            columnar.write_columns(args.out, out_cols, [32, 32] + list(widths), names=["oid", "group"] + names,
                                   descending=[False, False] + list(desc))
            info["write_ms"] = (time.perf_counter() - t0) * 1e3
            info["out"] = os.path.basename(args.out)
        print(json.dumps(info), flush=True)
        return info
    finally:
        ctx.close()


def cmd_exhaustive(args):
    import torch
    from . import engine, planner
    widths = parse_widths(args.widths) if args.widths else None
    host, widths, cards, names = load(args.inputfile, args.format, widths, args.nrows, args.nthreads)
    k = len(host)
    cols = _device_cols(host)
    n = len(host[0])
    S = sum(widths)
    ctx = engine.Context(0)
    try:
        model = _model(args, ctx)
        if cards is None:
            cards = [ctx.distinct_count(c) for c in cols]
        orders = [list(p) for p in itertools.permutations(range(k))] if not args.ordered else [list(range(k))]
        max_round = math.ceil((2 * S - 2) / 16)
        if args.max_rounds:
            max_round = min(max_round, args.max_rounds)
        # warm-up with the one-round-per-column plan (the reference runs it 10 times)
        for _ in range(args.warmup):
            planner.run_plan(ctx, cols, widths, list(range(k)), list(widths))
        runs = []
        done = False
        order = torch.empty(n, dtype=torch.uint32, device=cols[0].device)
        first = None
        for co in orders:
            for r in range(1, max_round + 1):
                for split in planner.split_exhaustive(S, r):
                    est, _ = planner.estimate_plan(model, n, widths, cards, co, split)
                    _, bd, tot, ng = planner.run_plan(ctx, cols, widths, co, split, order=order)
                    rec = {"col_order": co, "split": split, "rounds": r, "est_ms": est, "gpu_ms": tot,
                           "breakdown_ms": bd}
                    if args.check and args.ordered:
                        # every plan sorts by the same key with oid ties: one permutation
                        if first is None:
                            first = order.clone()
                        rec["same_order"] = bool(torch.equal(order, first))
                    runs.append(rec)
                    if args.plans:
                        print(json.dumps(rec), flush=True)
                if len(runs) > args.max_plans:  # the reference stops once past 1000 plans
                    done = True
                    break
            if done:
                break
        best = min(runs, key=lambda x: x["gpu_ms"])
        po, ps, _, _ = planner.optimize(model, n, widths, cards, ordered=bool(args.ordered))
        actual = sorted(x["gpu_ms"] for x in runs)
        pick = next((x for x in runs if x["col_order"] == po and x["split"] == ps), None)
        est = np.array([x["est_ms"] for x in runs])
        act = np.array([x["gpu_ms"] for x in runs])
        rank_corr = None
        if len(runs) > 2:
            re_, ra = np.argsort(np.argsort(est)), np.argsort(np.argsort(act))
            rank_corr = float(np.corrcoef(re_, ra)[0, 1])
        summary = {"file": os.path.basename(args.inputfile), "rows": n, "widths": widths,
                   "ordered": bool(args.ordered), "plans": len(runs), "best": best,
                   "optimizer_pick": {"col_order": po, "split": ps,
                                      "gpu_ms": pick["gpu_ms"] if pick else None,
                                      "rank": (actual.index(pick["gpu_ms"]) + 1) if pick else None,
                                      "regret": (pick["gpu_ms"] / best["gpu_ms"] - 1) if pick else None},
                   "est_actual_rank_correlation": rank_corr}
        if args.check and args.ordered:
            summary["all_same_order"] = all(x["same_order"] for x in runs)
        print(json.dumps(summary), flush=True)
        return summary, runs
    finally:
        ctx.close()


def generate(nrows, widths, cardinality_pct=100, group_num=None, seed=0x5EED):
    """GenAndSaveAsFile's columns: card = (2^w - 1) * pct / 100 (2^64 - 1 at
    w = 64); ceil(n / g) rows per drawn value, shuffled."""
    from .datagen import permutation, splitmix64
    cols = []
    for c, w in enumerate(widths):
        card = ((1 << w) - 1) * cardinality_pct // 100 if w < 64 else (1 << 64) - 1
        g = group_num or nrows
        each = -(-nrows // g)
        nv = -(-nrows // each)
        z = splitmix64(seed + 101 * c, nv)
        vals = z % np.uint64(card + 1) if card < (1 << 64) - 1 else z
        col = np.repeat(vals, each)[:nrows][permutation(seed + 211 * c, nrows)]
        cols.append(col.astype(columnar.text_dtype(w)))
    return cols


def cmd_gen(args):
    widths = parse_widths(args.widths)
    cols = generate(args.nrows, widths, args.cardinality, args.ngroups, args.seed)
    if args.format == "cols":
        cards = [int(np.unique(c).size) for c in cols]
        columnar.write_columns(args.out, cols, widths, cardinalities=cards)
    else:
        columnar.write_text(args.out, cols, widths, tpch=(args.format == "tpch"))
    print(json.dumps({"out": os.path.basename(args.out), "rows": args.nrows, "widths": widths,
                      "format": args.format}), flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m fastmulticolsort_amd.query", description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    for name in ("run", "exhaustive"):
        p = sub.add_parser(name)
        p.add_argument("-i", "--inputfile", required=True)
        p.add_argument("-w", "--bitwidth", dest="widths", default=None)
        p.add_argument("-o", "--ordered", type=int, default=1)
        p.add_argument("-n", "--nthreads", type=int, default=0)
        p.add_argument("--format", choices=["auto", "tpch", "dat", "cols"], default="auto")
        p.add_argument("--nrows", type=int, default=None, help="rows per column of a plain .dat file")
        p.add_argument("--model", default=None, help="cost model JSON (planner.CostModel.to_dict)")
        p.add_argument("--calibrate", action="store_true", help="measure the cost model on this GPU first")
        if name == "run":
            p.add_argument("-a", "--asc-desc", dest="asc_desc", default=None)
            p.add_argument("-p", "--comptype", choices=["ch", "st"], default="st")
            p.add_argument("--repeat", type=int, default=1)
            p.add_argument("--check", action="store_true", help="compare with a host lexsort / unique count")
            p.add_argument("--out", default=None, help="result as a binary column file")
        else:
            p.add_argument("--max-plans", type=int, default=1000)
            p.add_argument("--max-rounds", type=int, default=0)
            p.add_argument("--warmup", type=int, default=3)
            p.add_argument("--plans", action="store_true", help="one JSON line per plan")
            p.add_argument("--check", action="store_true", help="every plan's order equals the first plan's")
    g = sub.add_parser("gen")
    g.add_argument("-r", "--nrows", type=int, required=True)
    g.add_argument("-w", "--bitwidth", dest="widths", required=True)
    g.add_argument("-C", "--cardinality", type=int, default=100)
    g.add_argument("-g", "--ngroups", type=int, default=None)
    g.add_argument("--seed", type=int, default=0x5EED)
    g.add_argument("--format", choices=["tpch", "dat", "cols"], default="tpch")
    g.add_argument("--out", required=True)
    args = ap.parse_args(argv)
    if args.cmd == "run":
        return cmd_run(args)
    if args.cmd == "exhaustive":
        return cmd_exhaustive(args)
    return cmd_gen(args)


if __name__ == "__main__":
    main(sys.argv[1:])
