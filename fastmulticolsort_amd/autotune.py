"""Autotuner for the batched segmented sort.

Sweeps the bin thresholds and tile shapes of fmcs_tuning (small_max: the
largest segment packed into an LDS tile, which also picks the 512-, 1024-
or 2048-row tile; window: rows per packing window; msd_bits: digit width of an
MSD level), how the second round's column reaches the first round's order
(carry_next: moved along by the first round's radix passes, packed into
composite keys or as its own column, or gathered, in one pass or in passes
over source windows of gather_window_kb KiB) and the form of the batched
sort (med_max 0: every radix level sorts its own windows; else one
persistent size-binned launch after the levels, with single segments of up to
med_max rows sorted by one workgroup each) on a workload, timing whole
multi-round sorts, and writes the fastest setting as JSON.  The column move is
swept first with the sort settings fixed, then the sort grid with the best
move, then the persistent form's thresholds with the best sort settings.
Results never depend on the tuning, only the time does.
"""
import itertools
import json
import os
import time

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_PATH = os.path.join(HERE, "tuning_mi355x.json")

GRID = {
    "small_max": [256, 512, 1024],
    "window_frac": [1.0, 0.5],
    "msd_bits": [6, 7, 8],
}
GATHER_KB = [0, 262144]
MED_MAX = [2048, 4096, 8192, 16384]


def candidates(grid=GRID):
    for sm, wf, mb in itertools.product(grid["small_max"], grid["window_frac"], grid["msd_bits"]):
        yield {"small_max": sm, "window": max(1, int(sm * wf)), "msd_bits": mb}


def apply(ctx, setting):
    ctx.set_tuning(setting["small_max"], setting["window"], setting["msd_bits"], setting.get("gather_window_kb", 0),
                   setting.get("carry_next", 1), setting.get("med_max", 0))


def time_setting(ctx, run, setting, reps=3):
    import torch
    apply(ctx, setting)
    run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        run()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / reps


def tune(ctx, run, grid=GRID, gather_kb=GATHER_KB, med_max=MED_MAX, reps=3, log=None):
    """run(): one multi-round sort on resident data.  Returns (best, table)."""
    table = []

    def one(st):
        ms = time_setting(ctx, run, st, reps)
        table.append(dict(st, ms=round(ms, 4)))
        if log:
            log(f"tune {st} -> {ms:.3f} ms")
        return ms

    base = {k: v for k, v in ctx.tuning().items()}
    # how round 2's column reaches round-1 order: carried by round 1's
    # radix passes, or gathered (in one pass or source windows)
    moves = [dict(carry_next=1, gather_window_kb=0), dict(carry_next=2, gather_window_kb=0)] + \
        [dict(carry_next=0, gather_window_kb=g) for g in gather_kb]
    mv = min(moves, key=lambda m: one(dict(base, **m)))
    for st in candidates(grid):
        one(dict(st, med_max=0, **mv))
    # the persistent form (1024-row window tiles: small_max at most 512)
    lvl = min(table, key=lambda r: r["ms"])
    sm = min(lvl["small_max"], 512)
    for mm in med_max:
        one(dict(lvl, small_max=sm, window=min(lvl["window"], sm), med_max=mm))
    best = min(table, key=lambda r: r["ms"])
    apply(ctx, best)
    return best, table


def save(best, table, workload, path=DEFAULT_PATH):
    with open(path, "w") as f:
        json.dump({"workload": workload, "best": best, "table": table}, f, indent=1)


def load(ctx, path=DEFAULT_PATH):
    """Apply a saved tuning; returns it, or None when there is none."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        best = json.load(f)["best"]
    apply(ctx, best)
    return best
