"""Writes tests/golden/plan_timings.npz: GPU-measured times of massaging
plans over seeded multi-column requests, the data the learned plan picker
is trained and judged on (the measurement the reference's
CostModelComparison makes, src/stitch_composer.cpp:4337-4469).

Run on a GPU from the repository root:

    python tests/golden/make_plan_timings.py [--train 256] [--held-out 64] [--out tests/golden/plan_timings.npz]

Steps:
  1. calibrate the analytic cost model on the device (fmcs_calibrate);
  2. draw seeded requests (learned.make_request: 2-6 columns, widths 1-40,
     uniform or Zipf values, 2^12 .. 2^20 rows);
  3. per request, enumerate every candidate plan (1- and 2-round exhaustive,
     greedy 3..6 rounds, seeded PlanEnumRRS);
  4. time every plan on the GPU (device-driven mode, best of 5 runs after a
     warm-up), checking each plan's order against numpy's lexsort of the
     request's columns;
  5. store the times, the analytic estimates under the same calibrated
     constants, the plans, the requests and the constants.  The picker's
     features follow from a request, its plans and their estimates
     (learned.request_features), so they are not stored.

The file holds arrays only (no pickles): numpy.load(..., allow_pickle=False)
reads it.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from fastmulticolsort_amd import learned, planner  # noqa: E402
from fastmulticolsort_amd.engine import Context  # noqa: E402

SEED = 20261204
MAX_LOG2 = 20
ALL_PLANS = 1 << 20
REPS = 5


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=256)
    ap.add_argument("--held-out", type=int, default=64)
    ap.add_argument("--out", default=os.path.join(ROOT, "tests", "golden", "plan_timings.npz"))
    args = ap.parse_args(argv)
    ctx = Context(0)
    t0 = time.time()
    model = planner.calibrate(ctx, nrows=1 << 22)
    print(f"calibrated in {time.time() - t0:.1f}s", flush=True)
    rng = np.random.default_rng(SEED)
    MS, EST, R, P, reqs = [], [], [], [], []
    total = args.train + args.held_out
    for i in range(total):
        req = learned.make_request(rng, max_log2=MAX_LOG2)
        order = list(range(len(req["bws"])))
        plans = learned.candidate_plans(model, req["n"], req["bws"], req["cards"], order, seed=i, max_plans=ALL_PLANS)
        ms = learned.measure_plans(ctx, req, plans, reps=REPS, verify=True)
        est = [planner.estimate_plan(model, req["n"], req["bws"], req["cards"], order, p)[0] for p in plans]
        MS.extend(ms)
        EST.extend(est)
        R.extend([i] * len(plans))
        P.extend(",".join(map(str, p)) for p in plans)
        reqs.append(req)
        print(f"request {i}/{total}: n={req['n']} bws={req['bws']} plans={len(plans)} best={min(ms):.3f} ms "
              f"({time.time() - t0:.0f}s)", flush=True)
    np.savez_compressed(
        args.out,
        ms=np.array(MS, np.float64),
        est=np.array(EST, np.float64),
        req=np.array(R, np.int32),
        plan=np.array(P),
        requests=np.array(json.dumps(reqs)),
        model=np.array(json.dumps(model.to_dict())),
        meta=np.array(json.dumps({"seed": SEED, "train": args.train, "held_out": args.held_out,
                                  "max_log2": MAX_LOG2, "reps": REPS,
                                  "device": ctx.device_name() if hasattr(ctx, "device_name") else "gfx950"})))
    print(f"wrote {args.out}: {len(MS)} plans over {total} requests in {time.time() - t0:.0f}s", flush=True)


if __name__ == "__main__":
    main()
