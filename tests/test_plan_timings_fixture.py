"""The committed plan-timing fixture (tests/golden/plan_timings.npz, written
on MI355X by tests/golden/make_plan_timings.py) against the planner, on the
CPU: every request holds every candidate plan of the enumeration the picker
chooses from, and the stored analytic estimates are the library's own under
the stored calibrated constants."""
import json
import os

import numpy as np
import pytest

from fastmulticolsort_amd import learned, planner

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "plan_timings.npz")


def load():
    d = np.load(FIXTURE, allow_pickle=False)
    return {k: d[k] for k in d.files}


def test_fixture_shape():
    d = load()
    meta = json.loads(str(d["meta"]))
    reqs = json.loads(str(d["requests"]))
    assert len(reqs) == meta["train"] + meta["held_out"]
    n = len(d["ms"])
    assert len(d["est"]) == len(d["req"]) == len(d["plan"]) == n
    assert np.all(d["ms"] > 0) and np.all(d["est"] > 0)
    assert np.array_equal(np.unique(d["req"]), np.arange(len(reqs)))


@pytest.mark.parametrize("part", range(4))
def test_every_candidate_and_the_library_estimates(part):
    d = load()
    reqs = json.loads(str(d["requests"]))
    model = planner.CostModel.from_dict(json.loads(str(d["model"])))
    for r in range(part, len(reqs), 4):
        q = reqs[r]
        order = list(range(len(q["bws"])))
        m = np.where(d["req"] == r)[0]
        stored = [tuple(int(w) for w in d["plan"][i].split(",")) for i in m]
        cand = learned.candidate_plans(model, q["n"], q["bws"], q["cards"], order, seed=r, max_plans=1 << 20)
        assert sorted(stored) == sorted(tuple(p) for p in cand), r
        for i, p in zip(m, stored):
            e = planner.estimate_plan(model, q["n"], q["bws"], q["cards"], order, list(p))[0]
            assert abs(e - d["est"][i]) <= 1e-12 * e, (r, p)
