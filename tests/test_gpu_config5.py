"""Config 5 (the learned picker's request workload): every plan gives the
reference's result.

Seeded requests of learned.make_request (2-6 columns of 1-40 bits, distinct
counts log-uniform, uniform or Zipf(1.0) values; rows up to 2^20 here).
For each request every candidate plan (rounds <= 2 exhaustive, greedy 3-6
rounds, seeded RRS: the ExhaustiveSearch set, src/stitch_composer.cpp:
4471-4572), plans whose later rounds are 53-64 bits wide, and the plan the
MLP picks are run through fmcs_run_plan.  Each order must equal numpy's
stable lexsort of the columns, and the final groups the distinct tuples
(np.unique).  Bar: bit-exact.
"""
import numpy as np
import pytest
import torch

from fastmulticolsort_amd import learned, planner
from fastmulticolsort_amd.engine import MODE_BATCHED, MODE_DEVICE

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)


def requests(count=18, seed=20261204, max_log2=20):
    rng = np.random.default_rng(seed)
    return [learned.make_request(rng, max_log2=max_log2) for _ in range(count)]


def wide_tail_plans(S):
    """Plans whose later rounds are 53..64 bits wide (keys too wide to pack a
    segment index beside them inside an LDS tile)."""
    out = []
    for w in (64, 60, 53):
        r = -(-S // w)
        first = S - w * (r - 1)
        if r >= 2 and 0 < first <= 64:
            out.append([first] + [w] * (r - 1))
    return out


def expected_groups(host, exp):
    sorted_cols = [c[exp] for c in host]
    heads = np.zeros(len(exp), dtype=bool)
    heads[0] = True
    for c in sorted_cols:
        heads[1:] |= c[1:] != c[:-1]
    return (np.cumsum(heads) - 1).astype(np.uint32)


def test_requests_cover_the_config5_shapes():
    reqs = requests()
    assert sum(len(r["bws"]) >= 5 for r in reqs) >= 4
    assert sum(any(r["zipf"]) for r in reqs) >= 10
    assert sum(sum(r["bws"]) > 64 for r in reqs) >= 12  # more than one round needed
    assert max(r["n"] for r in reqs) > (1 << 19)


def test_every_candidate_plan_of_config5_requests(ctx):
    model = planner.default_cost_model()
    mlp = learned.Mlp(ctx, seed=1)  # its pick is one of the candidates; the pick never changes a result
    nplans = nwide = 0
    for i, req in enumerate(requests()):
        host = learned.request_columns(req)
        n, bws, cards = req["n"], req["bws"], req["cards"]
        cols = [torch.from_numpy(c).cuda() for c in host]
        exp = learned.expected_order(host)
        gexp = expected_groups(host, exp)
        nd = len(np.unique(np.stack(host, axis=1), axis=0))
        assert int(gexp[-1]) + 1 == nd
        ids = list(range(len(bws)))
        plans = learned.candidate_plans(model, n, bws, cards, ids, seed=i, max_plans=1 << 20)
        plans += [p for p in wide_tail_plans(sum(bws)) if p not in plans]
        k, _ = learned.pick(mlp, model, n, bws, cards, ids, plans)
        log(f"request {i}: n={n} bws={bws} zipf={req['zipf']} plans={len(plans)} mlp pick={plans[k]}")
        order = torch.empty(n, dtype=torch.uint32, device="cuda")
        groups = torch.empty(n, dtype=torch.uint32, device="cuda")
        for j, p in enumerate(plans):
            mode = MODE_DEVICE if j % 4 else MODE_BATCHED
            order.view(torch.int32).fill_(0)
            _, _, _, ng = planner.run_plan(ctx, cols, bws, ids, p, mode=mode, order=order, groups=groups)
            ctx.sync()
            assert np.array_equal(order.cpu().numpy(), exp), (i, p, mode)
            assert ng == nd, (i, p, ng, nd)
            assert np.array_equal(groups.cpu().numpy(), gexp), (i, p)
            nplans += 1
            nwide += any(w >= 53 for w in p[1:])
        # the learned picker's own path (planner.optimize_and_run with the MLP)
        res = planner.optimize_and_run(ctx, cols, bws, cards=cards, model=model, mode=MODE_DEVICE, mlp=mlp)
        assert np.array_equal(res["order"].cpu().numpy(), exp)
    log(f"{nplans} plans checked, {nwide} with a later round of 53-64 bits")
    assert nplans > 1500 and nwide >= 300


def test_measure_plans_verifies_orders(ctx):
    req = requests(count=1, seed=7, max_log2=16)[0]
    plans = learned.candidate_plans(planner.default_cost_model(), req["n"], req["bws"], req["cards"],
                                    list(range(len(req["bws"]))), seed=0, max_plans=6)
    ms = learned.measure_plans(ctx, req, plans, reps=1)
    assert len(ms) == len(plans) and all(t > 0 for t in ms)
