"""The learned cost model's kernels: forward, backward and Adam against an
fp64 restatement (pinned itself by central finite differences), and the
plan picker's gain over the analytic model on GPU-measured plans."""
import math

import numpy as np
import pytest
import torch

from fastmulticolsort_amd import learned, planner
from mlp_ref import fwd64, grad64, loss64

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("B", [1, 17, 300])
def test_forward_and_backward_match_fp64(ctx, B):
    m = learned.Mlp(ctx, nin=learned.NFEAT, h1=64, h2=64, seed=3)
    rng = np.random.default_rng(B)
    X = rng.uniform(0, 1, (B, learned.NFEAT)).astype(np.float32)
    t = rng.normal(0, 1, B).astype(np.float32)
    Xd, td = torch.from_numpy(X).cuda(), torch.from_numpy(t).cuda()
    y, loss = m.backward(Xd, td)
    p = m.params.cpu().numpy().astype(np.float64)
    ey, _ = fwd64(p, X.astype(np.float64), learned.NFEAT, 64, 64)
    np.testing.assert_allclose(y.cpu().numpy(), ey, rtol=1e-4, atol=1e-4)
    eg = grad64(p, X.astype(np.float64), t.astype(np.float64), learned.NFEAT, 64, 64)
    np.testing.assert_allclose(m.grad.cpu().numpy(), eg, rtol=2e-3, atol=2e-5)
    assert abs(loss.item() - loss64(p, X.astype(np.float64), t, learned.NFEAT, 64, 64)) < 1e-4 * max(1, loss.item())


def test_backward_is_deterministic(ctx):
    m = learned.Mlp(ctx, seed=1)
    X = torch.rand(777, learned.NFEAT, device="cuda")
    t = torch.rand(777, device="cuda")
    m.backward(X, t)
    g1 = m.grad.clone()
    m.backward(X, t)
    assert torch.equal(g1, m.grad)


def test_adam_step_matches_fp64(ctx):
    m = learned.Mlp(ctx, seed=2)
    X = torch.rand(64, learned.NFEAT, device="cuda")
    t = torch.rand(64, device="cuda")
    p0 = m.params.cpu().numpy().astype(np.float64)
    m.backward(X, t)
    g = m.grad.cpu().numpy().astype(np.float64)
    m.adam(lr=1e-2)
    mm = 0.1 * g
    vv = 0.001 * g * g
    exp = p0 - 1e-2 * (mm / 0.1) / (np.sqrt(vv / 0.001) + 1e-8)
    np.testing.assert_allclose(m.params.cpu().numpy(), exp, rtol=1e-5, atol=1e-6)


def test_training_fits_a_smooth_function(ctx):
    rng = np.random.default_rng(0)
    X = rng.uniform(0, 1, (2048, learned.NFEAT)).astype(np.float32)
    y = (np.sin(3 * X[:, 0]) + X[:, 3] * X[:, 5] + 0.5 * X[:, 7]).astype(np.float32)
    m = learned.Mlp(ctx, seed=4)
    hist = m.train(torch.from_numpy(X).cuda(), torch.from_numpy(y).cuda(), epochs=60, batch=128, lr=3e-3)
    assert hist[-1] < hist[0] / 10


def test_learned_picker_beats_analytic_on_measured_plans(ctx):
    """Train on GPU-measured plan times of seeded requests, then compare plan
    regret (measured time of the pick / best measured - 1) of the MLP pick
    and of the analytic cost model's pick on held-out requests.  Both use
    the device-calibrated cost model; the MLP is an ensemble of three seeds
    (mean prediction), which steadies a 56-request training set."""
    model = planner.calibrate(ctx, nrows=1 << 22)
    data = learned.collect(ctx, nreq=80, seed=11, max_log2=20, max_plans=24, model=model)
    X, y, req = data["X"], data["y"], data["req"]
    train = req < 56
    test = ~train
    Xt = torch.from_numpy(X[test]).cuda()
    pred = np.zeros(int(test.sum()), np.float64)
    for s in range(3):
        m = learned.Mlp(ctx, seed=s)
        m.train(torch.from_numpy(X[train]).cuda(), torch.from_numpy(y[train]).cuda(), epochs=200, batch=64, lr=3e-3)
        pred += m.forward(Xt).cpu().numpy()
    reg_mlp = learned.regret_table(pred, y[test], req[test])
    # analytic picks
    import json
    reqs = json.loads(data["requests"])
    est = []
    for x_plan, r in zip(data["plan"][test], req[test]):
        q = reqs[r]
        p = [int(w) for w in x_plan.split(",")]
        est.append(planner.estimate_plan(model, q["n"], q["bws"], q["cards"], list(range(len(q["bws"]))), p)[0])
    reg_an = learned.regret_table(np.log(np.array(est)), y[test], req[test])
    print("regret mlp mean %.3f analytic mean %.3f" % (reg_mlp.mean(), reg_an.mean()), flush=True)
    assert reg_mlp.mean() <= reg_an.mean() + 0.02
