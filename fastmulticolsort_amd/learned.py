"""Learned plan picker: an MLP cost model trained on GPU-measured plan times.

Per (request, plan) the features are the row count, the analytic cost
model's log-estimate of the plan relative to the mean over the request's
candidates, and the plan's rounds (width, bank, and the estimated group size
before the round from the base cardinalities, as the analytic model computes
it), padded to MAX_ROUNDS.  The network predicts log(ms) relative to the
mean over the request's candidates (only the ranking inside a request
matters); the planner runs the candidate plans' features through the MLP
forward kernel and takes the argmin.  With the analytic estimate as an
input the network learns a correction of the analytic model from measured
times.  Training
uses the project's own forward / backward / Adam kernels (include/fmcs_plan.h).
The analytic model plus exhaustive measurement over the candidates stays the
fallback and the judge of the learned pick's regret; the chosen plan never
changes a result, only the time.
"""
import ctypes as C
import json
import math
import os

import numpy as np

from . import _lib, planner

MAX_ROUNDS = 6
PER_ROUND = 7
NFEAT = 4 + PER_ROUND * MAX_ROUNDS
HERE = os.path.dirname(os.path.abspath(__file__))


class MlpShape(C.Structure):
    _fields_ = [("nin", C.c_uint32), ("h1", C.c_uint32), ("h2", C.c_uint32)]


_sig = False


def lib():
    global _sig
    L = planner.lib()
    if not _sig:
        vp, u32, f32 = C.c_void_p, C.c_uint32, C.c_float
        P = C.POINTER
        for name, args in {
            "fmcs_mlp_param_count": [P(MlpShape), P(u32)],
            "fmcs_mlp_forward": [vp, P(MlpShape), vp, vp, u32, vp, vp],
            "fmcs_mlp_backward": [vp, P(MlpShape), vp, vp, vp, vp, vp, u32, vp, vp],
            "fmcs_mlp_adam": [vp, u32, vp, vp, vp, vp, f32, f32, f32, f32, u32],
        }.items():
            fn = getattr(L, name)
            fn.restype = C.c_int
            fn.argtypes = args
        _sig = True
    return L


def bank_type(w):
    return 0 if w <= 16 else (1 if w <= 32 else 2)


def round_groups(n, bws, cards, col_order, split):
    """Estimated groups before each round: the card_per_bit products of
    getCardinalityFromStatistics (src/stitch_composer.cpp:288-553)."""
    cpb = [cards[c] ** (1.0 / bws[c]) for c in range(len(bws))]
    card = [1.0] * len(split)
    oi, ni = 0, 0
    ow, nw = bws[col_order[0]], split[0]
    while ni < len(split):
        m = min(ow, nw)
        ow -= m
        nw -= m
        if ni + 1 < len(split):
            card[ni] *= cpb[col_order[oi]] ** m
        if ow == 0:
            oi += 1
            if oi < len(bws):
                ow = bws[col_order[oi]]
        if nw == 0:
            ni += 1
            if ni < len(split):
                nw = split[ni]
    G, acc = [], 1.0
    for r in range(len(split)):
        G.append(min(acc, float(n)))
        acc *= card[r]
    return G


def features(n, bws, cards, col_order, split, rel_est=0.0):
    """rel_est: the analytic log-estimate of this plan minus the mean over
    the request's candidates."""
    f = np.zeros(NFEAT, np.float32)
    f[0] = math.log2(n) / 24.0
    f[1] = len(split) / MAX_ROUNDS
    f[2] = sum(bws) / 240.0
    f[3] = max(-4.0, min(4.0, rel_est)) / 4.0
    G = round_groups(n, bws, cards, col_order, split)
    for r, w in enumerate(split[:MAX_ROUNDS]):
        b = 4 + PER_ROUND * r
        f[b] = w / 64.0
        f[b + 1 + bank_type(w)] = 1.0
        f[b + 4] = math.log2(max(1.0, n / max(1.0, G[r]))) / 24.0
        # fixed-cost drivers: LSD passes (round 1) or segmented-sort levels,
        # and keys too wide for a packed segment index (comparison rank)
        f[b + 5] = ((w + 7) // 8 + (1 if r else 0)) / 9.0
        f[b + 6] = 1.0 if w > 53 else 0.0
    return f


class Mlp:
    """Device-resident MLP parameters with forward/backward/Adam kernels."""

    def __init__(self, ctx, nin=NFEAT, h1=64, h2=64, seed=0):
        import torch
        self.ctx = ctx
        self.shape = MlpShape(nin, h1, h2)
        cnt = C.c_uint32()
        assert lib().fmcs_mlp_param_count(C.byref(self.shape), C.byref(cnt)) == 0
        self.P = cnt.value
        rng = np.random.default_rng(seed)
        p = np.zeros(self.P, np.float32)
        o = 0
        for (fan_out, fan_in) in ((h1, nin), (h2, h1)):
            p[o:o + fan_out * fan_in] = rng.normal(0, math.sqrt(2.0 / fan_in), fan_out * fan_in)
            o += fan_out * fan_in + fan_out
        p[o:o + h2] = rng.normal(0, math.sqrt(1.0 / h2), h2)
        dev = "cuda"
        self.params = torch.from_numpy(p).to(dev)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.grad = torch.zeros_like(self.params)
        self.step = 0

    def _p(self, t):
        return C.c_void_p(t.data_ptr())

    def forward(self, x, acts=None):
        import torch
        y = torch.empty(x.shape[0], dtype=torch.float32, device=x.device)
        self.ctx._chk(lib().fmcs_mlp_forward(self.ctx.h, C.byref(self.shape), self._p(self.params), self._p(x),
                                             x.shape[0], self._p(y), self._p(acts) if acts is not None else None))
        return y

    def backward(self, x, target):
        import torch
        B = x.shape[0]
        acts = torch.empty(B, self.shape.h1 + self.shape.h2, dtype=torch.float32, device=x.device)
        y = self.forward(x, acts)
        loss = torch.zeros(1, dtype=torch.float32, device=x.device)
        self.ctx._chk(lib().fmcs_mlp_backward(self.ctx.h, C.byref(self.shape), self._p(self.params), self._p(x),
                                              self._p(acts), self._p(y), self._p(target), B, self._p(self.grad),
                                              self._p(loss)))
        return y, loss

    def adam(self, lr=3e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step += 1
        self.ctx._chk(lib().fmcs_mlp_adam(self.ctx.h, self.P, self._p(self.params), self._p(self.grad),
                                          self._p(self.m), self._p(self.v), lr, b1, b2, eps, self.step))

    def train(self, X, y, epochs=300, batch=256, lr=3e-3, seed=0):
        """Minibatch Adam on (X, y) device tensors; returns the loss history."""
        import torch
        n = X.shape[0]
        g = torch.Generator(device="cpu").manual_seed(seed)
        hist = []
        for ep in range(epochs):
            perm = torch.randperm(n, generator=g).to(X.device)
            tot = 0.0
            for i in range(0, n, batch):
                idx = perm[i:i + batch]
                xb, yb = X[idx].contiguous(), y[idx].contiguous()
                _, loss = self.backward(xb, yb)
                self.adam(lr)
                tot += float(loss.item()) * len(idx)
            hist.append(tot / n)
        return hist

    def save(self, path):
        np.save(path, self.params.cpu().numpy())

    def load(self, path):
        import torch
        self.params = torch.from_numpy(np.load(path)).to("cuda")


def candidate_plans(model, n, bws, cards, col_order, seed=0, max_plans=48):
    """Rounds <= 2 exhaustive, greedy for 3..6 rounds, seeded RRS visits
    (src/stitch_composer.cpp:3889-3923, src/plan_enum_rrs.cpp)."""
    S = sum(bws)
    plans = []
    for r in (1, 2):
        plans += planner.split_exhaustive(S, r, 64)
    for r in range(3, MAX_ROUNDS + 1):
        plans += planner.greedy_plans(model, n, bws, cards, col_order, r)
    plans += planner.plan_enum_rrs(model, n, bws, cards, col_order, min(MAX_ROUNDS, max(1, -(-S // 16))), seed)
    seen, out = set(), []
    for p in plans:
        t = tuple(int(w) for w in p)
        if t not in seen and len(t) <= MAX_ROUNDS and sum(t) == S:
            seen.add(t)
            out.append(list(t))
    if len(out) > max_plans:
        # keep the analytic model's best half and a seeded sample of the rest
        est = [planner.estimate_plan(model, n, bws, cards, col_order, p)[0] for p in out]
        idx = np.argsort(est)
        keep = list(idx[: max_plans // 2])
        rest = idx[max_plans // 2:]
        rng = np.random.default_rng(seed)
        keep += list(rng.choice(rest, size=max_plans - len(keep), replace=False))
        out = [out[i] for i in sorted(keep)]
    return out


def request_features(n, bws, cards, col_order, plans, est):
    """Feature rows of a request's candidate plans given their analytic
    estimates (ms): each plan's log-estimate enters relative to the mean
    over the candidates."""
    le = np.log(np.maximum(np.asarray(est, np.float64), 1e-9))
    rel = le - le.mean()
    return np.stack([features(n, bws, cards, col_order, p, r) for p, r in zip(plans, rel)])


def plan_features(model, n, bws, cards, col_order, plans):
    """Feature rows of a request's candidate plans and the analytic estimates."""
    est = np.array([planner.estimate_plan(model, n, bws, cards, col_order, p)[0] for p in plans])
    return request_features(n, bws, cards, col_order, plans, est), est


def relative_log(ms):
    """Training target: log time relative to the mean over the request's plans."""
    lt = np.log(np.maximum(np.asarray(ms, np.float64), 1e-3))
    return (lt - lt.mean()).astype(np.float32)


def pick(mlp, model, n, bws, cards, col_order, plans):
    import torch
    X, _ = plan_features(model, n, bws, cards, col_order, plans)
    y = mlp.forward(torch.from_numpy(X).cuda()).cpu().numpy()
    return int(np.argmin(y)), y


# ---- request workload: seeded multi-column requests ------------------------

def make_request(rng, min_log2=12, max_log2=24):
    """2-6 columns, widths U[1,40], distinct counts log-uniform in [2, 2^w],
    values uniform or Zipf(1.0) by a coin; rows log-uniform."""
    n = int(round(2 ** rng.uniform(min_log2, max_log2)))
    ncols = int(rng.integers(2, 7))
    bws = [int(rng.integers(1, 41)) for _ in range(ncols)]
    cards = [max(2, min(2 ** w, int(round(2 ** rng.uniform(1, w))))) if w > 1 else 2 for w in bws]
    zipf = [bool(rng.integers(0, 2)) for _ in range(ncols)]
    return {"n": n, "bws": bws, "cards": cards, "zipf": zipf, "seed": int(rng.integers(0, 2 ** 31))}


def request_columns(req):
    """Host columns of a request (banks by width)."""
    from .datagen import bank_dtype, splitmix64
    n = req["n"]
    cols = []
    for c, (w, card, z) in enumerate(zip(req["bws"], req["cards"], req["zipf"])):
        u = splitmix64(req["seed"] * 16 + c, n)
        if z:
            # P(rank k) ~ 1/k: rank = floor(card^x) - 1 for x uniform in [0, 1)
            x = (u >> np.uint64(11)).astype(np.float64) / float(1 << 53)
            rank = np.minimum(np.floor(np.power(float(card), x)).astype(np.int64) - 1, card - 1)
        else:
            rank = (u % np.uint64(card)).astype(np.int64)
        step = max(1, (1 << w) // card)
        vals = (rank.astype(np.uint64) * np.uint64(step)) & np.uint64((1 << w) - 1)
        cols.append(vals.astype(bank_dtype(w)))
    return cols


class PlanMismatch(AssertionError):
    """A plan's order differs from the lexicographic order of the columns."""


def expected_order(host_cols):
    """ORDER BY the columns with the oid tie-break (numpy's stable lexsort)."""
    return np.lexsort(tuple(reversed(host_cols))).astype(np.uint32)


def measure_plans(ctx, req, plans, reps=2, verify=True, mode=None):
    """GPU time (ms, HIP events) of every plan of a request, in the
    device-driven mode the request queue runs (mode None).  verify: every
    plan's order is checked against numpy's lexsort of the request's columns
    (PlanMismatch on a difference); the plan never changes the result."""
    import torch
    from .engine import MODE_DEVICE
    mode = MODE_DEVICE if mode is None else mode
    host = request_columns(req)
    cols = [torch.from_numpy(c).cuda() for c in host]
    order = torch.empty(req["n"], dtype=torch.uint32, device="cuda")
    order_cols = list(range(len(req["bws"])))
    exp = expected_order(host) if verify else None
    out = []
    planner.run_plan(ctx, cols, req["bws"], order_cols, plans[0], mode=mode, order=order)  # warm-up
    for p in plans:
        best = float("inf")
        for _ in range(reps):
            _, _, tot, _ = planner.run_plan(ctx, cols, req["bws"], order_cols, p, mode=mode, order=order)
            best = min(best, tot)
        if verify and not np.array_equal(order.cpu().numpy(), exp):
            raise PlanMismatch(f"plan {p} of request n={req['n']} bws={req['bws']}: order differs from lexsort")
        out.append(best)
    return out


def collect(ctx, nreq, seed=5, max_log2=22, max_plans=32, log=None, model=None):
    """Measured (features, ms) over seeded requests; returns a dict of arrays.
    model: the cost model of the analytic-estimate feature (None: default)."""
    rng = np.random.default_rng(seed)
    if model is None:
        model = planner.default_cost_model()
    X, Y, R, P = [], [], [], []
    reqs = []
    for i in range(nreq):
        req = make_request(rng, max_log2=max_log2)
        order = list(range(len(req["bws"])))
        plans = candidate_plans(model, req["n"], req["bws"], req["cards"], order, seed=i, max_plans=max_plans)
        ms = measure_plans(ctx, req, plans)
        feats, _ = plan_features(model, req["n"], req["bws"], req["cards"], order, plans)
        X.extend(list(feats))
        Y.extend(list(relative_log(ms)))
        R.extend([i] * len(plans))
        P.extend(",".join(map(str, p)) for p in plans)
        reqs.append(req)
        if log:
            log(f"request {i}: n={req['n']} bws={req['bws']} plans={len(plans)} best={min(ms):.3f} ms")
    return {"X": np.stack(X), "y": np.array(Y, np.float32), "req": np.array(R, np.int32),
            "plan": np.array(P), "requests": json.dumps(reqs)}


def regret_table(pred, y, req):
    """Per request: measured time of the picked plan over the best measured
    plan, minus one (y: log times, absolute or relative within a request)."""
    out = []
    for r in np.unique(req):
        m = req == r
        t = np.exp(y[m])
        out.append(t[np.argmin(pred[m])] / t.min() - 1.0)
    return np.array(out)
