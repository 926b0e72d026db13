"""Host-side massaging planner (include/fmcs_plan.h) from Python.

Plan = split: the widths of the new rounds, MSB first.  The cost model is
the reference's structure (src/stitch_composer.cpp:2186-2378) with per-row
GPU times instead of CPU cycle tables.
"""
import ctypes as C

from . import _lib

GB = 32


class CostModel(C.Structure):
    _fields_ = [("radix_pass_ns", C.c_double * 3), ("segsort_ns", (C.c_double * GB) * 3),
                ("gather_ns", C.c_double * 3), ("boundary_ns", C.c_double * 3),
                ("stitch_ns_per_byte", C.c_double), ("round_overhead_ms", C.c_double),
                ("pass_overhead_ms", C.c_double), ("level_overhead_ms", C.c_double),
                ("carry_next", C.c_uint32)]

    def to_dict(self):
        return {"radix_pass_ns": list(self.radix_pass_ns),
                "segsort_ns": [list(r) for r in self.segsort_ns],
                "gather_ns": list(self.gather_ns), "boundary_ns": list(self.boundary_ns),
                "stitch_ns_per_byte": self.stitch_ns_per_byte, "round_overhead_ms": self.round_overhead_ms,
                "pass_overhead_ms": self.pass_overhead_ms, "level_overhead_ms": self.level_overhead_ms,
                "carry_next": self.carry_next}

    @classmethod
    def from_dict(cls, d):
        m = cls()
        for i in range(3):
            m.radix_pass_ns[i] = d["radix_pass_ns"][i]
            m.gather_ns[i] = d["gather_ns"][i]
            m.boundary_ns[i] = d["boundary_ns"][i]
            for j in range(GB):
                m.segsort_ns[i][j] = d["segsort_ns"][i][j]
        m.stitch_ns_per_byte = d["stitch_ns_per_byte"]
        m.round_overhead_ms = d["round_overhead_ms"]
        m.pass_overhead_ms = d.get("pass_overhead_ms", 0.015)
        m.level_overhead_ms = d.get("level_overhead_ms", 0.05)
        m.carry_next = d.get("carry_next", 1)
        return m


SortCostFn = C.CFUNCTYPE(C.c_double, C.c_void_p, C.c_double, C.c_uint32)
StitchCostFn = C.CFUNCTYPE(C.c_double, C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.c_uint64)
ReconCostFn = C.CFUNCTYPE(C.c_double, C.c_void_p, C.POINTER(C.c_uint32), C.c_uint32, C.c_uint64)


class CostPrimitives(C.Structure):
    """fmcs_cost_primitives: sort_cost(user, group_rows, bank_type),
    stitch_cost(user, assemble9, shift3, nrows), reconstruct_cost(user,
    bytewidths, nrounds, nrows)."""
    _fields_ = [("sort_cost", SortCostFn), ("stitch_cost", StitchCostFn), ("reconstruct_cost", ReconCostFn),
                ("user", C.c_void_p)]

    @classmethod
    def from_addresses(cls, sort, stitch, recon):
        """Primitives at raw C function addresses (e.g. the reference's estimators)."""
        return cls(C.cast(sort, SortCostFn), C.cast(stitch, StitchCostFn), C.cast(recon, ReconCostFn), None)


_sig_done = False


def lib():
    global _sig_done
    L = _lib.load()
    if not _sig_done:
        u32, u64, i32, vp = C.c_uint32, C.c_uint64, C.c_int, C.c_void_p
        P = C.POINTER
        sigs = {
            "fmcs_cost_model_default": [P(CostModel)],
            "fmcs_estimate_plan": [P(CostModel), u64, u32, P(u32), P(u64), P(u32), u32, P(u32), i32,
                                   P(C.c_double), P(C.c_double)],
            "fmcs_split_exhaustive": [u32, u32, u32, P(u32), u64, P(u64)],
            "fmcs_enum_banksize_category": [u32, P(u32), u64, P(u64)],
            "fmcs_greedy_plans": [P(CostModel), u64, u32, P(u32), P(u64), P(u32), u32, P(u32), u64, P(u64), P(u64)],
            "fmcs_plan_enum_rrs": [P(CostModel), u64, u32, P(u32), P(u64), P(u32), u32, u64, P(u32), u64, P(u64),
                                   P(u64)],
            "fmcs_optimize": [P(CostModel), u64, u32, P(u32), P(u64), i32, P(u32), P(u32), P(u32), P(C.c_double),
                              P(u64)],
            "fmcs_run_plan": [vp, u32, P(vp), P(u32), P(u32), P(u32), u32, P(u32), i32, u64, vp, vp,
                              P(C.c_double), P(C.c_double), P(u32)],
            "fmcs_run_plan_dir": [vp, u32, P(vp), P(u32), P(u32), P(C.c_uint8), P(u32), u32, P(u32), i32, u64, vp,
                                  vp, P(C.c_double), P(C.c_double), P(u32)],
            "fmcs_run_plan_hash": [vp, u32, P(vp), P(u32), P(u32), P(u32), u32, P(u32), i32, u64, vp, vp,
                                   P(C.c_double), P(u32)],
            "fmcs_calibrate": [vp, u64, P(CostModel)],
            "fmcs_optimize_and_run": [vp, P(CostModel), u32, P(vp), P(u32), P(u32), P(u64), i32, i32, u64, vp, vp,
                                      P(u32), P(u32), P(u32), P(C.c_double), P(C.c_double), P(C.c_double), P(u32)],
            "fmcs_plan_statistics": [u32, P(u32), P(u64), P(u32), u32, P(u32), P(C.c_double), P(u32), P(u32)],
            "fmcs_plan_statistics_u64": [u32, P(u32), P(u64), P(u32), u32, P(u32), P(u64), P(u64)],
            "fmcs_estimate_group_num": [u32, u32, P(u32), P(u64), P(u32), P(C.c_double)],
            "fmcs_column_cost_prim": [P(CostPrimitives), u64, C.c_double, u32, P(C.c_double)],
            "fmcs_estimate_plan_prim": [P(CostPrimitives), u64, u32, P(u32), P(u64), P(u32), u32, P(u32), i32,
                                        P(C.c_double)],
            "fmcs_greedy_plans_prim": [P(CostPrimitives), u64, u32, P(u32), P(u64), P(u32), u32, P(u32), u64,
                                       P(u64), P(u64)],
            "fmcs_plan_enum_rrs_prim": [P(CostPrimitives), u64, u32, P(u32), P(u64), P(u32), u32, u64, P(u32), u64,
                                        P(u64), P(u64)],
        }
        for name, args in sigs.items():
            fn = getattr(L, name)
            fn.restype = C.c_int
            fn.argtypes = args
        _sig_done = True
    return L


def _u32(xs):
    xs = list(int(x) for x in xs)
    return (C.c_uint32 * max(1, len(xs)))(*xs)


def _u64(xs):
    xs = list(int(x) for x in xs)
    return (C.c_uint64 * max(1, len(xs)))(*xs)


def _chk(rc, what):
    if rc != 0:
        raise _lib.FmcsError(rc, what)


def default_cost_model():
    m = CostModel()
    _chk(lib().fmcs_cost_model_default(C.byref(m)), "cost model")
    return m


def estimate_plan(model, nrows, bw, card, col_order, split, truncate=False):
    est = C.c_double(0)
    parts = (C.c_double * 4)()
    _chk(lib().fmcs_estimate_plan(C.byref(model), nrows, len(bw), _u32(bw), _u64(card), _u32(col_order), len(split),
                                  _u32(split), int(truncate), C.byref(est), parts), "estimate_plan")
    return est.value, {"stitch": parts[0], "reconstruct": parts[1], "sort": parts[2], "extract": parts[3]}


def split_exhaustive(S, rounds, maxv=64):
    cnt = C.c_uint64(0)
    L = lib()
    _chk(L.fmcs_split_exhaustive(S, rounds, maxv, None, 0, C.byref(cnt)), "split_exhaustive")
    n = cnt.value
    out = (C.c_uint32 * max(1, n * rounds))()
    _chk(L.fmcs_split_exhaustive(S, rounds, maxv, out, n, C.byref(cnt)), "split_exhaustive")
    return [list(out[i * rounds:(i + 1) * rounds]) for i in range(n)]


def bank_categories(rounds):
    cnt = C.c_uint64(0)
    L = lib()
    _chk(L.fmcs_enum_banksize_category(rounds, None, 0, C.byref(cnt)), "bank categories")
    n = cnt.value
    out = (C.c_uint32 * (n * rounds))()
    _chk(L.fmcs_enum_banksize_category(rounds, out, n, C.byref(cnt)), "bank categories")
    return [list(out[i * rounds:(i + 1) * rounds]) for i in range(n)]


def _unflatten(buf, words):
    plans, i = [], 0
    while i < words:
        k = buf[i]
        plans.append(list(buf[i + 1:i + 1 + k]))
        i += 1 + k
    return plans


def greedy_plans(model, nrows, bw, card, col_order, rounds):
    np_, nw = C.c_uint64(0), C.c_uint64(0)
    L = lib()
    args = (C.byref(model), nrows, len(bw), _u32(bw), _u64(card), _u32(col_order), rounds)
    _chk(L.fmcs_greedy_plans(*args, None, 0, C.byref(np_), C.byref(nw)), "greedy")
    buf = (C.c_uint32 * max(1, nw.value))()
    _chk(L.fmcs_greedy_plans(*args, buf, nw.value, C.byref(np_), C.byref(nw)), "greedy")
    return _unflatten(buf, nw.value)


def plan_enum_rrs(model, nrows, bw, card, col_order, max_round, seed):
    np_, nw = C.c_uint64(0), C.c_uint64(0)
    L = lib()
    args = (C.byref(model), nrows, len(bw), _u32(bw), _u64(card), _u32(col_order), max_round, seed)
    _chk(L.fmcs_plan_enum_rrs(*args, None, 0, C.byref(np_), C.byref(nw)), "rrs")
    buf = (C.c_uint32 * max(1, nw.value))()
    _chk(L.fmcs_plan_enum_rrs(*args, buf, nw.value, C.byref(np_), C.byref(nw)), "rrs")
    return _unflatten(buf, nw.value)


def optimize(model, nrows, bw, card, ordered=True):
    n = len(bw)
    order = (C.c_uint32 * n)()
    split = (C.c_uint32 * 64)()
    ns = C.c_uint32(0)
    cost = C.c_double(0)
    costed = C.c_uint64(0)
    _chk(lib().fmcs_optimize(C.byref(model), nrows, n, _u32(bw), _u64(card), int(bool(ordered)), order, split,
                             C.byref(ns), C.byref(cost), C.byref(costed)), "optimize")
    return list(order), list(split[: ns.value]), cost.value, costed.value


# ---- composer-level execution (GPU) -------------------------------------

def plan_statistics(bw, card, col_order, split):
    """(cardinality of each new column, assemble[3][3] flat, shift[3]) as
    getCardinalityFromStatistics computes them."""
    k = len(split)
    cd = (C.c_double * k)()
    a = (C.c_uint32 * 9)()
    sh = (C.c_uint32 * 3)()
    _chk(lib().fmcs_plan_statistics(len(bw), _u32(bw), _u64(card), _u32(col_order), k, _u32(split), cd, a, sh),
         "plan_statistics")
    return list(cd), list(a), list(sh)


def plan_statistics_u64(bw, card, col_order, split):
    """(uint64 cardinalities as PlanEnumRRS holds them, groups before each round)."""
    k = len(split)
    cu = (C.c_uint64 * k)()
    g = (C.c_uint64 * k)()
    _chk(lib().fmcs_plan_statistics_u64(len(bw), _u32(bw), _u64(card), _u32(col_order), k, _u32(split), cu, g),
         "plan_statistics_u64")
    return list(cu), list(g)


def estimate_group_num(bitpos, bw, card, col_order):
    out = C.c_double()
    _chk(lib().fmcs_estimate_group_num(bitpos, len(bw), _u32(bw), _u64(card), _u32(col_order), C.byref(out)),
         "estimate_group_num")
    return out.value


def column_cost_prim(prims, nrows, g, bank_type):
    out = C.c_double()
    _chk(lib().fmcs_column_cost_prim(C.byref(prims), nrows, g, bank_type, C.byref(out)), "column_cost_prim")
    return out.value


def estimate_plan_prim(prims, nrows, bw, card, col_order, split, truncate=False):
    out = C.c_double()
    _chk(lib().fmcs_estimate_plan_prim(C.byref(prims), nrows, len(bw), _u32(bw), _u64(card), _u32(col_order),
                                       len(split), _u32(split), int(bool(truncate)), C.byref(out)),
         "estimate_plan_prim")
    return out.value


def greedy_plans_prim(prims, nrows, bw, card, col_order, rounds):
    cap = 1 << 16
    buf = (C.c_uint32 * cap)()
    npl, words = C.c_uint64(), C.c_uint64()
    _chk(lib().fmcs_greedy_plans_prim(C.byref(prims), nrows, len(bw), _u32(bw), _u64(card), _u32(col_order), rounds,
                                      buf, cap, C.byref(npl), C.byref(words)), "greedy_plans_prim")
    return _unflatten(buf, words.value)


def plan_enum_rrs_prim(prims, nrows, bw, card, col_order, max_round, seed):
    cap = 1 << 19
    buf = (C.c_uint32 * cap)()
    npl, words = C.c_uint64(), C.c_uint64()
    _chk(lib().fmcs_plan_enum_rrs_prim(C.byref(prims), nrows, len(bw), _u32(bw), _u64(card), _u32(col_order),
                                       max_round, seed, buf, cap, C.byref(npl), C.byref(words)), "plan_enum_rrs_prim")
    return _unflatten(buf, words.value)


def run_plan(ctx, cols, bws, col_order, split, mode=0, order=None, groups=None, descending=None):
    """StitchComposer::run_single_plan on device tensors; returns
    (order, breakdown_ms[6], total_ms, ngroups).  descending: a flag per
    column (None: all ascending)."""
    import torch
    from .engine import bank_of
    n = cols[0].numel()
    if order is None:
        order = torch.empty(n, dtype=torch.uint32, device=cols[0].device)
    k = len(cols)
    cp = (C.c_void_p * k)(*[c.data_ptr() for c in cols])
    bd = (C.c_double * 6)()
    tot = C.c_double(0)
    ng = C.c_uint32(0)
    if descending is None:
        rc = lib().fmcs_run_plan(ctx.h, k, cp, _u32([bank_of(c) for c in cols]), _u32(bws), _u32(col_order),
                                 len(split), _u32(split), mode, n, C.c_void_p(order.data_ptr()),
                                 C.c_void_p(groups.data_ptr()) if groups is not None else None, bd, C.byref(tot),
                                 C.byref(ng))
    else:
        desc = (C.c_uint8 * k)(*[1 if d else 0 for d in descending])
        rc = lib().fmcs_run_plan_dir(ctx.h, k, cp, _u32([bank_of(c) for c in cols]), _u32(bws), desc,
                                     _u32(col_order), len(split), _u32(split), mode, n, C.c_void_p(order.data_ptr()),
                                     C.c_void_p(groups.data_ptr()) if groups is not None else None, bd,
                                     C.byref(tot), C.byref(ng))
    ctx._chk(rc)
    return order, list(bd), tot.value, ng.value


def optimize_and_run(ctx, cols, bws, cards=None, ordered=True, model=None, mode=0, order=None, groups=None,
                     mlp=None):
    """StitchComposer::optimizeAndRun (src/stitch_composer.cpp:3839-4013) on
    device tensors: search the plans, run the best.  cards: distinct values
    per column (None: counted on the device).  model: a cost model (None:
    the default; calibrate(ctx) gives the device's).  mlp: a trained
    learned.Mlp picks among the learned module's candidate plans instead of
    the analytic search (ordered only).  Returns a dict with the order, the
    chosen column order and split, the search time and the plan's GPU time."""
    import time
    import torch
    from .engine import bank_of
    n = cols[0].numel()
    if order is None:
        order = torch.empty(n, dtype=torch.uint32, device=cols[0].device)
    k = len(cols)
    if mlp is not None:
        if not ordered:
            raise ValueError("the learned picker covers ORDER BY (ordered) only")
        from . import learned
        t0 = time.perf_counter()
        if cards is None:
            cards = [ctx.distinct_count(c) for c in cols]
        m = model if model is not None else default_cost_model()
        ids = list(range(k))
        plans = learned.candidate_plans(m, n, bws, cards, ids, seed=0)
        pick, _ = learned.pick(mlp, m, n, bws, cards, ids, plans)
        opt_ms = (time.perf_counter() - t0) * 1e3
        _, bd, tot, ng = run_plan(ctx, cols, bws, ids, plans[pick], mode=mode, order=order, groups=groups)
        return {"order": order, "col_order": ids, "split": list(plans[pick]), "opt_ms": opt_ms, "breakdown": bd,
                "total_ms": tot, "ngroups": ng}
    m = model if model is not None else default_cost_model()
    cp = (C.c_void_p * k)(*[c.data_ptr() for c in cols])
    bo = (C.c_uint32 * k)()
    bs = (C.c_uint32 * 64)()
    ns = C.c_uint32(0)
    opt = C.c_double(0)
    bd = (C.c_double * 6)()
    tot = C.c_double(0)
    ng = C.c_uint32(0)
    rc = lib().fmcs_optimize_and_run(ctx.h, C.byref(m), k, cp, _u32([bank_of(c) for c in cols]), _u32(bws),
                                     _u64(cards) if cards is not None else None, int(bool(ordered)), mode, n,
                                     C.c_void_p(order.data_ptr()),
                                     C.c_void_p(groups.data_ptr()) if groups is not None else None, bo, bs,
                                     C.byref(ns), C.byref(opt), bd, C.byref(tot), C.byref(ng))
    ctx._chk(rc)
    return {"order": order, "col_order": list(bo), "split": list(bs[: ns.value]), "opt_ms": opt.value,
            "breakdown": list(bd), "total_ms": tot.value, "ngroups": ng.value}


def run_plan_hash(ctx, cols, bws, col_order, split, hash_type=1):
    import torch
    from .engine import bank_of
    n = cols[0].numel()
    order = torch.empty(n, dtype=torch.uint32, device=cols[0].device)
    gid = torch.empty(n, dtype=torch.uint32, device=cols[0].device)
    k = len(cols)
    cp = (C.c_void_p * k)(*[c.data_ptr() for c in cols])
    tot = C.c_double(0)
    ng = C.c_uint32(0)
    rc = lib().fmcs_run_plan_hash(ctx.h, k, cp, _u32([bank_of(c) for c in cols]), _u32(bws), _u32(col_order),
                                  len(split), _u32(split), hash_type, n, C.c_void_p(order.data_ptr()),
                                  C.c_void_p(gid.data_ptr()), C.byref(tot), C.byref(ng))
    ctx._chk(rc)
    return order, gid, tot.value, ng.value


def calibrate(ctx, nrows=1 << 24):
    m = CostModel()
    ctx._chk(lib().fmcs_calibrate(ctx.h, nrows, C.byref(m)))
    return m
