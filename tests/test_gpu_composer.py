"""Composer-level execution on the GPU: every massaging plan gives the
same result (the plans re-cut one bit string), equal to the oracle."""
import numpy as np
import pytest
import torch

import oracle_bind as ob
from fastmulticolsort_amd import datagen, planner
from fastmulticolsort_amd.engine import MODE_BATCHED, MODE_DEVICE, MODE_UNBINNED

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_two_round_sweep_all_plans_identical(ctx):
    """TwoRoundsExhaustive (src/stitch_composer.cpp:4197-4258): shift bits
    s in [-w2, w1], split (w1-s, w2+s) without zero widths."""
    n = 1 << 16
    (c1, c2), bws = datagen.config1(n)
    exp = np.lexsort((c2, c1)).astype(np.uint32)
    cols = [dev(c1), dev(c2)]
    for s in range(-bws[1], bws[0] + 1):
        split = [w for w in (bws[0] - s, bws[1] + s) if w > 0]
        order, bd, tot, ng = planner.run_plan(ctx, cols, bws, [0, 1], split)
        assert np.array_equal(order.cpu().numpy(), exp), split
        assert tot > 0 and bd[0] >= 0


@pytest.mark.parametrize("mode", [MODE_BATCHED, MODE_UNBINNED])
def test_config4_plans_match_oracle(ctx, mode):
    n = 1 << 18
    cols, bws = datagen.config4(n)
    stacked = [c.astype(np.uint64) for c in cols]
    exp = np.lexsort(tuple(reversed(stacked))).astype(np.uint32)
    dcols = [dev(c) for c in cols]
    m = planner.default_cost_model()
    card = [3, 2, 2526, 1 << 27]
    plans = [[42], [21, 21], [3, 12, 27], [16, 16, 10]] + \
        planner.plan_enum_rrs(m, n, bws, card, [0, 1, 2, 3], 6, 1)[:12]
    for split in plans:
        if mode == MODE_UNBINNED and len(split) > 1 and max(split[1:]) > 40:
            continue
        order, _, _, _ = planner.run_plan(ctx, dcols, bws, [0, 1, 2, 3], split, mode=mode)
        assert np.array_equal(order.cpu().numpy(), exp), split


def test_plan_groups_are_final_groups(ctx):
    n = 1 << 15
    (c1, c2), bws = datagen.config1(n)
    groups = torch.empty(n, dtype=torch.uint32, device="cuda")
    order, _, _, ng = planner.run_plan(ctx, [dev(c1), dev(c2)], bws, [0, 1], [25, 25], groups=groups)
    o = order.cpu().numpy()
    key = (c1.astype(np.uint64) << np.uint64(33)) | c2
    ks = key[o]
    exp = np.r_[0, np.cumsum(ks[1:] != ks[:-1])].astype(np.uint32)
    assert np.array_equal(groups.cpu().numpy(), exp) and ng == len(np.unique(key))


def test_hash_plans_count_groups_exactly(ctx):
    n = 1 << 18
    cols, bws = datagen.config4(n)
    dcols = [dev(c) for c in cols]
    stacked = np.stack([c.astype(np.uint64) for c in cols], axis=1)
    nd = len(np.unique(stacked, axis=0))
    for split in ([42], [3, 12, 27], [2, 1, 12, 27], [8, 8, 26]):
        for ht in (0, 1):
            order, gid, tot, ng = planner.run_plan_hash(ctx, dcols, bws, [0, 1, 2, 3], split, hash_type=ht)
            assert ng == nd, (split, ht)


def test_calibrate_produces_positive_costs(ctx):
    m = planner.calibrate(ctx, 1 << 20)
    d = m.to_dict()
    for b in range(3):
        assert d["radix_pass_ns"][b] > 0 and d["gather_ns"][b] > 0 and d["boundary_ns"][b] > 0
        assert all(x > 0 for x in d["segsort_ns"][b][:21])
    assert d["stitch_ns_per_byte"] > 0


@pytest.mark.parametrize("dt", [np.uint8, np.uint16, np.uint32, np.uint64])
@pytest.mark.parametrize("n", [0, 1, 5000, 100003])
def test_distinct_count_is_exact(ctx, dt, n):
    bits = 8 * np.dtype(dt).itemsize
    v = datagen.uniform_bits(31 + n, n, bits).astype(dt) if n else np.zeros(0, dt)
    if n > 10:
        v[::3] = v[0]  # many repeats
        v = v % dt(max(2, n // 7)) if bits > 8 else v
    assert ctx.distinct_count(dev(v)) == np.unique(v).size


@pytest.mark.parametrize("mode", [MODE_BATCHED, MODE_DEVICE])
def test_optimize_and_run_order_by(ctx, mode):
    # StitchComposer::optimizeAndRun: search, then run the best plan; with
    # cardinalities given and counted on the device
    n = 1 << 16
    (c1, c2), bws = datagen.config1(n)
    exp = np.lexsort((c2, c1)).astype(np.uint32)
    cols = [dev(c1), dev(c2)]
    for cards in ([8192, 1 << 33], None):
        r = planner.optimize_and_run(ctx, cols, bws, cards=cards, ordered=True, mode=mode)
        assert np.array_equal(r["order"].cpu().numpy(), exp)
        assert sum(r["split"]) == sum(bws) and r["col_order"] == [0, 1]
        assert r["opt_ms"] >= 0 and r["total_ms"] > 0


def test_optimize_and_run_group_by(ctx):
    # ordered = 0 (GROUP BY): the search may permute the columns; the rows
    # come out sorted by the chosen column order, and the groups are the
    # distinct tuples
    n = 1 << 16
    c0 = datagen.uniform_below(41, n, 3).astype(np.uint16)
    c1 = datagen.uniform_below(42, n, 2000).astype(np.uint32)
    c2 = datagen.uniform_below(43, n, 5).astype(np.uint16)
    cols_h = [c0, c1, c2]
    bws = [2, 11, 3]
    groups = torch.empty(n, dtype=torch.uint32, device="cuda")
    r = planner.optimize_and_run(ctx, [dev(c) for c in cols_h], bws, ordered=False, groups=groups)
    co = r["col_order"]
    assert sorted(co) == [0, 1, 2]
    o = r["order"].cpu().numpy()
    perm = [cols_h[c].astype(np.uint64) for c in co]
    assert np.array_equal(o, np.lexsort(tuple(reversed(perm))).astype(np.uint32))
    s = [p[o] for p in perm]
    head = np.r_[True, np.any(np.stack([x[1:] != x[:-1] for x in s]), axis=0)]
    assert np.array_equal(groups.cpu().numpy(), (np.cumsum(head) - 1).astype(np.uint32))
    assert r["ngroups"] == int(head.sum()) == len({(a, b, c) for a, b, c in zip(c0, c1, c2)})


def test_optimize_and_run_learned_picker(ctx):
    # the learned picker chooses among the candidates; any choice gives the
    # same order
    from fastmulticolsort_amd import learned
    n = 1 << 15
    (c1, c2), bws = datagen.config1(n)
    exp = np.lexsort((c2, c1)).astype(np.uint32)
    mlp = learned.Mlp(ctx, seed=0)
    r = planner.optimize_and_run(ctx, [dev(c1), dev(c2)], bws, mlp=mlp, mode=MODE_DEVICE)
    assert np.array_equal(r["order"].cpu().numpy(), exp)
    assert sum(r["split"]) == sum(bws)
    with pytest.raises(ValueError):
        planner.optimize_and_run(ctx, [dev(c1), dev(c2)], bws, mlp=mlp, ordered=False)


@pytest.mark.parametrize("mode", [MODE_BATCHED, MODE_DEVICE])
def test_descending_columns(ctx, mode):
    # per-column direction through the stitch (a descending column's bits
    # inverted within its width); plans whose rounds mix an inverted and a
    # plain source; ties keep ascending oids
    n = 1 << 16
    (c1, c2), bws = datagen.config1(n)
    c2 = c2.copy()
    c2[::5] = c2[0]  # ties in the second column too
    cols = [dev(c1), dev(c2)]
    m1, m2 = np.uint64((1 << bws[0]) - 1), np.uint64((1 << bws[1]) - 1)
    k1, k2 = c1.astype(np.uint64), c2.astype(np.uint64)
    for desc in ([True, False], [False, True], [True, True]):
        a = m1 - k1 if desc[0] else k1
        b = m2 - k2 if desc[1] else k2
        exp = np.lexsort((b, a)).astype(np.uint32)
        for split in ([50], [17, 33], [21, 29], [9, 41], [16, 16, 18]):
            order, _, _, _ = planner.run_plan(ctx, cols, bws, [0, 1], split, mode=mode, descending=desc)
            assert np.array_equal(order.cpu().numpy(), exp), (desc, split)
