"""GPU parity of the narrowed-key levels of the batched segmented sort.

In a last round (sorted keys not wanted) with 64-bit keys of which at most
32 bits are left after the first MSD level, every later level moves u32 keys.
These cases run a two-round ORDER BY whose second column is a u64 of 33 to 41
bits (41: no narrowing) in groups large enough for several MSD levels, with
uniform keys, heavy duplicates and all-equal keys (every level down to the
last key bit, the odd last level moving its rows back), through the packed
and the carried first round, eager device-driven and graph-replayed.

Bar: the order equals np.lexsort of (c2, c1), bit-exact.
"""
import numpy as np
import pytest
import torch

from fastmulticolsort_amd import datagen
from fastmulticolsort_amd.engine import MODE_DEVICE, MODE_GRAPH

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def log(*a):
    print(*a, flush=True)


def _columns(n, w1, w2, kind, seed):
    rng = np.random.default_rng(seed)
    # a few huge groups, some medium ones, many small ones
    sizes = np.concatenate([[n // 3, n // 5, n // 9], rng.integers(1, 3000, 200)])
    sizes = sizes[np.cumsum(sizes) <= n]
    sizes = np.append(sizes, n - sizes.sum())
    cand = np.unique(rng.integers(0, 1 << w1, 4 * len(sizes)))
    rng.shuffle(cand)
    codes = cand[: len(sizes)].astype(np.uint32)
    c1 = np.repeat(codes, sizes)
    c1 = c1[datagen.permutation(seed + 1, n)]
    c2 = datagen.uniform_bits(seed + 2, n, w2).astype(np.uint64)
    if kind == "dups":
        c2 = c2 % np.uint64(5)
    elif kind == "equal":
        c2[:] = c2[0]
    elif kind == "top":  # only the top 8 bits vary: one child bucket per key value
        c2 = (c2 >> np.uint64(w2 - 8)) << np.uint64(w2 - 8)
    return c1.astype(np.uint32), c2


@pytest.mark.parametrize("w2", [33, 36, 40, 41])
@pytest.mark.parametrize("kind", ["uniform", "dups", "equal", "top"])
@pytest.mark.parametrize("w1,carry", [(12, 1), (30, 1), (12, 2)])
def test_narrowed_last_round(ctx, w2, kind, w1, carry):
    n = 300007
    c1, c2 = _columns(n, w1, w2, kind, 17)
    exp = np.lexsort((c2, c1)).astype(np.uint32)
    old = ctx.tuning()
    try:
        ctx.set_tuning(old["small_max"], old["window"], old["msd_bits"], old["gather_window_kb"], carry, 0)
        for mode in (MODE_DEVICE, MODE_GRAPH):
            log("narrowed last round", w1, w2, kind, carry, mode)
            order = torch.empty(n, dtype=torch.uint32, device="cuda")
            d1, d2 = dev(c1), dev(c2)
            ctx.multi_round_sort([d1, d2], [w1, w2], order=order, mode=mode, want_count=False)
            ctx.sync()
            assert np.array_equal(order.cpu().numpy(), exp)
            if mode == MODE_GRAPH:  # a replay with new contents in the same buffers
                c2b = c2[::-1].copy()
                d2.copy_(dev(c2b))
                ctx.multi_round_sort([d1, d2], [w1, w2], order=order, mode=mode, want_count=False)
                ctx.sync()
                assert np.array_equal(order.cpu().numpy(), np.lexsort((c2b, c1)).astype(np.uint32))
    finally:
        ctx.set_tuning(old["small_max"], old["window"], old["msd_bits"], old["gather_window_kb"], old["carry_next"],
                       old["med_max"])


def test_narrowed_rounds_keep_group_counts(ctx):
    # three rounds: the middle round's sorted keys feed its boundary (no
    # narrowing there); the last round narrows
    n = 200003
    c1, c2 = _columns(n, 10, 36, "uniform", 5)
    c3 = datagen.uniform_bits(9, n, 38).astype(np.uint64)
    c2 = (c2 % np.uint64(7)).astype(np.uint64)
    order, ng = ctx.multi_round_sort([dev(c1), dev(c2), dev(c3)], [10, 36, 38], mode=MODE_DEVICE)
    assert np.array_equal(order.cpu().numpy(), np.lexsort((c3, c2, c1)).astype(np.uint32))
    assert ng == len(np.unique(np.stack([c1.astype(np.uint64), c2]), axis=1)[0])
