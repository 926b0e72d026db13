"""Device-driven and graph-captured multi-round sort: parity with the
oracle and with the eager batched path, replay after replay.

The graph is replayed several times with new column contents written into
the same device buffers, with eager work and host synchronisations between
the replays; every replay is checked.  Bar: bit-exact.
"""
import numpy as np
import pytest
import torch

import oracle_bind as ob
from fastmulticolsort_amd import datagen
from fastmulticolsort_amd._lib import FmcsError
from fastmulticolsort_amd.engine import MODE_BATCHED, MODE_DEVICE, MODE_GRAPH

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.cpu().numpy()


def log(*a):
    print(*a, flush=True)


def lex(cols):
    return np.lexsort(tuple(reversed(cols))).astype(np.uint32)


def test_graph_replays_match_eager_and_oracle(ctx):
    n = 1 << 18
    (c1, c2), bws = datagen.config3(n, ngroups=1 << 10)
    d1, d2 = dev(c1), dev(c2)
    order = torch.empty(n, dtype=torch.uint32, device="cuda")
    groups = torch.empty(n, dtype=torch.uint32, device="cuda")
    eager = torch.empty(n, dtype=torch.uint32, device="cuda")
    for rep in range(5):
        # new contents in the same buffers: a replay must sort what is there now
        k1 = datagen.zipf_column(100 + rep, n, 1 << 10, 1.1,
                                 datagen.permutation(7 + rep, 1 << 20)[: 1 << 10].astype(np.uint32)).astype(np.uint32)
        k2 = datagen.uniform_bits(200 + rep, n, 40).astype(np.uint64)
        if rep == 3:
            k2[::3] = k2[0]  # many equal keys: ties broken by oid
        d1.copy_(dev(k1))
        d2.copy_(dev(k2))
        log("graph replay", rep, n)
        _, ng = ctx.multi_round_sort([d1, d2], bws, order=order, groups=groups, mode=MODE_GRAPH,
                                     want_count=(rep % 2 == 0))
        torch.cuda.synchronize()
        ctx.multi_round_sort([d1, d2], bws, order=eager, mode=MODE_BATCHED)
        torch.cuda.synchronize()
        exp = lex([k1, k2])
        got = host(order)
        assert np.array_equal(got, exp), f"replay {rep}"
        assert np.array_equal(got, host(eager))
        eo, eg, eng, _ = ob.multi_round_sort([k1, k2])
        assert np.array_equal(got, eo)
        # dense group ids of the last round over (c1, c2)
        s1, s2 = k1[got], k2[got]
        head = np.r_[True, (s1[1:] != s1[:-1]) | (s2[1:] != s2[:-1])]
        assert np.array_equal(host(groups), (np.cumsum(head) - 1).astype(np.uint32))
        if rep % 2 == 0:
            assert ng == int(head.sum())


def test_graph_recaptures_on_tuning_and_shape_change(ctx):
    old = ctx.tuning()
    try:
        for (n, sm, w, mb, gkb, cn) in [(1 << 16, 512, 512, 8, 0, 1), (1 << 16, 256, 128, 6, 64, 0),
                                        (1 << 16, 256, 128, 6, 64, 1), (50001, 256, 128, 6, 0, 0),
                                        (50001, 512, 512, 8, 96, 0), (50001, 512, 512, 8, 96, 1)]:
            (c1, c2), bws = datagen.config3(n, ngroups=1 << 8)
            ctx.set_tuning(sm, w, mb, gkb, cn)
            d1, d2 = dev(c1), dev(c2)
            order = torch.empty(n, dtype=torch.uint32, device="cuda")
            log("graph recapture", n, sm, w, mb, gkb, cn)
            for rep in range(3):
                ctx.multi_round_sort([d1, d2], bws, order=order, mode=MODE_GRAPH, want_count=False)
                torch.cuda.synchronize()
                assert np.array_equal(host(order), lex([c1, c2])), (n, sm, rep)
    finally:
        ctx.set_tuning(old["small_max"], old["window"], old["msd_bits"], old["gather_window_kb"], old["carry_next"])


@pytest.mark.parametrize("mode", [MODE_DEVICE, MODE_GRAPH])
def test_carried_second_column_matches_gathered(ctx, mode):
    # Round 1's radix passes carry round 2's column (every key / payload bank
    # pair; 1, 2 and 3 passes; tiny inputs; a third round gathered after a
    # carried second), packed into u64 composites where the widths allow
    # (carry_next 1: u16/u32 first column, u64 second, widths summing to at
    # most 64, 1 to 4 passes) or as a separate column (carry_next 2): same
    # order and group ids as gathering the column (carry_next 0).
    old = ctx.tuning()
    cases = [([4, 12], 5000), ([12, 30], 70001), ([20, 40], 1 << 17), ([30, 16], 33333), ([40, 60], 20000),
             ([16, 2], 4097), ([8, 8], 3), ([1, 1], 2), ([20, 40, 10], 40000), ([50, 9, 33], 30001),
             ([16, 48], 50001), ([8, 56], 20000), ([31, 33], 40000), ([20, 44], 3), ([24, 40, 10], 30001),
             ([1, 63], 4097)]
    try:
        for bws, n in cases:
            cols = []
            for i, w in enumerate(bws):
                c = datagen.uniform_bits(300 + i, n, w)
                if i == 0:
                    c = c % np.uint64(max(2, n // 37)) if w > 8 else c  # groups of ~37 rows
                cols.append(c.astype(datagen.bank_dtype(w)))
            exp = lex(cols)
            s = [c[exp].astype(np.uint64) for c in cols]
            head = np.r_[True, np.any(np.stack([x[1:] != x[:-1] for x in s]), axis=0)] if n > 1 else np.ones(n, bool)
            egroups = (np.cumsum(head) - 1).astype(np.uint32)
            dc = [dev(c) for c in cols]
            for cn in (1, 2, 0):
                ctx.set_tuning(old["small_max"], old["window"], old["msd_bits"], old["gather_window_kb"], cn)
                order = torch.empty(n, dtype=torch.uint32, device="cuda")
                groups = torch.empty(n, dtype=torch.uint32, device="cuda")
                log("carried column", mode, bws, n, cn)
                for rep in range(2):
                    _, ng = ctx.multi_round_sort(dc, bws, order=order, groups=groups, mode=mode)
                    torch.cuda.synchronize()
                    assert np.array_equal(host(order), exp), (bws, n, cn, rep)
                    assert np.array_equal(host(groups), egroups), (bws, n, cn, rep)
                    assert ng == int(head.sum())
    finally:
        ctx.set_tuning(old["small_max"], old["window"], old["msd_bits"], old["gather_window_kb"], old["carry_next"])


@pytest.mark.parametrize("mode", [MODE_DEVICE, MODE_GRAPH])
def test_2048_row_tiles_and_single_windows(ctx, mode):
    # small_max 1024 selects 2048-row tiles: segments of 513..1024 rows are
    # packed into LDS tiles.  With small_max 512 (1024-row tiles), segments
    # of 513..1024 and 1025..2048 rows are sorted alone in 1024- and
    # 2048-row tiles.  Distinct keys (bucket sort), heavy duplicates (the
    # radix tile sort takes the refused tiles) and 60-bit keys (no room for a
    # packed segment index: comparison rank) are all checked.
    old = ctx.tuning()
    n = 1 << 18
    cases = [("zipf", *datagen.config3(n, ngroups=1 << 8)[0], [20, 40])]
    for gs in (800, 1500):
        g = datagen.uniform_below(21 + gs, n, n // gs).astype(np.uint32)
        # wide keys with heavy duplicates: windows the bucket sort refuses go
        # to the radix tile sort over the bits that vary inside them (a
        # constant high part), or to the comparison rank (three distinct
        # full-width values)
        wide_dup = (np.uint64(0xABCDE) << np.uint64(40)) | datagen.uniform_below(25, n, 3).astype(np.uint64)
        three = datagen.splitmix64(26, 3)
        cases += [(f"distinct{gs}", g, datagen.uniform_bits(22, n, 40), [20, 40]),
                  (f"duplicates{gs}", g, datagen.uniform_below(23, n, 3).astype(np.uint64), [20, 40]),
                  (f"wide{gs}", g, datagen.uniform_bits(24, n, 60), [20, 60]),
                  (f"wide_dup{gs}", g, wide_dup, [20, 60]),
                  (f"wide_dup64_{gs}", g, three[datagen.uniform_below(27, n, 3).astype(np.int64)], [20, 64])]
    try:
        for name, c1, c2, bws in cases:
            c1 = c1.astype(np.uint32)
            c2 = c2.astype(np.uint64)
            exp = lex([c1, c2])
            eo, _, eng, _ = ob.multi_round_sort([c1, c2])
            assert np.array_equal(exp, eo)
            d1, d2 = dev(c1), dev(c2)
            for sm, w in ((1024, 1024), (1024, 512), (512, 512)):
                ctx.set_tuning(sm, w, 8, 0, 1)
                order = torch.empty(n, dtype=torch.uint32, device="cuda")
                log("2048-row tiles / single windows", mode, name, sm, w)
                for rep in range(2):
                    _, ng = ctx.multi_round_sort([d1, d2], bws, order=order, mode=mode)
                    torch.cuda.synchronize()
                    assert np.array_equal(host(order), exp), (name, w, rep)
                    assert ng == eng
    finally:
        ctx.set_tuning(old["small_max"], old["window"], old["msd_bits"], old["gather_window_kb"], old["carry_next"])


@pytest.mark.parametrize("mode", [MODE_BATCHED, MODE_DEVICE, MODE_GRAPH])
def test_packing_window_after_a_parent_that_fills_its_tiles(ctx, mode):
    # Regression: a second-level parent bucket of exactly 4096 rows (one
    # whole scatter tile) whose digits leave empty buckets at its end.  Those
    # empty buckets start at the parent's end; their packing window must stay
    # in the parent's range, not take the next parent's first window (which
    # would stretch that window back over the rows between the two parents,
    # past the LDS tile).
    rng = np.random.default_rng(1)

    def part(m, top, lim):
        return ((np.uint64(top) << np.uint64(32)) | (rng.integers(0, lim, m).astype(np.uint64) << np.uint64(24))
                | rng.integers(0, 1 << 24, m).astype(np.uint64))

    for lead in (0, 5):  # lead: rows of a small group before the first parent
        c2 = np.concatenate([rng.integers(0, 1 << 40, lead).astype(np.uint64), part(4096, 0x10, 0x80),
                             rng.integers(0, 1 << 40, 1000).astype(np.uint64), part(3000, 0x20, 0x100)])
        c1 = np.concatenate([np.full(lead, 0), np.full(4096, 1), np.full(1000, 2), np.full(3000, 3)]).astype(np.uint32)
        n = c1.size
        order = torch.empty(n, dtype=torch.uint32, device="cuda")
        log("packing window after a tile-filling parent", mode, lead)
        _, ng = ctx.multi_round_sort([dev(c1), dev(c2)], [20, 40], order=order, mode=mode)
        torch.cuda.synchronize()
        exp = lex([c1, c2])
        assert np.array_equal(host(order), exp)
        eo, _, eng, _ = ob.multi_round_sort([c1, c2])
        assert np.array_equal(host(order), eo) and ng == eng


@pytest.mark.parametrize("mode", [MODE_DEVICE, MODE_GRAPH])
def test_device_modes_four_rounds_config4_shape(ctx, mode):
    n = 1 << 17
    cols, bws = datagen.config4(n)
    log("device modes config4-shape", mode, n)
    dc = [dev(c) for c in cols]
    order = torch.empty(n, dtype=torch.uint32, device="cuda")
    for rep in range(2):
        _, ng = ctx.multi_round_sort(dc, bws, order=order, mode=mode)
        eo, _, eng, _ = ob.multi_round_sort(cols)
        assert np.array_equal(host(order), eo)
        assert np.array_equal(host(order), lex(cols))
        assert ng == eng


@pytest.mark.parametrize("mode", [MODE_DEVICE, MODE_GRAPH])
def test_device_modes_wide_keys_and_single_rows(ctx, mode):
    # 60-bit second column: levels whose key bits leave no room for a packed
    # segment index (single rows only), many MSD levels; plus tiny inputs
    for n in (1, 2, 3, 4097, 70001):
        c1 = datagen.uniform_below(5, n, 3).astype(np.uint32)
        c2 = datagen.uniform_bits(6, n, 60).astype(np.uint64)
        c2[::5] = c2[0]
        order = torch.empty(n, dtype=torch.uint32, device="cuda")
        log("device modes wide", mode, n)
        _, ng = ctx.multi_round_sort([dev(c1), dev(c2)], [2, 60], order=order, mode=mode)
        assert np.array_equal(host(order), lex([c1, c2])), n
        eo, _, eng, _ = ob.multi_round_sort([c1, c2])
        assert ng == eng


def test_packed_composite_reports_second_column_wider_than_declared(ctx):
    # the packed form (20-bit u32 first column, 40-bit u64 second) checks the
    # second column too: a 41-bit value would corrupt the first column's bits
    n = 10000
    c1 = datagen.uniform_bits(8, n, 20).astype(np.uint32)
    c2 = datagen.uniform_bits(9, n, 40).astype(np.uint64)
    c2[n // 2] |= np.uint64(1 << 40)
    order = torch.empty(n, dtype=torch.uint32, device="cuda")
    with pytest.raises(FmcsError):
        ctx.multi_round_sort([dev(c1), dev(c2)], [20, 40], order=order, mode=MODE_DEVICE)
    c2[n // 2] &= np.uint64((1 << 40) - 1)
    ctx.multi_round_sort([dev(c1), dev(c2)], [20, 40], order=order, mode=MODE_DEVICE)
    assert np.array_equal(host(order), lex([c1, c2]))


def test_device_mode_reports_keys_wider_than_declared(ctx):
    n = 10000
    c1 = datagen.uniform_bits(8, n, 24).astype(np.uint32)  # declared as 20 bits below
    c2 = datagen.uniform_bits(9, n, 40).astype(np.uint64)
    order = torch.empty(n, dtype=torch.uint32, device="cuda")
    with pytest.raises(FmcsError):
        ctx.multi_round_sort([dev(c1), dev(c2)], [20, 40], order=order, mode=MODE_DEVICE)
    # a later good call is unaffected
    c1 &= (1 << 20) - 1
    ctx.multi_round_sort([dev(c1), dev(c2)], [20, 40], order=order, mode=MODE_DEVICE)
    assert np.array_equal(host(order), lex([c1, c2]))
