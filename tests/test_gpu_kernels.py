"""GPU parity of the HIP kernels against the CPU oracle, through the C ABI.

Bar: bit-exact.  Sorted orders are compared under the oid tie-break.
"""
import numpy as np
import pytest
import torch

import oracle_bind as ob
from fastmulticolsort_amd import datagen
from fastmulticolsort_amd._lib import FmcsError
from fastmulticolsort_amd.engine import (MODE_BATCHED, MODE_DEVICE, MODE_PLAIN, MODE_UNBINNED, construct_entries)

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.cpu().numpy()


def log(*a):
    print(*a, flush=True)


# ---- stitch ---------------------------------------------------------------

STITCH_CASES = [
    ([20, 36], [24, 32], 1000),             # bit_borrow shape (tests/misc_test.cpp:1263-1295)
    ([12, 20, 30], [62], 4099),             # 62-bit stitch
    ([17, 33], [10, 40], 777),              # borrow 7 bits forward
    ([17, 33], [50], 4096),                 # single round
    ([2, 1, 12, 27], [3, 12, 27], 5000),
    ([2, 1, 12, 27], [16, 16, 10], 3333),
    ([63, 20], [31, 21, 21, 10], 2048),     # stitchTest layout (tests/misc_test.cpp:741-773)
    ([5, 7], [1, 11], 1),
    ([64], [64], 100),
]


@pytest.mark.parametrize("bws,split,n", STITCH_CASES)
def test_stitch_matches_oracle(ctx, bws, split, n):
    log("stitch", bws, split, n)
    cols = [datagen.uniform_bits(1000 + i, n, w).astype(datagen.bank_dtype(w)) for i, w in enumerate(bws)]
    ents = construct_entries(bws, list(range(len(bws))), split)
    assert ents == ob.construct_entries(bws, list(range(len(bws))), split)
    exp = ob.stitch(cols, bws, ents, split)
    dsts = [torch.zeros(n, dtype=torch.from_numpy(e[:0]).dtype, device="cuda") for e in exp]
    ctx.stitch([dev(c) for c in cols], bws, ents, dsts)
    for d, e in zip(dsts, exp):
        assert np.array_equal(host(d), e)


def test_stitch_bit_borrow_known_answer(ctx):
    # tests/misc_test.cpp:1263-1295: 0x12345 (20b) | 0x678912345 (36b) -> 24+32
    c1 = np.full(8, 0x12345, np.uint32)
    c2 = np.full(8, 0x678912345, np.uint64)
    ents = construct_entries([20, 36], [0, 1], [24, 32])
    d1 = torch.zeros(8, dtype=torch.uint32, device="cuda")
    d2 = torch.zeros(8, dtype=torch.uint32, device="cuda")
    ctx.stitch([dev(c1), dev(c2)], [20, 36], ents, [d1, d2])
    assert host(d1)[0] == 0x123456 and host(d2)[0] == 0x78912345


# ---- gather / boundaries ---------------------------------------------------

@pytest.mark.parametrize("dt", [np.uint16, np.uint32, np.uint64])
@pytest.mark.parametrize("n", [1, 7, 4096, 100003])
def test_gather(ctx, dt, n):
    v = datagen.splitmix64(7, n).astype(dt)
    order = datagen.permutation(3, n).astype(np.uint32)
    out = ctx.gather(dev(v), dev(order))
    assert np.array_equal(host(out), v[order])


_WDT = {2: np.uint16, 4: np.uint32, 8: np.uint64}


@pytest.mark.parametrize("widths", [[2, 4, 4], [8, 8, 8], [4, 4], [2, 2], [8, 2, 2, 4], [2] * 9, [8], [4, 8, 2]])
@pytest.mark.parametrize("n", [1, 3, 4, 5, 1027, 100003])
@pytest.mark.parametrize("repeats", [False, True])
def test_gather_columns(ctx, widths, n, repeats):
    # several columns by one order through packed row records (groups of
    # 2..8 columns in 8- or 16-byte records, lone columns through the column
    # gather), ragged tails of the 4-row vectors, orders with repeated rows
    log("gather_columns", widths, n, repeats)
    cols = [datagen.splitmix64(11 + k, n).astype(_WDT[w]) for k, w in enumerate(widths)]
    if repeats:
        order = (datagen.splitmix64(5, n) % np.uint64(n)).astype(np.uint32)
    else:
        order = datagen.permutation(6, n).astype(np.uint32)
    outs = ctx.gather_columns([dev(c) for c in cols], dev(order))
    for c, o in zip(cols, outs):
        assert np.array_equal(host(o), c[order])


def test_gather_columns_unaligned_and_errors(ctx):
    # columns whose pointers are not 16-byte aligned take the column gather
    n = 5003
    base = [datagen.splitmix64(21 + k, n + 1).astype(dt) for k, dt in enumerate([np.uint16, np.uint32, np.uint16])]
    dbase = [dev(b) for b in base]
    cols = [d[1:] for d in dbase]
    order = datagen.permutation(9, n).astype(np.uint32)
    outs = ctx.gather_columns(cols, dev(order))
    for b, o in zip(base, outs):
        assert np.array_equal(host(o), b[1:][order])
    import ctypes as C
    lib = ctx.lib
    nul = (C.c_void_p * 1)(None)
    bad = (C.c_uint32 * 1)(3)
    assert lib.fmcs_gather_columns(ctx.h, n, 0, nul, bad, None, C.c_void_p(dev(order).data_ptr()), nul) != 0
    assert lib.fmcs_gather_columns(ctx.h, n, 1, nul, bad, None, C.c_void_p(dev(order).data_ptr()), nul) != 0


@pytest.mark.parametrize("widths,bits", [([2, 4, 4], [12, 20, 30]), ([8, 4], [40, 24]), ([2, 2, 4, 8], [16, 9, 7, 32]),
                                         ([2, 4, 4], [16, 32, 32]), ([8, 8], [33, 31]), ([4] * 9, [7] * 9),
                                         ([8, 2], [64, 1])])
@pytest.mark.parametrize("n", [1, 5, 1027, 100003])
def test_gather_columns_bits(ctx, widths, bits, n):
    # declared widths: columns summing to at most 64 bits share an 8-byte
    # record at bit offsets (taken when it needs no more records a row),
    # values of exactly their declared width, ragged tails, repeated rows
    log("gather_columns bits", widths, bits, n)
    cols = [(datagen.splitmix64(31 + k, n) & np.uint64((1 << b) - 1)).astype(_WDT[w])
            for k, (w, b) in enumerate(zip(widths, bits))]
    order = (datagen.splitmix64(7, n) % np.uint64(n)).astype(np.uint32)
    outs = ctx.gather_columns([dev(c) for c in cols], dev(order), bits=bits)
    for c, o in zip(cols, outs):
        assert np.array_equal(host(o), c[order])


def test_gather_columns_bits_too_wide(ctx):
    # a value above its declared width: FMCS_E_ARG and no output written
    n = 4099
    cols = [(datagen.splitmix64(41 + k, n) & np.uint64((1 << b) - 1)).astype(dt)
            for k, (dt, b) in enumerate([(np.uint16, 12), (np.uint32, 20), (np.uint32, 30)])]
    cols[1][777] = 1 << 20
    outs = [torch.zeros(n, dtype=torch.from_numpy(c).dtype, device="cuda") for c in cols]
    with pytest.raises(FmcsError):
        ctx.gather_columns([dev(c) for c in cols], dev(datagen.permutation(2, n).astype(np.uint32)), outs=outs,
                           bits=[12, 20, 30])
    assert not any(host(o).any() for o in outs)
    with pytest.raises(FmcsError):
        ctx.gather_columns([dev(c) for c in cols], dev(np.arange(n, dtype=np.uint32)), bits=[12, 33, 30])


@pytest.mark.parametrize("dt", [np.uint16, np.uint32, np.uint64])
@pytest.mark.parametrize("kb", [1, 3, 64])
def test_gather_windowed(ctx, dt, kb):
    # passes over source windows of kb KiB (a window edge inside a 4-row
    # vector, a last partial window) give the one-pass result
    n = 100003
    old = ctx.tuning()
    try:
        ctx.set_tuning(old["small_max"], old["window"], old["msd_bits"], kb)
        v = datagen.splitmix64(8, n).astype(dt)
        order = datagen.permutation(4, n).astype(np.uint32)
        order[: n // 2] = np.sort(order[: n // 2])  # ascending runs, as in a sorted group
        out = ctx.gather(dev(v), dev(order))
        assert np.array_equal(host(out), v[order])
    finally:
        ctx.set_tuning(old["small_max"], old["window"], old["msd_bits"], old["gather_window_kb"], old["carry_next"])


@pytest.mark.parametrize("dt", [np.uint16, np.uint32, np.uint64])
@pytest.mark.parametrize("n", [1, 2, 8191, 8192, 8193, 300001])
@pytest.mark.parametrize("with_in_group", [False, True])
def test_group_offsets_without_ids(ctx, dt, n, with_in_group):
    # no per-row ids wanted: the count pass writes every row's head bit and
    # the write pass reads those bits instead of the keys (8192-row tiles)
    log("boundary offsets only", dt, n, with_in_group)
    import ctypes as C
    v = np.sort(datagen.uniform_below(13, n, max(2, n // 7)).astype(dt))
    ing = np.sort(datagen.uniform_below(14, n, max(2, n // 40))).astype(np.uint32)
    exp, ng = ob.extract_group_info(v, ing) if with_in_group else ob.extract_group_info(v)
    heads = np.flatnonzero(np.r_[1, np.diff(exp.astype(np.int64)) != 0])
    dv, di = dev(v), dev(ing)
    off = torch.full((n + 1,), 0xFFFFFFFF, dtype=torch.uint32, device="cuda")
    cnt = C.c_uint32(0)
    rc = ctx.lib.fmcs_extract_group_info(ctx.h, v.dtype.itemsize, n, C.c_void_p(dv.data_ptr()),
                                         C.c_void_p(di.data_ptr()) if with_in_group else None, None,
                                         C.c_void_p(off.data_ptr()), C.byref(cnt))
    assert rc == 0
    assert cnt.value == ng
    assert np.array_equal(host(off)[: ng + 1], np.r_[heads, n].astype(np.uint32))


@pytest.mark.parametrize("dt", [np.uint16, np.uint32, np.uint64])
@pytest.mark.parametrize("n", [1, 2, 4095, 4096, 4097, 300001])
def test_extract_group_info(ctx, dt, n):
    log("boundary", dt, n)
    v = np.sort(datagen.uniform_below(11, n, max(2, n // 5)).astype(dt))
    ing = np.sort(datagen.uniform_below(12, n, max(2, n // 50))).astype(np.uint32)
    exp0, ng0 = ob.extract_group_info(v)
    exp1, ng1 = ob.extract_group_info(v, ing)
    off = torch.zeros(n + 1, dtype=torch.uint32, device="cuda")
    g0, m0 = ctx.extract_group_info(dev(v), seg_offsets=off)
    assert m0 == ng0 and np.array_equal(host(g0), exp0)
    heads = np.flatnonzero(np.r_[1, np.diff(exp0.astype(np.int64)) != 0])
    assert np.array_equal(host(off)[: ng0 + 1], np.r_[heads, n].astype(np.uint32))
    g1, m1 = ctx.extract_group_info(dev(v), in_group=dev(ing))
    assert m1 == ng1 and np.array_equal(host(g1), exp1)


# ---- sort_pairs (LSD onesweep) ---------------------------------------------

SORT_CASES = [
    (np.uint16, 16, 1), (np.uint16, 16, 4096), (np.uint16, 12, 70001),
    (np.uint32, 17, 4097), (np.uint32, 32, 262145), (np.uint32, 20, 1 << 20),
    (np.uint64, 33, 100000), (np.uint64, 62, 300007), (np.uint64, 64, 65536), (np.uint64, 40, 12345),
]


@pytest.mark.parametrize("dt,bits,n", SORT_CASES)
def test_sort_pairs(ctx, dt, bits, n):
    log("sort_pairs", dt, bits, n)
    k = datagen.uniform_bits(21, n, bits).astype(dt)
    if n > 10:
        k[: n // 3] = k[0]  # heavy ties: stability is visible
    oids = np.arange(n, dtype=np.uint32)
    kd, od = dev(k), dev(oids)
    ctx.sort_pairs(kd, od, bits)
    exp = np.argsort(k, kind="stable").astype(np.uint32)
    assert np.array_equal(host(od), exp)
    assert np.array_equal(host(kd), k[exp])


@pytest.mark.parametrize("dt,bits,bad_bit", [(np.uint32, 20, 21), (np.uint32, 20, 31), (np.uint16, 12, 15),
                                              (np.uint64, 33, 63), (np.uint64, 40, 41)])
def test_sort_pairs_reports_keys_wider_than_declared(ctx, dt, bits, bad_bit):
    # any bit at or above the declared width is refused, also one far above
    # the last digit (the digits are 6-8 bits wide, split evenly)
    n = 50000
    log("sort_pairs over-wide key", dt, bits, bad_bit)
    k = datagen.uniform_bits(4, n, bits).astype(dt)
    k[n // 2] |= dt(1) << dt(bad_bit)
    with pytest.raises(FmcsError):
        ctx.sort_pairs(dev(k), dev(np.arange(n, dtype=np.uint32)), bits)
    k[n // 2] &= dt((1 << bits) - 1)
    kd, od = dev(k), dev(np.arange(n, dtype=np.uint32))
    ctx.sort_pairs(kd, od, bits)
    assert np.array_equal(host(od), np.argsort(k, kind="stable").astype(np.uint32))


@pytest.mark.parametrize("dt,bits,n", SORT_CASES + [(np.uint64, 62, 3 * 2048 + 5), (np.uint32, 31, 2048 * 9)])
def test_sort_pairs_merge(ctx, dt, bits, n):
    # bitonic tiles + merge-path passes: odd run counts, a partial last
    # tile, heavy ties (ordered by oid), incoming oids not ascending
    log("sort_pairs_merge", dt, bits, n)
    k = datagen.uniform_bits(22, n, bits).astype(dt)
    if n > 10:
        k[: n // 3] = k[0]
    oids = datagen.permutation(23, n).astype(np.uint32)
    kd, od = dev(k), dev(oids)
    ctx.sort_pairs_merge(kd, od)
    exp = np.lexsort((oids, k))
    assert np.array_equal(host(od), oids[exp])
    assert np.array_equal(host(kd), k[exp])


def test_sort_pairs_matches_reference_mergesort(ctx):
    if ob.ref() is None:
        pytest.skip("reference library not built here")
    n = 50000
    k = datagen.uniform_bits(5, n, 30).astype(np.uint32)
    o = np.arange(n, dtype=np.uint32)
    rk, ro = ob.ref_mergesort(k, o)
    kd, od = dev(k), dev(o)
    ctx.sort_pairs(kd, od, 30)
    assert np.array_equal(host(kd), rk)
    assert np.array_equal(host(od), ob.canonical_order(rk, ro))


# ---- segmented sort ----------------------------------------------------------

def _segments(n, seed, kind):
    if kind == "zipf":
        sizes = datagen.zipf_counts(n, max(1, n // 40), 1.1)
    elif kind == "tiny":
        sizes = datagen.zipf_counts(n, max(1, n // 3), 0.3)
    elif kind == "one":
        sizes = np.array([n])
    elif kind == "alt":
        # large segments separated by runs of small ones: MSD bucket lists
        # have gaps between parents that windows must not span
        pat = np.array([3000, 5, 9, 1, 0, 2500, 40, 1300, 7, 1100, 1025, 2, 60000])
        reps = n // int(pat.sum()) + 1
        sizes = np.tile(pat, reps)
        cs = np.cumsum(sizes)
        k = int(np.searchsorted(cs, n))
        sizes = sizes[: k + 1].copy()
        sizes[-1] -= int(cs[k] - n)
    else:  # mixed: huge + many small + singles
        rest = n - n // 2
        small = datagen.zipf_counts(rest, max(1, rest // 20), 0.8)
        sizes = np.r_[small[: len(small) // 2], n // 2, small[len(small) // 2:]]
    off = np.r_[0, np.cumsum(sizes)].astype(np.uint32)
    return off


@pytest.mark.parametrize("mode", [MODE_BATCHED, MODE_DEVICE, MODE_UNBINNED, MODE_PLAIN])
@pytest.mark.parametrize("dt,bits", [(np.uint64, 40), (np.uint32, 20), (np.uint16, 9), (np.uint64, 60)])
@pytest.mark.parametrize("kind", ["zipf", "tiny", "one", "mixed", "alt"])
def test_segmented_sort(ctx, mode, dt, bits, kind):
    n = 200003 if mode != MODE_PLAIN else 40009
    off = _segments(n, 3, kind)
    nseg = len(off) - 1
    # MODE_UNBINNED with (segment, key) wider than 64 bits (60-bit keys over
    # many segments) takes two stable sorts, by key then by segment
    log("segsort", mode, dt, bits, kind, n, nseg)
    k = datagen.uniform_bits(31, n, bits).astype(dt)
    k[::7] = k[0]
    oids = datagen.permutation(9, n).astype(np.uint32)
    # within a segment incoming rows are in ascending oid order
    for s in range(nseg):
        oids[off[s]:off[s + 1]] = np.sort(oids[off[s]:off[s + 1]])
    kd, od = dev(k), dev(oids)
    info = ctx.segmented_sort(kd, od, dev(off), nseg, bits, mode=mode, want_info=True)
    seg = np.repeat(np.arange(nseg), np.diff(off.astype(np.int64)))
    idx = np.lexsort((oids, k, seg))
    assert np.array_equal(host(od), oids[idx])
    assert np.array_equal(host(kd), k[idx])
    sz = np.diff(off.astype(np.int64))
    assert info.ngroups == nseg and info.ngroups_gtOne == int((sz > 1).sum())


# ---- sort_by_group / multi-round ----------------------------------------------

@pytest.mark.parametrize("mode", [MODE_BATCHED, MODE_DEVICE, MODE_UNBINNED, MODE_PLAIN])
def test_multi_round_sort_config1_shape(ctx, mode):
    n = 1 << 16 if mode != MODE_PLAIN else 1 << 14
    (c1, c2), bws = datagen.config1(n)
    log("multi_round config1-shape", mode, n)
    order, ng = ctx.multi_round_sort([dev(c1), dev(c2)], bws, mode=mode)
    eo, eg, eng, _ = ob.multi_round_sort([c1, c2])
    assert np.array_equal(host(order), eo)
    assert ng == eng
    assert np.array_equal(host(order), np.lexsort((c2, c1)).astype(np.uint32))


def test_sort_by_group_rounds_match_oracle(ctx):
    n = 1 << 17
    (c1, c2), bws = datagen.config3(n, ngroups=1 << 10)
    order = dev(np.arange(n, dtype=np.uint32))
    o0 = np.arange(n, dtype=np.uint32)
    g1, ng1, _ = ctx.sort_by_group(dev(c1), order, 20)
    eo1, eg1, eng1, _, _ = ob.sort_by_group(c1, o0)
    assert np.array_equal(host(order), eo1) and np.array_equal(host(g1), eg1) and ng1 == eng1
    g2, ng2, info = ctx.sort_by_group(dev(c2), order, 40, in_group=g1, want_info=True)
    eo2, eg2, eng2, _, einfo = ob.sort_by_group(c2, eo1, eg1)
    assert np.array_equal(host(order), eo2) and np.array_equal(host(g2), eg2) and ng2 == eng2
    assert (info.ngroups, info.ngroups_gtOne, info.sum_group_size, info.ngroups_incache) == \
        (einfo.ngroups, einfo.ngroups_gtOne, einfo.sum_group_size, einfo.ngroups_incache)


def test_multi_round_matches_reference(ctx):
    if ob.ref() is None:
        pytest.skip("reference library not built here")
    n = 1 << 15
    (c1, c2), bws = datagen.config1(n)
    ro, rg = ob.ref_sort_rounds([c1, c2], bws)
    order, _ = ctx.multi_round_sort([dev(c1), dev(c2)], bws)
    keys = np.stack([c1[ro], c2[ro]])
    assert np.array_equal(host(order), ob.canonical_order(keys, ro))


# ---- synthetic column generator ----------------------------------------------

@pytest.mark.parametrize("dt,w,card", [(np.uint16, 12, 2526), (np.uint32, 27, 1 << 20), (np.uint64, 40, 3),
                                       (np.uint64, 40, 1 << 40)])
def test_fill_column(ctx, dt, w, card):
    n = 100003
    out = torch.empty(n, dtype={np.uint16: torch.uint16, np.uint32: torch.uint32, np.uint64: torch.uint64}[dt],
                      device="cuda")
    u = datagen.splitmix64(77, n)
    step = max(1, (1 << w) // card)
    mask = np.uint64((1 << w) - 1)
    # uniform: exact
    ctx.fill_column(out, 77, card, False, w)
    exp = ((u % np.uint64(card)) * np.uint64(step)) & mask
    assert np.array_equal(host(out).astype(np.uint64), exp)
    # Zipf(1): floor(card^x) - 1 in double precision (a last-ulp difference of
    # exp2 may move a row across a rank edge: allow a handful)
    ctx.fill_column(out, 77, card, True, w)
    x = (u >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    rank = np.minimum(np.maximum(np.floor(np.exp2(x * np.log2(float(card)))), 1).astype(np.uint64) - 1,
                      np.uint64(card - 1))
    got = host(out).astype(np.uint64)
    assert (got != ((rank * np.uint64(step)) & mask)).sum() <= 5
    assert got.max() <= int(mask)
