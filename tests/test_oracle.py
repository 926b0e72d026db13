"""Pins the CPU oracle: known answers from the reference's tests, closed
forms, numpy restatements, and (where it was built here) the reference
itself.  CPU only."""
import itertools
import math

import numpy as np
import pytest

import oracle_bind as ob
from fastmulticolsort_amd import datagen


def test_bank_choice():
    # src/stitch_composer.cpp:622-650 and :883-890
    assert [ob.bank_bytes(w) for w in (1, 8, 16, 17, 32, 33, 64)] == [2, 2, 2, 4, 4, 8, 8]
    assert [ob.bank_bytes(w, True) for w in (1, 8, 9, 16)] == [1, 1, 2, 2]


def test_entries_bit_borrow_known_answer():
    # tests/misc_test.cpp:1263-1295: 20b|36b -> 24|32
    ents = ob.construct_entries([20, 36], [0, 1], [24, 32])
    assert ents == [(0, 0, 1, 20), (0, 1, 1, 4), (1, 1, 5, 36)]
    d = ob.stitch([np.array([0x12345], np.uint32), np.array([0x678912345], np.uint64)], [20, 36], ents, [24, 32])
    assert int(d[0][0]) == 0x123456 and int(d[1][0]) == 0x78912345


def test_entries_stitch_layout_known_answer():
    # disabled stitchTest (tests/misc_test.cpp:741-773): old {63,20} -> new {31,21,21,10}
    ents = ob.construct_entries([63, 20], [0, 1], [31, 21, 21, 10])
    assert ents == [(0, 0, 1, 31), (1, 0, 32, 52), (2, 0, 53, 63), (2, 1, 1, 10), (3, 1, 11, 20)]


@pytest.mark.parametrize("bws", [[17, 33], [12, 20, 30], [2, 1, 12, 27], [5, 7, 9, 11, 13]])
def test_stitch_is_a_bijection_of_the_concatenated_key(bws):
    """Every split re-cuts the same bit string: concatenating the new
    columns gives back the concatenation of the old ones."""
    n = 500
    cols = [datagen.uniform_bits(77 + i, n, w).astype(datagen.bank_dtype(w)) for i, w in enumerate(bws)]
    S = sum(bws)
    full = [0] * n
    for c, w in zip(cols, bws):
        full = [(f << w) | int(v) for f, v in zip(full, c)]
    rng = np.random.default_rng(5)
    for _ in range(8):
        r = int(rng.integers(1, min(5, S) + 1))
        cuts = sorted(rng.choice(np.arange(1, S), size=r - 1, replace=False)) if r > 1 else []
        split = list(np.diff([0] + list(cuts) + [S]).astype(int))
        if max(split) > 64:
            continue
        ents = ob.construct_entries(bws, list(range(len(bws))), split)
        new = ob.stitch(cols, bws, ents, split)
        back = [0] * n
        for c, w in zip(new, split):
            back = [(f << int(w)) | int(v) for f, v in zip(back, c)]
        assert back == full


def test_multi_round_sort_equals_lexsort():
    n = 20000
    (c1, c2), _ = datagen.config1(n)
    order, groups, ng, info = ob.multi_round_sort([c1, c2])
    assert np.array_equal(order, np.lexsort((c2, c1)).astype(np.uint32))
    # round-1 groups are the dense ranks of c1 in sorted order
    assert np.array_equal(groups[0], np.unique(c1, return_inverse=True)[1][order].astype(np.uint32))
    assert ng == len(np.unique(c1))
    sizes = np.bincount(groups[0])
    assert info.ngroups == len(sizes) and info.ngroups_gtOne == int((sizes > 1).sum())
    assert info.sum_group_size == n


def test_plans_give_identical_orders():
    """Massaging is a bijection: every two-round plan of the 17|33 sweep
    (src/stitch_composer.cpp:4225-4257) yields the same final order."""
    n = 5000
    (c1, c2), bws = datagen.config1(n)
    base = np.lexsort((c2, c1)).astype(np.uint32)
    for s in range(-33, 18, 5):
        split = [w for w in (17 - s, 33 + s) if w > 0]
        ents = ob.construct_entries(bws, [0, 1], split)
        new = ob.stitch([c1, c2], bws, ents, split)
        order, _, _, _ = ob.multi_round_sort(new)
        assert np.array_equal(order, base), s


def test_extract_group_info_closed_form():
    v = np.sort(datagen.uniform_below(3, 1000, 50)).astype(np.uint64)
    g, ng = ob.extract_group_info(v)
    assert np.array_equal(g, np.r_[0, np.cumsum(np.diff(v) != 0)].astype(np.uint32))
    assert ng == len(np.unique(v))


def test_hash_subgroup_known_answer():
    # tests/subgroup_test.cpp:92-155: two-round hashing partitions rows like
    # one-round hashing of col1*8192+col2; first-appearance one-round ids
    # are [0,1,1,2,3,0,4,5,5]
    c1 = np.array([100, 200, 200, 200, 100, 100, 300, 100, 100], np.uint32)
    c2 = np.array([1000, 2000, 2000, 1000, 2000, 1000, 3000, 4000, 4000], np.uint32)
    comb = (c1.astype(np.uint64) * 8192 + c2).astype(np.uint64)
    _, gid1, _, t1 = ob.multi_round_hash([comb])
    assert list(gid1) == [0, 1, 1, 2, 3, 0, 4, 5, 5] and t1 == 6
    order, gid, rg, total = ob.multi_round_hash([c1, c2])
    assert total == 6
    per_row = np.empty(9, np.uint32)
    per_row[order] = gid
    # same partition as the one-round ids
    pairs = {(int(a), int(b)) for a, b in zip(per_row, gid1)}
    assert len(pairs) == 6
    # ids in (segment, first appearance) order: round-1 groups 100,200,300
    assert list(per_row) == [0, 3, 3, 4, 1, 0, 5, 2, 2]


def _hash_ref_py(cols):
    """Dict-based restatement of multiRoundHashing_reordered for checking."""
    n = len(cols[0])
    order = np.arange(n)
    groups = None
    for r, col in enumerate(cols):
        vals = col[order] if groups is not None else col
        seg = groups if groups is not None else np.zeros(n, np.int64)
        gid = np.empty(n, np.int64)
        nxt = 0
        start = 0
        for e in range(1, n + 1):
            if e == n or seg[e] != seg[start]:
                seen = {}
                for i in range(start, e):
                    v = int(vals[i])
                    if v not in seen:
                        seen[v] = nxt
                        nxt += 1
                    gid[i] = seen[v]
                start = e
        if r == len(cols) - 1:
            return order, gid, nxt
        idx = np.argsort(gid, kind="stable")
        order = order[idx]
        groups = gid[idx]


def test_multi_round_hash_matches_dict_restatement():
    n = 3000
    cols = [datagen.uniform_below(1, n, 7).astype(np.uint16), datagen.uniform_below(2, n, 5).astype(np.uint32),
            datagen.uniform_below(3, n, 11).astype(np.uint64)]
    order, gid, _, total = ob.multi_round_hash(cols)
    eo, eg, et = _hash_ref_py(cols)
    assert total == et and np.array_equal(order, eo) and np.array_equal(gid, eg)


def test_hashmask_closed_form():
    for cnt in [1, 2, 3, 5, 100, 1000, 4096, 5000, 6000, 1 << 20, (1 << 20) + 400000]:
        m = ob.hashmask(cnt)
        assert m & (m - 1) == 0
        lo = 1 << (cnt.bit_length() - 1)
        assert m in (lo, 2 * lo)
        assert (m == 2 * lo) == (lo + lo - cnt < 2 * (cnt - lo))


@pytest.mark.parametrize("S,r", [(42, 1), (42, 2), (42, 3), (42, 4), (50, 2), (10, 5)])
def test_split_enumeration_counts(S, r):
    cnt, plans = ob.split_exhaustive(S, r, 64)
    assert cnt == math.comb(S - 1, r - 1)
    assert all(sum(p) == S and min(p) > 0 for p in plans)
    # descending-first order (src/stitch_composer.cpp:223-228)
    assert [tuple(p) for p in plans] == sorted((tuple(p) for p in plans), reverse=True)


def test_split_enumeration_respects_max():
    cnt, plans = ob.split_exhaustive(100, 2, 64)
    assert cnt == len([a for a in range(1, 100) if a <= 64 and 100 - a <= 64])
    brute = [(a, b) for a in range(64, 0, -1) for b in [100 - a] if 0 < b <= 64]
    assert [tuple(p) for p in plans] == brute


def test_zipf_counts_profile():
    c = datagen.zipf_counts(1 << 20, 1 << 10, 1.1)
    assert c.sum() == 1 << 20 and c.min() >= 1 and (np.diff(c) <= 1).all()


@pytest.mark.skipif(ob.ref() is None, reason="reference library not built")
def test_oracle_matches_reference_rounds():
    n = 1 << 15
    (c1, c2), bws = datagen.config1(n)
    ro, rg = ob.ref_sort_rounds([c1, c2], bws)
    eo, eg, _, _ = ob.multi_round_sort([c1, c2])
    keys = np.stack([c1[ro], c2[ro]])
    assert np.array_equal(ob.canonical_order(keys, ro), eo)
    assert np.array_equal(rg, eg)


@pytest.mark.skipif(ob.ref() is None, reason="reference library not built")
@pytest.mark.parametrize("nthreads", [1, 3, 8])
def test_multithreaded_reference_rounds_match_the_single_threaded(nthreads):
    """The all-core CPU comparator: the reference's kernels under a parallel
    driver give the single-threaded round loop's groups, and its order up to
    the order of equal keys; config-1 and a Zipf first column with groups
    large enough to take the partitioned path."""
    cases = [datagen.config1(1 << 15)]
    z = datagen.zipf_column(11, 1 << 16, 64, 1.1, np.arange(64, dtype=np.uint32) * np.uint32(977))
    cases.append(([z, datagen.uniform_bits(12, 1 << 16, 40)], [16, 40]))
    cases.append(([datagen.uniform_bits(13, 1 << 16, 12).astype(np.uint16),
                   datagen.uniform_bits(14, 1 << 16, 20).astype(np.uint32),
                   datagen.uniform_bits(15, 1 << 16, 30).astype(np.uint32)], [12, 20, 30]))
    for cols, bws in cases:
        cols = [np.ascontiguousarray(c) for c in cols]
        ro, rg = ob.ref_sort_rounds(cols, bws)
        mo, mg = ob.ref_sort_rounds_mt(cols, bws, nthreads)
        keys = np.stack([c[mo].astype(np.uint64) for c in cols])
        keys_r = np.stack([c[ro].astype(np.uint64) for c in cols])
        assert np.array_equal(keys, keys_r)  # the same sorted tuples
        assert np.array_equal(ob.canonical_order(keys, mo), ob.canonical_order(keys_r, ro))
        assert np.array_equal(mg, rg)


@pytest.mark.skipif(ob.ref() is None, reason="reference library not built")
@pytest.mark.parametrize("bws,split", [([20, 36], [24, 32]), ([12, 20, 30], [62]), ([17, 33], [10, 40]),
                                       ([2, 1, 12, 27], [16, 16, 10])])
def test_oracle_stitch_matches_reference(bws, split):
    n = 4096
    cols = [datagen.uniform_bits(9 + i, n, w).astype(datagen.bank_dtype(w)) for i, w in enumerate(bws)]
    ents = ob.construct_entries(bws, list(range(len(bws))), split)
    a = ob.stitch(cols, bws, ents, split)
    b = ob.stitch(cols, bws, ents, split, use_ref=True)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.skipif(ob.ref() is None, reason="reference library not built")
@pytest.mark.parametrize("dt,bits", [(np.uint16, 16), (np.uint32, 32), (np.uint64, 64)])
def test_reference_mergesort_sorts(dt, bits):
    n = 20001
    k = datagen.uniform_bits(4, n, bits).astype(dt)
    o = np.arange(n, dtype=np.uint32)
    rk, ro = ob.ref_mergesort(k, o)
    assert np.array_equal(rk, np.sort(k)) and np.array_equal(k[ro], rk)
