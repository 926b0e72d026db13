"""GPU hash grouping (the GROUP BY alternative) against the oracle."""
import numpy as np
import pytest
import torch

import oracle_bind as ob
from fastmulticolsort_amd import datagen

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.cpu().numpy()


def test_hash_subgroup_known_answer(ctx):
    # tests/subgroup_test.cpp:92-155 (two rounds), expected ids derived from
    # first appearance within the round-1 groups 100, 200, 300
    c1 = np.array([100, 200, 200, 200, 100, 100, 300, 100, 100], np.uint32)
    c2 = np.array([1000, 2000, 2000, 1000, 2000, 1000, 3000, 4000, 4000], np.uint32)
    order, gid, ng = ctx.multi_round_hash([dev(c1), dev(c2)])
    per_row = np.empty(9, np.uint32)
    per_row[host(order)] = host(gid)
    assert ng == 6 and list(per_row) == [0, 3, 3, 4, 1, 0, 5, 2, 2]


COLSETS = [
    ("u8-u16", lambda n: [datagen.uniform_below(1, n, 3).astype(np.uint8), datagen.uniform_below(2, n, 40).astype(np.uint16)]),
    ("u32-u64", lambda n: [datagen.uniform_below(3, n, 1000).astype(np.uint32), datagen.uniform_below(4, n, 7).astype(np.uint64)]),
    ("three", lambda n: [datagen.uniform_below(5, n, 5).astype(np.uint16), datagen.uniform_below(6, n, 9).astype(np.uint32),
                         datagen.uniform_bits(7, n, 40).astype(np.uint64)]),
    ("one", lambda n: [datagen.uniform_below(8, n, 12345).astype(np.uint32)]),
]


@pytest.mark.parametrize("name,make", COLSETS)
@pytest.mark.parametrize("n", [1, 1000, 200001])
def test_multi_round_hash_matches_oracle(ctx, name, make, n):
    print("hash", name, n, flush=True)
    cols = make(n)
    order, gid, ng = ctx.multi_round_hash([dev(c) for c in cols])
    eo, eg, _, et = ob.multi_round_hash(cols)
    assert ng == et
    assert np.array_equal(host(order), eo)
    assert np.array_equal(host(gid), eg)


@pytest.mark.parametrize("n", [1000, 150001])
def test_non_reordered_hash_first_appearance(ctx, n):
    cols = [datagen.uniform_below(11, n, 6).astype(np.uint16), datagen.uniform_below(12, n, 50).astype(np.uint32)]
    order, gid, ng = ctx.multi_round_hash([dev(c) for c in cols], hash_type=0)
    tup = cols[0].astype(np.uint64) * 1000 + cols[1]
    _, first_idx, inv = np.unique(tup, return_index=True, return_inverse=True)
    rank = np.argsort(np.argsort(first_idx))
    assert ng == len(first_idx)
    assert np.array_equal(host(gid), rank[inv].astype(np.uint32))
    assert np.array_equal(host(order), np.arange(n, dtype=np.uint32))


def test_config4_group_count_exact(ctx):
    n = 1 << 20
    cols, bws = datagen.config4(n)
    order, gid, ng = ctx.multi_round_hash([dev(c) for c in cols])
    stacked = np.stack([c.astype(np.uint64) for c in cols], axis=1)
    assert ng == len(np.unique(stacked, axis=0))
    # the hash partition equals the sort partition: rows with equal tuples
    # share an id and ids are dense
    o = host(order)
    g = host(gid)
    per_row = np.empty(n, np.uint32)
    per_row[o] = g
    key = ((stacked[:, 0] << 40) | (stacked[:, 1] << 39) | (stacked[:, 2] << 27) | stacked[:, 3])
    u, inv = np.unique(key, return_inverse=True)
    pairs = np.unique(np.stack([inv, per_row]), axis=1)
    assert pairs.shape[1] == len(u) == ng


# ---- the packed path: composites of at most 64 bits -------------------------

M64 = (1 << 64) - 1


def _inv_xorshift(y, s):
    x = y
    for _ in range(64 // s + 1):
        x = y ^ (x >> s)
    return x & M64


def unmix(h):
    """Inverse of the bijective mixer the packed hash path applies to a
    composite (SplitMix64's finalizer)."""
    x = _inv_xorshift(h, 31)
    x = (x * pow(0x94D049BB133111EB, -1, 1 << 64)) & M64
    x = _inv_xorshift(x, 27)
    x = (x * pow(0xBF58476D1CE4E5B9, -1, 1 << 64)) & M64
    return _inv_xorshift(x, 30)


def first_appearance_ids(keys, seg=None):
    """Dense ids of (seg, key) in order of first appearance (restated with numpy)."""
    tup = keys.astype(np.uint64) if seg is None else np.stack([seg.astype(np.uint64), keys.astype(np.uint64)], 1)
    _, first, inv = np.unique(tup, return_index=True, return_inverse=True, axis=0 if seg is not None else None)
    rank = np.empty(len(first), np.int64)
    rank[np.argsort(first, kind="stable")] = np.arange(len(first))
    return rank[inv.ravel()].astype(np.uint32), len(first)


@pytest.mark.parametrize("n", [1, 2, 255, 256, 4097, 1 << 17, (1 << 17) + 1, 600001])
def test_distinct_count_and_ids_across_partition_widths(ctx, n):
    # 8-bit partitions up to 2^17 rows, 16-bit ones above; ids from the one-round hash
    print("packed hash, distinct", n, flush=True)
    k = datagen.uniform_below(41, n, max(2, n // 3)).astype(np.uint32)
    assert ctx.distinct_count(dev(k)) == len(np.unique(k))
    gid, _, _, ng = ctx.hash_round(dev(k), None, last_round=True)
    eg, en = first_appearance_ids(k)
    assert ng == en and np.array_equal(host(gid), eg)


@pytest.mark.parametrize("dt", [np.uint8, np.uint16, np.uint32])
def test_heavy_hitters_take_the_global_table(ctx, dt):
    # three values over 2^20 rows fill a few partitions far beyond an LDS
    # window; one rare value and a million distinct ones beside them in round 2
    n = 1 << 20
    print("packed hash, heavy hitters", np.dtype(dt).name, flush=True)
    c1 = datagen.uniform_below(43, n, 3).astype(dt)
    c2 = datagen.uniform_below(44, n, 1 << 31).astype(np.uint32)
    c2[: n // 2] = 7  # half the rows one value
    order, gid, ng = ctx.multi_round_hash([dev(c1), dev(c2)])
    eo, eg, _, et = ob.multi_round_hash([c1, c2])
    assert ng == et
    assert np.array_equal(host(order), eo) and np.array_equal(host(gid), eg)
    assert ctx.distinct_count(dev(c1)) == 3


def test_the_composite_whose_hash_is_the_empty_marker(ctx):
    # the one composite the mixer maps to all ones, as a u64 key and as
    # (segment, u32 key); in a small window and in a large one
    c0 = unmix(M64)
    for n, reps in ((5000, 7), (1 << 18, 1 << 17)):
        print("packed hash, marker composite", n, reps, flush=True)
        k = datagen.uniform_bits(45, n, 64).astype(np.uint64)
        idx = datagen.permutation(46, n)[:reps]
        k[idx] = np.uint64(c0)
        gid, _, _, ng = ctx.hash_round(dev(k), None, last_round=True)
        eg, en = first_appearance_ids(k)
        assert ng == en and np.array_equal(host(gid), eg)
        seg = datagen.uniform_below(47, n, 5).astype(np.uint32)
        key = datagen.uniform_below(48, n, 1000).astype(np.uint32)
        seg[idx] = np.uint32(c0 >> 32)
        key[idx] = np.uint32(c0 & 0xFFFFFFFF)
        order = torch.arange(n, dtype=torch.int64).to(torch.int32).view(torch.uint32).cuda()
        gid, _, _, ng = ctx.hash_round(dev(key), order, in_group=dev(seg), last_round=True)
        eg, en = first_appearance_ids(key, seg)
        assert ng == en and np.array_equal(host(gid), eg)


def test_local_form_table_overflow_runs_again(ctx):
    # the form is chosen from every stride-th 2048-row chunk (about 256 of
    # them): make those chunks one repeated value and the others distinct, so
    # the sample calls for a small table and its first pass overflows it
    n = 1 << 22
    nch = n // 2048
    stride = max(1, nch // 256)
    k = (np.arange(n, dtype=np.uint64) * np.uint64(2654435761) + np.uint64(12345)).astype(np.uint32)
    for c in range(0, nch, stride):
        k[c * 2048:(c + 1) * 2048] = 7
    print("packed hash, local-form overflow", n, stride, flush=True)
    gid, _, _, ng = ctx.hash_round(dev(k), None, last_round=True)
    eg, en = first_appearance_ids(k)
    assert ng == en and np.array_equal(host(gid), eg)


@pytest.mark.parametrize("n", [(1 << 21) + 5, 1 << 22])
def test_long_scans_number_the_groups_in_both_forms(ctx, n):
    # from 2^21 rows the heads are numbered by reduce-then-scan: the packed
    # form scans its head flags (F[p] == p) directly, the chained form (u64
    # keys under segments) scans a flag array; a ragged tail ends the last chunk
    print("long scans", n, flush=True)
    k32 = datagen.uniform_below(51, n, 1 << 20).astype(np.uint32)
    gid, _, _, ng = ctx.hash_round(dev(k32), None, last_round=True)
    eg, en = first_appearance_ids(k32)
    assert ng == en and np.array_equal(host(gid), eg)
    seg = datagen.uniform_below(52, n, 3).astype(np.uint32)
    k64 = datagen.uniform_below(53, n, 1 << 18).astype(np.uint64) << np.uint64(40)
    order = torch.arange(n, dtype=torch.int64).to(torch.int32).view(torch.uint32).cuda()
    gid, _, _, ng = ctx.hash_round(dev(k64), order, in_group=dev(seg), last_round=True)
    eg, en = first_appearance_ids(k64, seg)
    assert ng == en and np.array_equal(host(gid), eg)
