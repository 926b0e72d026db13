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
