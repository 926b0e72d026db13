"""GPU parity of the persistent size-binned segmented sort (tuning med_max set).

The levels split only segments above med_max rows; one k_seg_bins launch then
sorts every level's windows and single segments, those above one 1024-row
window tile (up to med_max rows) by one workgroup each: its own MSD level,
then its child windows, children still above 1024 rows on a per-block stack.
Windows the bucket sort refuses go to the radix tile sort (and the rank for
keys too wide for a segment index) after the launch.

Bar: bit-exact against np.lexsort of (segment, key, oid) and against the
level-synchronous form of the same call.
"""
import numpy as np
import pytest
import torch

from fastmulticolsort_amd import datagen
from fastmulticolsort_amd.engine import FmcsError, MODE_BATCHED, MODE_DEVICE, MODE_GRAPH

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.cpu().numpy()


def log(*a):
    print(*a, flush=True)


def _sizes(kind, seed):
    rng = np.random.default_rng(seed)
    if kind == "bins":
        # every bin: packed small, single (513..1024), medium (1025..16384), huge
        parts = [rng.integers(0, 513, 400), rng.integers(513, 1025, 60), rng.integers(1025, 16385, 40),
                 np.array([20000, 70001, 150000])]
        sizes = np.concatenate(parts)
        rng.shuffle(sizes)
    elif kind == "medium":
        sizes = rng.integers(1025, 16385, 40)
    elif kind == "zipf":
        sizes = datagen.zipf_counts(300000, 8000, 1.1)
    else:  # edges: bin boundaries exactly
        sizes = np.array([0, 1, 512, 513, 1024, 1025, 2048, 2049, 4096, 16384, 16385, 0, 3, 1023, 8191, 1])
        sizes = np.tile(sizes, 3)
    return sizes.astype(np.int64)


def _keys(kind, n, dt, bits, seed):
    k = datagen.uniform_bits(seed, n, bits).astype(dt)
    if kind == "dups":       # heavy duplicates: bucket bins overflow -> radix tile sort
        k = (k % dt(3)).astype(dt)
    elif kind == "lowbits":  # only the low 10 bits vary: a medium's top digits are constant
        k = (k & dt((1 << min(10, bits)) - 1)).astype(dt)
    elif kind == "equal":    # one value: every level keeps the whole segment
        k[:] = k[0]
    else:
        k[::7] = k[0]
    return k


def _run(ctx, k, oids, off, bits, mode, med_max):
    old = ctx.tuning()
    try:
        ctx.set_tuning(min(old["small_max"], 512), min(old["window"], 512), old["msd_bits"], med_max=med_max)
        kd, od = dev(k), dev(oids)
        info = ctx.segmented_sort(kd, od, dev(off), len(off) - 1, bits, mode=mode, want_info=True)
        ctx.sync()  # reports a key above the declared width (the device-driven forms' error word)
        return host(kd), host(od), info
    finally:
        ctx.set_tuning(old["small_max"], old["window"], old["msd_bits"], old["gather_window_kb"], old["carry_next"],
                       old["med_max"])


CASES = [("bins", "uniform"), ("bins", "dups"), ("medium", "lowbits"), ("medium", "equal"), ("zipf", "uniform"),
         ("edges", "uniform"), ("edges", "dups")]


@pytest.mark.parametrize("med_max", [1024, 4096, 16384])
@pytest.mark.parametrize("dt,bits", [(np.uint64, 40), (np.uint32, 20), (np.uint16, 9), (np.uint64, 60),
                                     (np.uint64, 64)])
@pytest.mark.parametrize("seg,kk", CASES)
def test_persistent_segmented_sort(ctx, med_max, dt, bits, seg, kk):
    sizes = _sizes(seg, 11)
    n = int(sizes.sum())
    off = np.r_[0, np.cumsum(sizes)].astype(np.uint32)
    nseg = len(sizes)
    log("persistent segsort", med_max, np.dtype(dt).name, bits, seg, kk, n, nseg)
    k = _keys(kk, n, dt, bits, 31)
    oids = datagen.permutation(9, n).astype(np.uint32)
    for s in range(nseg):  # within a segment incoming rows are in ascending oid order
        oids[off[s]:off[s + 1]] = np.sort(oids[off[s]:off[s + 1]])
    gk, go, info = _run(ctx, k, oids, off, bits, MODE_DEVICE, med_max)
    seg_id = np.repeat(np.arange(nseg), sizes)
    idx = np.lexsort((oids, k, seg_id))
    assert np.array_equal(go, oids[idx])
    assert np.array_equal(gk, k[idx])
    assert info.ngroups == nseg and info.ngroups_gtOne == int((sizes > 1).sum())
    # the level-synchronous form gives the same rows
    lk, lo_, _ = _run(ctx, k, oids, off, bits, MODE_DEVICE, 0)
    assert np.array_equal(lo_, go) and np.array_equal(lk, gk)


def test_persistent_reports_wide_keys(ctx):
    # a key above the declared width inside a medium segment (level 0)
    sizes = np.array([300, 5000, 700, 20000])
    n = int(sizes.sum())
    off = np.r_[0, np.cumsum(sizes)].astype(np.uint32)
    k = datagen.uniform_bits(3, n, 30).astype(np.uint64)
    oids = np.arange(n, dtype=np.uint32)
    for at in (310, 20, 5400, 9000):  # medium, window, single, huge
        kb = k.copy()
        kb[at] = np.uint64(1) << np.uint64(35)
        log("persistent wide key at", at)
        with pytest.raises(FmcsError):
            _run(ctx, kb, oids, off, 30, MODE_DEVICE, 8192)
    # the error is reported once; a clean call afterwards succeeds
    gk, go, _ = _run(ctx, k, oids, off, 30, MODE_DEVICE, 8192)
    seg_id = np.repeat(np.arange(len(sizes)), sizes)
    assert np.array_equal(go, oids[np.lexsort((oids, k, seg_id))])


def test_persistent_multi_round_graph(ctx):
    n = 1 << 20
    (c1, c2), bws = datagen.config3(n, ngroups=1 << 12)
    old = ctx.tuning()
    try:
        ctx.set_tuning(old["small_max"], old["window"], old["msd_bits"], med_max=8192)
        d1, d2 = dev(c1), dev(c2)
        order = torch.empty(n, dtype=torch.uint32, device="cuda")
        exp = np.lexsort((c2, c1)).astype(np.uint32)
        for rep in range(4):
            log("persistent graph replay", rep)
            ctx.multi_round_sort([d1, d2], bws, order=order, mode=MODE_GRAPH, want_count=False)
            ctx.sync()
            assert np.array_equal(host(order), exp)
            # eager work between replays
            ctx.multi_round_sort([d1, d2], bws, mode=MODE_BATCHED)
        order2, ng = ctx.multi_round_sort([d1, d2], bws, mode=MODE_DEVICE)
        assert np.array_equal(host(order2), exp)
        assert ng == len(np.unique(c1))  # the count of the round before the last (the last extracts no groups)
    finally:
        ctx.set_tuning(old["small_max"], old["window"], old["msd_bits"], old["gather_window_kb"], old["carry_next"],
                       old["med_max"])


def test_probe_handoff_runs(ctx):
    ms = ctx.probe_handoff(1024, 256)
    log("handoff probe", ms)
    assert len(ms) == 4 and all(m > 0 for m in ms)
    with pytest.raises(FmcsError):
        ctx.probe_handoff(0, 256)
