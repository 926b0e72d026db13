"""Buffer arguments at the C ABI: what the reference would pass.

The reference's buffers are pageable host memory (malloc_aligned,
src/common.h:19-31).  Without staging every entry point must refuse them
with FMCS_E_ARG before it launches anything; with staging
(fmcs_ctx_set_host_staging) the same calls on host arrays must give the
results of the device calls, bit for bit.  Pinned host memory is used in
place.  Every host-pointer call below is made only after fmcs_pointer_kind
has classified that pointer as host memory.

Also here: keys wider than their declared width in every segmented-sort
mode, and the device-driven error word surviving later calls until a sync
reports it.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from fastmulticolsort_amd import _lib, datagen
from fastmulticolsort_amd._lib import (FMCS_E_ARG, MEM_DEVICE, MEM_HOST, MEM_PINNED, Entry, FmcsError)
from fastmulticolsort_amd.engine import (MODE_BATCHED, MODE_DEVICE, MODE_GRAPH, MODE_PLAIN, MODE_UNBINNED,
                                         construct_entries)

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)


def hp(a):
    """Address of a numpy array (pageable host memory)."""
    return C.c_void_p(a.ctypes.data)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def lex(cols):
    return np.lexsort(tuple(reversed(cols))).astype(np.uint32)


@pytest.fixture
def staging(ctx):
    ctx.set_host_staging(True)
    yield ctx
    ctx.set_host_staging(False)


def err(ctx):
    return ctx.lib.fmcs_last_error(ctx.h).decode()


def test_pointer_kinds(ctx):
    d = torch.zeros(1000, dtype=torch.uint32, device="cuda")
    assert ctx.pointer_kind(d.data_ptr()) == MEM_DEVICE
    assert ctx.pointer_kind(d.data_ptr() + 400) == MEM_DEVICE  # inside an allocation
    h = np.zeros(1000, np.uint32)
    assert ctx.pointer_kind(h.ctypes.data) == MEM_HOST
    p = torch.zeros(1000, dtype=torch.uint32).pin_memory()
    assert ctx.pointer_kind(p.data_ptr()) == MEM_PINNED


def test_host_buffers_refused_without_staging(ctx):
    lib = ctx.lib
    ctx.set_host_staging(False)
    n = 5000
    vals = np.arange(n, dtype=np.uint32)[::-1].copy()
    order = np.arange(n, dtype=np.uint32)
    out = np.full(n, 7, np.uint32)
    for a in (vals, order, out):
        assert ctx.pointer_kind(a.ctypes.data) == MEM_HOST
    log("gather, all host")
    assert lib.fmcs_gather(ctx.h, 4, n, hp(vals), hp(order), hp(out)) == FMCS_E_ARG
    assert "pageable host memory" in err(ctx) and "values_d" in err(ctx)
    log("gather, host output only")
    dv, do = dev(vals), dev(order)
    assert lib.fmcs_gather(ctx.h, 4, n, C.c_void_p(dv.data_ptr()), C.c_void_p(do.data_ptr()), hp(out)) == FMCS_E_ARG
    assert "out_d" in err(ctx)
    assert (out == 7).all()
    log("sort_by_group, host column")
    og = np.zeros(n, np.uint32)
    ng = C.c_uint32(0)
    assert lib.fmcs_sort_by_group(ctx.h, 4, 13, n, hp(vals), C.c_void_p(do.data_ptr()), None, hp(og), 0,
                                  MODE_BATCHED, None, None, C.byref(ng)) == FMCS_E_ARG
    log("multi_round_sort, host columns")
    cols = (C.c_void_p * 1)(vals.ctypes.data)
    banks = (C.c_uint32 * 1)(4)
    bws = (C.c_uint32 * 1)(13)
    for mode in (MODE_BATCHED, MODE_DEVICE, MODE_GRAPH):
        assert lib.fmcs_multi_round_sort(ctx.h, 1, cols, banks, bws, n, C.c_void_p(do.data_ptr()), None, mode,
                                         None) == FMCS_E_ARG
    assert np.array_equal(do.cpu().numpy(), order)  # nothing ran


def test_noalloc_never_stages(staging):
    ctx = staging
    n = 1000
    k = np.arange(n, dtype=np.uint32)[::-1].copy()
    o = np.arange(n, dtype=np.uint32)
    assert ctx.pointer_kind(k.ctypes.data) == MEM_HOST
    assert ctx.lib.fmcs_sort_pairs_noalloc(ctx.h, 4, 10, n, hp(k), hp(o)) == FMCS_E_ARG
    assert "not with this entry point" in err(ctx)


def test_staged_sort_by_group_rounds_like_the_reference_binding(staging):
    # the reference's round loop (src/composer.cpp:48-135) over host arrays:
    # round 1 sorts c1 from the identity order, round 2 sorts c2 within c1's groups
    ctx = staging
    lib = ctx.lib
    n = 200003
    (c1, c2), bws = datagen.config1(n)
    for mode in (MODE_BATCHED, MODE_DEVICE):
        log("staged sortByGroup rounds, mode", mode)
        order = np.arange(n, dtype=np.uint32)
        g1 = np.zeros(n, np.uint32)
        ng = C.c_uint32(0)
        for a in (c1, c2, order, g1):
            assert ctx.pointer_kind(a.ctypes.data) == MEM_HOST
        assert lib.fmcs_sort_by_group(ctx.h, 4, bws[0], n, hp(c1), hp(order), None, hp(g1), 0, mode, None, None,
                                      C.byref(ng)) == 0, err(ctx)
        exp1 = np.argsort(c1, kind="stable").astype(np.uint32)
        assert np.array_equal(order, exp1)
        _, first = np.unique(c1[exp1], return_index=True)
        assert ng.value == len(first)
        assert lib.fmcs_sort_by_group(ctx.h, 8, bws[1], n, hp(c2), hp(order), hp(g1), None, 1, mode, None, None,
                                      C.byref(ng)) == 0, err(ctx)
        ctx.sync()
        assert np.array_equal(order, lex([c1, c2]))


@pytest.mark.parametrize("mode", [MODE_BATCHED, MODE_DEVICE, MODE_GRAPH])
def test_staged_multi_round_sort(staging, mode):
    ctx = staging
    lib = ctx.lib
    n = 300007
    (c1, c2), bws = datagen.config1(n)
    cols = (C.c_void_p * 2)(c1.ctypes.data, c2.ctypes.data)
    banks = (C.c_uint32 * 2)(4, 8)
    bw = (C.c_uint32 * 2)(*bws)
    for rep in range(2):  # the graph mode replays the second time
        log("staged multi_round_sort mode", mode, "call", rep)
        order = np.zeros(n, np.uint32)
        groups = np.zeros(n, np.uint32)
        ng = C.c_uint32(0)
        assert lib.fmcs_multi_round_sort(ctx.h, 2, cols, banks, bw, n, hp(order), hp(groups), mode,
                                         C.byref(ng)) == 0, err(ctx)
        exp = lex([c1, c2])
        assert np.array_equal(order, exp)
        t1, t2 = c1[exp], c2[exp]
        heads = np.r_[True, (t1[1:] != t1[:-1]) | (t2[1:] != t2[:-1])]
        assert np.array_equal(groups, (np.cumsum(heads) - 1).astype(np.uint32))
        assert ng.value == int(heads.sum())


def test_staged_kernel_seams(staging):
    ctx = staging
    lib = ctx.lib
    n = 100001
    # stitch 17|33 -> 20|30 from host sources into host destinations
    (c1, c2), bws = datagen.config1(n)
    split = [20, 30]
    ents = construct_entries(bws, [0, 1], split)
    d1 = np.zeros(n, np.uint32)
    d2 = np.zeros(n, np.uint32)
    sp = (C.c_void_p * 2)(c1.ctypes.data, c2.ctypes.data)
    sb = (C.c_uint32 * 2)(4, 8)
    sw = (C.c_uint32 * 2)(*bws)
    ea = (Entry * len(ents))(*[Entry(*e) for e in ents])
    dp = (C.c_void_p * 2)(d1.ctypes.data, d2.ctypes.data)
    db = (C.c_uint32 * 2)(4, 4)
    log("staged stitch")
    assert lib.fmcs_stitch(ctx.h, n, 2, sp, sb, sw, len(ents), ea, 2, dp, db) == 0, err(ctx)
    full = (c1.astype(np.uint64) << np.uint64(33)) | c2
    assert np.array_equal(d1, (full >> np.uint64(30)).astype(np.uint32))
    assert np.array_equal(d2, (full & np.uint64((1 << 30) - 1)).astype(np.uint32))
    # gather
    log("staged gather")
    perm = datagen.permutation(3, n).astype(np.uint32)
    out = np.zeros(n, np.uint64)
    assert lib.fmcs_gather(ctx.h, 8, n, hp(c2), hp(perm), hp(out)) == 0, err(ctx)
    assert np.array_equal(out, c2[perm])
    # sort pairs, then group boundaries with an offsets array of exactly ngroups + 1 words
    log("staged sort_pairs + extract_group_info")
    k = c1.copy()
    o = np.arange(n, dtype=np.uint32)
    assert lib.fmcs_sort_pairs(ctx.h, 4, 17, n, hp(k), hp(o)) == 0, err(ctx)
    exp = np.argsort(c1, kind="stable")
    assert np.array_equal(o, exp) and np.array_equal(k, c1[exp])
    heads = np.r_[True, k[1:] != k[:-1]]
    ngx = int(heads.sum())
    off = np.full(ngx + 1 + 16, 0xDEADBEEF, np.uint32)  # 16 guard words after the array's ngroups + 1
    gid = np.zeros(n, np.uint32)
    ng = C.c_uint32(0)
    assert lib.fmcs_extract_group_info(ctx.h, 4, n, hp(k), None, hp(gid), hp(off), C.byref(ng)) == 0, err(ctx)
    assert ng.value == ngx
    assert np.array_equal(off[: ngx + 1], np.r_[np.flatnonzero(heads), n].astype(np.uint32))
    assert (off[ngx + 1:] == 0xDEADBEEF).all()
    assert np.array_equal(gid, (np.cumsum(heads) - 1).astype(np.uint32))
    # segmented sort of c2 within those groups, in place on host arrays
    log("staged segmented_sort")
    keys = c2[exp].copy()
    oids = o.copy()
    offs = off[: ngx + 1].copy()
    assert lib.fmcs_segmented_sort(ctx.h, 8, 33, n, hp(keys), hp(oids), hp(offs), ngx, MODE_BATCHED, None) == 0
    assert np.array_equal(oids, lex([c1, c2]))
    # hash grouping
    log("staged multi_round_hash")
    hcols = (C.c_void_p * 2)(c1.ctypes.data, c2.ctypes.data)
    ho = np.zeros(n, np.uint32)
    hg = np.zeros(n, np.uint32)
    assert lib.fmcs_multi_round_hash(ctx.h, 2, hcols, sb, n, 1, hp(ho), hp(hg), C.byref(ng)) == 0, err(ctx)
    assert ng.value == len(np.unique(full))


def test_staged_run_plan(staging):
    ctx = staging
    from fastmulticolsort_amd import planner
    lib = planner.lib()
    n = 65536 + 77
    (c1, c2), bws = datagen.config1(n)
    cols = (C.c_void_p * 2)(c1.ctypes.data, c2.ctypes.data)
    banks = (C.c_uint32 * 2)(4, 8)
    bw = (C.c_uint32 * 2)(*bws)
    order = (C.c_uint32 * 2)(0, 1)
    split = (C.c_uint32 * 2)(24, 26)
    out = np.zeros(n, np.uint32)
    ms = (C.c_double * 6)()
    tot = C.c_double(0)
    ng = C.c_uint32(0)
    log("staged fmcs_run_plan")
    rc = lib.fmcs_run_plan(ctx.h, 2, cols, banks, bw, order, 2, split, MODE_DEVICE, n, hp(out), None, ms,
                           C.byref(tot), C.byref(ng))
    assert rc == 0, err(ctx)
    assert np.array_equal(out, lex([c1, c2]))


def test_pinned_buffers_used_in_place(ctx):
    ctx.set_host_staging(False)
    n = 50000
    vals = torch.from_numpy(datagen.uniform_bits(4, n, 40).astype(np.uint64)).pin_memory()
    perm = torch.from_numpy(datagen.permutation(5, n).astype(np.uint32)).pin_memory()
    assert ctx.pointer_kind(vals.data_ptr()) == MEM_PINNED
    out = ctx.gather(vals, perm, out=torch.empty(n, dtype=torch.uint64, device="cuda"))
    assert np.array_equal(out.cpu().numpy(), vals.numpy()[perm.numpy()])


# ---- keys wider than declared ------------------------------------------------

def _segments(n):
    sizes = datagen.zipf_counts(n, 300, 1.1)  # small windows, single windows and large segments
    return np.r_[0, np.cumsum(sizes)].astype(np.uint32)


@pytest.mark.parametrize("where", ["small", "large"])
@pytest.mark.parametrize("mode", [MODE_BATCHED, MODE_PLAIN, MODE_UNBINNED, MODE_DEVICE])
def test_segmented_sort_reports_keys_wider_than_declared(ctx, mode, where):
    n = 60000
    off = _segments(n)
    sizes = np.diff(off)
    seg = int(np.argmin(sizes)) if where == "small" else int(np.argmax(sizes))
    log("oversized key, mode", mode, where, "segment of", sizes[seg], "rows")
    keys = datagen.uniform_bits(6, n, 30).astype(np.uint64)
    keys[off[seg]] |= np.uint64(1 << 30)
    kd, od = dev(keys), dev(np.arange(n, dtype=np.uint32))
    if mode == MODE_DEVICE:
        ctx.segmented_sort(kd, od, dev(off), len(sizes), 30, mode=mode)  # no synchronisation: no error yet
        with pytest.raises(FmcsError):
            ctx.sync()
        ctx.sync()  # reported once
    else:
        with pytest.raises(FmcsError):
            ctx.segmented_sort(kd, od, dev(off), len(sizes), 30, mode=mode)
    # the same data within the width sorts
    keys[off[seg]] &= np.uint64((1 << 30) - 1)
    kd, od = dev(keys), dev(np.arange(n, dtype=np.uint32))
    ctx.segmented_sort(kd, od, dev(off), len(sizes), 30, mode=mode)
    ctx.sync()
    gid = np.repeat(np.arange(len(sizes)), sizes)
    assert np.array_equal(od.cpu().numpy(), np.lexsort((keys, gid)).astype(np.uint32))


def test_device_key_error_survives_a_later_call(ctx):
    # an unsynchronised device-driven call with a key above its width, then a
    # good call: the next sync still reports the first call's error
    n = 20000
    c1 = datagen.uniform_bits(8, n, 24).astype(np.uint32)  # declared 20 bits
    c2 = datagen.uniform_bits(9, n, 40).astype(np.uint64)
    good = c1 & np.uint32((1 << 20) - 1)
    order = torch.empty(n, dtype=torch.uint32, device="cuda")
    ctx.multi_round_sort([dev(c1), dev(c2)], [20, 40], order=order, mode=MODE_DEVICE, want_count=False)
    ctx.multi_round_sort([dev(good), dev(c2)], [20, 40], order=order, mode=MODE_DEVICE, want_count=False)
    with pytest.raises(FmcsError):
        ctx.sync()
    ctx.sync()
    assert np.array_equal(order.cpu().numpy(), lex([good, c2]))


@pytest.mark.parametrize("carry_next", [0, 2])
def test_device_round2_reports_keys_wider_than_declared(ctx, carry_next):
    # a gathered or carried (not packed) second column is checked by the
    # segmented sort's first level
    n = 30000
    c1 = datagen.uniform_bits(10, n, 12).astype(np.uint32)
    c2 = datagen.uniform_bits(11, n, 40).astype(np.uint64)
    c2[n // 3] |= np.uint64(1 << 45)
    t = ctx.tuning()
    try:
        ctx.set_tuning(t["small_max"], t["window"], t["msd_bits"], carry_next=carry_next)
        order = torch.empty(n, dtype=torch.uint32, device="cuda")
        with pytest.raises(FmcsError):
            ctx.multi_round_sort([dev(c1), dev(c2)], [12, 40], order=order, mode=MODE_DEVICE)
        c2[n // 3] &= np.uint64((1 << 40) - 1)
        ctx.multi_round_sort([dev(c1), dev(c2)], [12, 40], order=order, mode=MODE_DEVICE)
        assert np.array_equal(order.cpu().numpy(), lex([c1, c2]))
    finally:
        ctx.set_tuning(t["small_max"], t["window"], t["msd_bits"], carry_next=t["carry_next"])
