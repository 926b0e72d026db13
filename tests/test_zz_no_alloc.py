"""fmcs_sort_pairs_noalloc against fmcs_sort_pairs, through the C ABI.

Bar: bit-exact keys and oids (both are the same stable LSD radix sort), on
two calls in a row, the second with a smaller input, and with no device
memory allocated by the calls.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from fastmulticolsort_amd import _lib, datagen

pytestmark = pytest.mark.gpu

MAX_ROWS = 1 << 20


def log(*a):
    print(*a, flush=True)


def sort_pairs(ctx, keys, oids, bits, noalloc):
    fn = ctx.lib.fmcs_sort_pairs_noalloc if noalloc else ctx.lib.fmcs_sort_pairs
    return fn(ctx.h, keys.element_size(), bits, keys.numel(), C.c_void_p(keys.data_ptr()),
              C.c_void_p(oids.data_ptr()))


def keys_for(seed, n, dt, bits, kind):
    k = datagen.uniform_bits(seed, n, bits).astype(dt)
    if kind == "ties" and n > 10:
        k[: n // 3] = k[0]
    elif kind == "low8":  # one pass with more than one digit: an odd pass count
        k &= dt(0xFF)
    elif kind == "low16":  # two such passes
        k &= dt(0xFFFF)
    return k


def check_pair(ctx, k, oids, bits):
    """noalloc result == fmcs_sort_pairs result == numpy's stable argsort."""
    ka, oa = torch.from_numpy(k).cuda(), torch.from_numpy(oids).cuda()
    kb, ob = ka.clone(), oa.clone()
    assert sort_pairs(ctx, ka, oa, bits, True) == 0, ctx.lib.fmcs_last_error(ctx.h)
    assert sort_pairs(ctx, kb, ob, bits, False) == 0, ctx.lib.fmcs_last_error(ctx.h)
    torch.cuda.synchronize()
    got_k, got_o = ka.cpu().numpy(), oa.cpu().numpy()
    assert np.array_equal(got_k, kb.cpu().numpy())
    assert np.array_equal(got_o, ob.cpu().numpy())
    exp = np.argsort(k, kind="stable")
    assert np.array_equal(got_o, oids[exp])
    assert np.array_equal(got_k, k[exp])


CASES = [(np.uint16, 12, "ties"), (np.uint32, 32, "ties"), (np.uint32, 20, "uniform"), (np.uint64, 40, "ties"),
         (np.uint64, 64, "uniform"), (np.uint64, 33, "low8"), (np.uint32, 32, "low16")]


@pytest.mark.parametrize("dt,bits,kind", CASES)
def test_noalloc_equals_sort_pairs_two_calls(ctx, dt, bits, kind):
    # first call at the largest size, the second smaller, ragged, with
    # incoming oids not ascending
    for call, n in enumerate([MAX_ROWS, 123457]):
        log("noalloc", np.dtype(dt).name, bits, kind, "call", call, "n", n)
        k = keys_for(31 + call, n, dt, bits, kind)
        oids = datagen.permutation(41 + call, n).astype(np.uint32)
        check_pair(ctx, k, oids, bits)


def test_noalloc_between_allocating_calls(ctx):
    # the arena is set up per call: an allocating sort between two
    # allocation-free ones leaves both correct
    for i, n in enumerate([5000, 70001, 4096]):
        log("noalloc interleaved", n)
        k = keys_for(51 + i, n, np.uint64, 44, "ties")
        check_pair(ctx, k, np.arange(n, dtype=np.uint32), 44)


def test_noalloc_small_and_rejected_sizes(ctx):
    lib = ctx.lib
    for n in (1, 2, 3, 4095, 4097):
        log("noalloc small", n)
        k = keys_for(61, n, np.uint32, 30, "uniform")
        check_pair(ctx, k, np.arange(n, dtype=np.uint32), 30)
    empty_k = torch.empty(0, dtype=torch.uint32, device="cuda")
    empty_o = torch.empty(0, dtype=torch.uint32, device="cuda")
    assert sort_pairs(ctx, empty_k, empty_o, 30, True) == 0
    # one row over the limit: refused before any launch
    big_k = torch.zeros(MAX_ROWS + 1, dtype=torch.uint32, device="cuda")
    big_o = torch.zeros(MAX_ROWS + 1, dtype=torch.uint32, device="cuda")
    assert sort_pairs(ctx, big_k, big_o, 30, True) == _lib.FMCS_E_ARG
    assert b"FMCS_NOALLOC_MAX_ROWS" in lib.fmcs_last_error(ctx.h)
    assert sort_pairs(ctx, big_k[:8], big_o[:8], 33, True) == _lib.FMCS_E_ARG  # keybits above a u32 bank


def test_noalloc_allocates_no_device_memory(ctx):
    n = MAX_ROWS
    k = torch.from_numpy(keys_for(71, n, np.uint64, 64, "uniform")).cuda()
    o = torch.from_numpy(np.arange(n, dtype=np.uint32)).cuda()
    assert sort_pairs(ctx, k, o, 64, True) == 0  # first call on this context's device
    ctx.trim()  # drop the context's workspace: an allocating path would have to grow it again
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for _ in range(3):
        assert sort_pairs(ctx, k, o, 64, True) == 0
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    log("free bytes before/after", free0, free1)
    # fmcs_sort_pairs would take 2 x 12 MiB of workspace at this size
    assert free0 - free1 < (4 << 20)
