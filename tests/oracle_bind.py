"""ctypes bindings to the CPU oracle (oracle/libfmcs_oracle.so) and, when it
was built in this container, to the reference's own code (oracle/_ref/libref.so).

Test infrastructure only: these are the checkers, never the thing measured.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "libfmcs_oracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref.so")

_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")


class GroupInfo(C.Structure):
    _fields_ = [("ngroups", C.c_uint32), ("sum_group_size", C.c_uint32),
                ("ngroups_incache", C.c_uint32), ("ngroups_outofcache", C.c_uint32),
                ("ngroups_gtOne", C.c_uint32)]


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            import subprocess
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        _lib = C.CDLL(ORACLE_SO)
        _lib.orc_bank_bytes.restype = C.c_uint32
        _lib.orc_construct_entries.restype = C.c_int
        _lib.orc_extract_group_info.restype = C.c_uint32
        _lib.orc_sort_by_group.restype = C.c_uint32
        _lib.orc_multi_round_sort.restype = C.c_uint32
        _lib.orc_subgroup_range.restype = C.c_uint32
        _lib.orc_multi_round_hash.restype = C.c_uint32
        _lib.orc_hashmask.restype = C.c_uint64
        _lib.orc_hashmask.argtypes = [C.c_uint64]
        _lib.orc_split_exhaustive.restype = C.c_uint64
    return _lib


def ref():
    """The reference library, or None where it was not built (e.g. GPU box)."""
    global _ref
    if _ref is None and os.path.exists(REF_SO):
        _ref = C.CDLL(REF_SO)
    return _ref


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _ptrs(arrs):
    return (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


def _u32(xs):
    a = np.ascontiguousarray(xs, dtype=np.uint32)
    return a


def bank_bytes(width, allow_u8=False):
    return lib().orc_bank_bytes(C.c_uint32(width), C.c_int(int(allow_u8)))


BANK_DTYPE = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}


def construct_entries(bw, col_order, split):
    bw = _u32(bw); col_order = _u32(col_order); split = _u32(split)
    cap = 256
    out = [np.zeros(cap, np.uint32) for _ in range(4)]
    n = lib().orc_construct_entries(C.c_uint32(len(bw)), _p(bw), _p(col_order),
                                    C.c_uint32(len(split)), _p(split),
                                    *[_p(o) for o in out], C.c_uint32(cap))
    if n < 0:
        raise ValueError("invalid split")
    return [tuple(int(o[i]) for o in out) for i in range(n)]  # (newcol, src, start, end)


def stitch(cols, bws, entries, split, use_ref=False):
    n = len(cols[0])
    src = [np.ascontiguousarray(c) for c in cols]
    src_bank = _u32([c.dtype.itemsize for c in src])
    src_bw = _u32(bws)
    dst_bank = _u32([bank_bytes(w) for w in split])
    dst = [np.zeros(n, BANK_DTYPE[int(b)]) for b in dst_bank]
    e = np.array(entries, dtype=np.uint32).reshape(-1, 4)
    cols4 = [np.ascontiguousarray(e[:, i]) for i in range(4)]
    fn = ref().ref_stitch if use_ref else lib().orc_stitch
    fn(C.c_uint64(n), _ptrs(src), _p(src_bank), _p(src_bw), C.c_uint32(len(e)),
       *[_p(c) for c in cols4], C.c_uint32(len(dst)), _ptrs(dst), _p(dst_bank))
    return dst


def gather(values, order):
    out = np.empty_like(values)
    lib().orc_gather(C.c_uint32(values.dtype.itemsize), C.c_uint64(len(values)), _p(values),
                     _p(order), _p(out))
    return out


def extract_group_info(sorted_values, in_group=None):
    n = len(sorted_values)
    out = np.zeros(n, np.uint32)
    ng = lib().orc_extract_group_info(C.c_uint32(sorted_values.dtype.itemsize), C.c_uint64(n),
                                      _p(sorted_values),
                                      None if in_group is None else _p(in_group), _p(out))
    return out, ng


def sort_by_group(column, order, in_group=None, last_round=False):
    n = len(column)
    order = order.copy()
    out_group = np.zeros(n, np.uint32)
    sorted_out = np.empty_like(column)
    info = GroupInfo()
    ng = lib().orc_sort_by_group(C.c_uint32(column.dtype.itemsize), C.c_uint64(n), _p(column),
                                 _p(order), None if in_group is None else _p(in_group),
                                 _p(out_group), C.c_int(int(last_round)), _p(sorted_out),
                                 C.byref(info))
    return order, (None if last_round else out_group), ng, sorted_out, info


def multi_round_sort(cols):
    n = len(cols[0])
    cols = [np.ascontiguousarray(c) for c in cols]
    banks = _u32([c.dtype.itemsize for c in cols])
    order = np.zeros(n, np.uint32)
    r = len(cols)
    groups = np.zeros(max(r - 1, 1) * n, np.uint32)
    info = GroupInfo()
    ng = lib().orc_multi_round_sort(C.c_uint32(r), _ptrs(cols), _p(banks), C.c_uint64(n),
                                    _p(order), _p(groups), C.byref(info))
    return order, groups[: (r - 1) * n].reshape(r - 1, n), ng, info


def multi_round_hash(cols):
    n = len(cols[0])
    cols = [np.ascontiguousarray(c) for c in cols]
    banks = _u32([c.dtype.itemsize for c in cols])
    order = np.zeros(n, np.uint32)
    gid = np.zeros(n, np.uint32)
    r = len(cols)
    rg = np.zeros(max(r - 1, 1) * n, np.uint32)
    total = lib().orc_multi_round_hash(C.c_uint32(r), _ptrs(cols), _p(banks), C.c_uint64(n),
                                       _p(order), _p(gid), _p(rg))
    return order, gid, rg[: (r - 1) * n].reshape(r - 1, n), total


def hashmask(cnt):
    return int(lib().orc_hashmask(C.c_uint64(cnt)))


def split_exhaustive(S, rounds, maxv=64, cap=200000):
    out = np.zeros(cap * rounds, np.uint32)
    cnt = lib().orc_split_exhaustive(C.c_uint32(S), C.c_uint32(rounds), C.c_uint32(maxv), _p(out),
                                     C.c_uint64(cap))
    k = min(cnt, cap)
    return int(cnt), out[: k * rounds].reshape(k, rounds)


# ---- the reference itself (this container only) ----

def ref_sort_rounds(cols, bitwidths):
    n = len(cols[0])
    cols = [np.ascontiguousarray(c) for c in cols]
    banks = _u32([c.dtype.itemsize for c in cols])
    bws = _u32(bitwidths)
    order = np.zeros(n, np.uint32)
    r = len(cols)
    groups = np.zeros(max(r - 1, 1) * n, np.uint32)
    rc = ref().ref_sort_rounds(C.c_uint32(r), _ptrs(cols), _p(banks), _p(bws), C.c_uint64(n),
                               _p(order), _p(groups))
    assert rc == 0
    return order, groups[: (r - 1) * n].reshape(r - 1, n)


def ref_sort_rounds_mt(cols, bitwidths, nthreads):
    """ref_sort_rounds on nthreads threads (the reference's kernels under a
    parallel driver): (order, groups of the rounds before the last)."""
    n = len(cols[0])
    cols = [np.ascontiguousarray(c) for c in cols]
    banks = _u32([c.dtype.itemsize for c in cols])
    bws = _u32(bitwidths)
    order = np.zeros(n, np.uint32)
    r = len(cols)
    groups = np.zeros(max(r - 1, 1) * n, np.uint32)
    rc = ref().ref_sort_rounds_mt(C.c_uint32(nthreads), C.c_uint32(r), _ptrs(cols), _p(banks), _p(bws),
                                  C.c_uint64(n), _p(order), _p(groups))
    assert rc == 0
    return order, groups[: (r - 1) * n].reshape(r - 1, n)


def ref_mergesort(keys, oids):
    keys = keys.copy(); oids = oids.copy()
    rc = ref().ref_mergesort(C.c_uint32(keys.dtype.itemsize), _p(keys), _p(oids),
                             C.c_uint64(len(keys)))
    assert rc == 0
    return keys, oids


def ref_read_text(path, widths, nrows=0, tpch=False):
    """DataGenerator::ReadFromFile / ReadFromFile_tpch (src/generator.h:67-136)
    of the reference itself: (columns, cardinalities).  nrows: the plain
    layout's row count (the TPC-H layout reads its header)."""
    k = len(widths)
    if tpch:
        # the header gives the row count; read it first to size the outputs
        with open(path) as f:
            nrows = int(f.read(64).split()[0])
    dts = [np.uint16 if w <= 16 else (np.uint32 if w <= 32 else np.uint64) for w in widths]
    cols = [np.zeros(nrows, dt) for dt in dts]
    card = np.zeros(k, np.uint64)
    nr = C.c_uint64(nrows)
    w = _u32(widths)
    rc = ref().ref_read_text(str(path).encode(), C.c_int(int(bool(tpch))), C.c_uint32(k), _p(w),
                             C.byref(nr), _ptrs(cols), _p(card))
    assert rc == 0 and nr.value == nrows
    return cols, [int(x) for x in card]


def canonical_order(full_key_sorted_runs, order):
    """Sort oids ascending within runs of equal full keys: the reference's
    tie order is arbitrary, the contract is the oid tie-break."""
    out = order.copy()
    keys = full_key_sorted_runs
    n = len(order)
    if n == 0:
        return out
    brk = np.flatnonzero(np.any(keys[:, 1:] != keys[:, :-1], axis=0)) + 1 if keys.ndim == 2 \
        else np.flatnonzero(keys[1:] != keys[:-1]) + 1
    starts = np.concatenate([[0], brk])
    ends = np.concatenate([brk, [n]])
    for s, e in zip(starts, ends):
        if e - s > 1:
            out[s:e] = np.sort(out[s:e])
    return out


# ---- the reference's planner (oracle/_ref) ----------------------------------

def _ref_planner_sigs(r):
    u32, u64, vp, dbl = C.c_uint32, C.c_uint64, C.c_void_p, C.c_double
    P = C.POINTER
    sigs = {
        "ref_planner_new": (vp, [u64, u32, P(u32), P(u64)]),
        "ref_planner_free": (None, [vp]),
        "ref_card_stats": (None, [vp, P(u32), u32, P(u32), P(dbl), P(u32), P(u32)]),
        "ref_group_num": (dbl, [vp, u32, P(u32)]),
        "ref_column_cost": (dbl, [vp, dbl, u32]),
        "ref_greedy_plans": (C.c_int, [vp, u32, P(u32), P(u32), u64, P(u64), P(u64)]),
        "ref_plan_cost": (dbl, [vp, P(u32), u32, P(u32)]),
        "ref_rrs_cost": (dbl, [u64, u32, P(u32), u32, P(u32), P(u64), u32, P(u32), P(u64), P(u32), P(u32)]),
        "ref_plan_enum_rrs": (C.c_int, [u64, u32, P(u32), u32, P(u32), P(u64), P(u64), P(u32), u64, P(u64),
                                        P(u64)]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(r, name)
        fn.restype = res
        fn.argtypes = args


def _a32(xs):
    return (C.c_uint32 * len(xs))(*[int(x) for x in xs])


def _a64(xs):
    return (C.c_uint64 * len(xs))(*[int(x) for x in xs])


def _unflatten(buf, words):
    out, i = [], 0
    while i < words:
        k = buf[i]
        out.append([int(x) for x in buf[i + 1:i + 1 + k]])
        i += 1 + k
    return out


class RefPlanner:
    """The reference's StitchComposer planner over columns of the given widths
    and cardinalities (no values)."""

    def __init__(self, n, bws, cards):
        r = ref()
        _ref_planner_sigs(r)
        self.r, self.n, self.bws, self.cards = r, n, list(bws), list(cards)
        self.h = r.ref_planner_new(n, len(bws), _a32(bws), _a64(cards))

    def close(self):
        if self.h:
            self.r.ref_planner_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def card_stats(self, order, split):
        card = (C.c_double * len(split))()
        a = (C.c_uint32 * 9)()
        s = (C.c_uint32 * 3)()
        self.r.ref_card_stats(self.h, _a32(order), len(split), _a32(split), card, a, s)
        return list(card), list(a), list(s)

    def group_num(self, bitpos, order):
        return self.r.ref_group_num(self.h, bitpos, _a32(order))

    def column_cost(self, g, bank_type):
        return self.r.ref_column_cost(self.h, g, bank_type)

    def greedy_plans(self, rounds, order):
        buf = (C.c_uint32 * 100000)()
        npl, words = C.c_uint64(), C.c_uint64()
        assert self.r.ref_greedy_plans(self.h, rounds, _a32(order), buf, len(buf), C.byref(npl), C.byref(words)) == 0
        return _unflatten(buf, words.value)

    def plan_cost(self, order, split):
        return self.r.ref_plan_cost(self.h, _a32(order), len(split), _a32(split))

    def rrs_cost(self, order, split, max_round):
        card = (C.c_uint64 * len(split))()
        a = (C.c_uint32 * 9)()
        s = (C.c_uint32 * 3)()
        c = self.r.ref_rrs_cost(self.n, max_round, _a32(order), len(self.bws), _a32(self.bws), _a64(self.cards),
                                len(split), _a32(split), card, a, s)
        return c, list(card), list(a), list(s)

    def plan_enum_rrs(self, order, max_round):
        """(seed the reference's engine took, visited plans)"""
        buf = (C.c_uint32 * 400000)()
        seed, npl, words = C.c_uint64(), C.c_uint64(), C.c_uint64()
        assert self.r.ref_plan_enum_rrs(self.n, max_round, _a32(order), len(self.bws), _a32(self.bws),
                                        _a64(self.cards), C.byref(seed), buf, len(buf), C.byref(npl),
                                        C.byref(words)) == 0
        return seed.value, _unflatten(buf, words.value)


def ref_cost_callbacks():
    """Addresses of the reference's static estimators, in the callback
    signatures of fmcs_cost_primitives (sort, stitch, reconstruct)."""
    r = ref()
    return tuple(C.cast(getattr(r, n), C.c_void_p).value for n in ("ref_cb_sort", "ref_cb_stitch",
                                                                      "ref_cb_reconstruct"))
