"""Python face of the C ABI, over torch device tensors.

torch is used only for device memory; every computation runs in the HIP
kernels of libfmcs.so.  Arguments are torch tensors on the GPU (uint16 /
uint32 / uint64 banks, uint32 oids and group ids).
"""
import ctypes as C

import torch

from . import _lib
from ._lib import Entry, FmcsError, GroupInfo, Tuning, MODE_BATCHED, MODE_PLAIN, MODE_UNBINNED, MODE_DEVICE, MODE_GRAPH

_BANK = {torch.uint8: 1, torch.uint16: 2, torch.uint32: 4, torch.uint64: 8}
_DTYPE = {1: torch.uint8, 2: torch.uint16, 4: torch.uint32, 8: torch.uint64}


def bank_of(t):
    return _BANK[t.dtype]


def dtype_of_bank(b):
    return _DTYPE[b]


def bank_bytes(width, allow_u8=False):
    """src/stitch_composer.cpp:622-650 (sort path) and :883-890 (hash path)."""
    return int(_lib.load().fmcs_bank_bytes(width, int(allow_u8)))


def construct_entries(bw, col_order, split):
    """Step a of run_single_plan (src/stitch_composer.cpp:571-612)."""
    lib = _lib.load()
    nc, ns = len(bw), len(split)
    a_bw = (C.c_uint32 * nc)(*bw)
    a_ord = (C.c_uint32 * nc)(*col_order)
    a_sp = (C.c_uint32 * ns)(*split)
    cap = 256
    ents = (Entry * cap)()
    nout = C.c_uint32(0)
    rc = lib.fmcs_construct_entries(nc, a_bw, a_ord, ns, a_sp, ents, cap, C.byref(nout))
    if rc != 0:
        raise FmcsError(rc, "invalid split for the given widths")
    return [(e.newcol, e.src, e.start, e.end) for e in ents[: nout.value]]


def _p(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


class Context:
    """fmcs_ctx on one device (default: the current torch device)."""

    def __init__(self, device=None):
        self.lib = _lib.load()
        if device is None:
            device = torch.cuda.current_device()
        self.device = device
        h = C.c_void_p()
        rc = self.lib.fmcs_ctx_create(device, C.byref(h))
        if rc != 0:
            msg = self.lib.fmcs_last_error(None)
            raise FmcsError(rc, "fmcs_ctx_create failed: " + (msg.decode() if msg else ""))
        self.h = h

    def close(self):
        if self.h:
            self.lib.fmcs_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != 0:
            msg = self.lib.fmcs_last_error(self.h)
            raise FmcsError(rc, msg.decode() if msg else "")

    def sync(self):
        self._chk(self.lib.fmcs_ctx_sync(self.h))

    def pointer_kind(self, ptr):
        """_lib.MEM_* kind of the memory at address ptr (an int)."""
        k = C.c_uint32(0)
        self._chk(self.lib.fmcs_pointer_kind(self.h, C.c_void_p(ptr), C.byref(k)))
        return k.value

    def set_host_staging(self, on):
        """Stage pageable host buffers through device memory (else they are rejected)."""
        self._chk(self.lib.fmcs_ctx_set_host_staging(self.h, int(bool(on))))

    def set_stream(self, stream):
        self._chk(self.lib.fmcs_ctx_set_stream(self.h, C.c_void_p(stream)))

    def trim(self):
        self._chk(self.lib.fmcs_ctx_trim(self.h))

    def tuning(self):
        t = Tuning()
        self._chk(self.lib.fmcs_ctx_get_tuning(self.h, C.byref(t)))
        return {"small_max": t.small_max, "window": t.window, "msd_bits": t.msd_bits,
                "gather_window_kb": t.gather_window_kb, "carry_next": t.carry_next, "med_max": t.med_max}

    def set_tuning(self, small_max, window, msd_bits, gather_window_kb=None, carry_next=None, med_max=None):
        """med_max 0: level-synchronous batched sort; 1024..16384: the
        persistent size-binned launch with single segments up to med_max rows
        sorted by one workgroup each."""
        cur = self.tuning()
        if gather_window_kb is None:
            gather_window_kb = cur["gather_window_kb"]
        if carry_next is None:
            carry_next = cur["carry_next"]
        if med_max is None:
            med_max = cur["med_max"]
        t = Tuning(small_max, window, msd_bits, gather_window_kb, int(carry_next), int(med_max))
        self._chk(self.lib.fmcs_ctx_set_tuning(self.h, C.byref(t)))

    # ---- kernels -------------------------------------------------------
    def stitch(self, srcs, src_bws, entries, dsts):
        """Fused ShiftLeftBatch + AssembleFromColumn into dsts (device tensors)."""
        n = srcs[0].numel()
        ns, nd, ne = len(srcs), len(dsts), len(entries)
        sp = (C.c_void_p * ns)(*[s.data_ptr() for s in srcs])
        sb = (C.c_uint32 * ns)(*[bank_of(s) for s in srcs])
        sw = (C.c_uint32 * ns)(*src_bws)
        ents = (Entry * ne)(*[Entry(*e) for e in entries])
        dp = (C.c_void_p * nd)(*[d.data_ptr() for d in dsts])
        db = (C.c_uint32 * nd)(*[bank_of(d) for d in dsts])
        self._chk(self.lib.fmcs_stitch(self.h, n, ns, sp, sb, sw, ne, ents, nd, dp, db))

    def gather(self, values, order, out=None):
        if out is None:
            out = torch.empty_like(values)
        self._chk(self.lib.fmcs_gather(self.h, bank_of(values), values.numel(), _p(values), _p(order), _p(out)))
        return out

    def gather_columns(self, columns, order, outs=None, bits=None):
        """outs[k][i] = columns[k][order[i]] for every column, through row records;
        bits: the columns' declared widths (bit-packed records), or None
        (see fmcs_gather_columns)."""
        if outs is None:
            outs = [torch.empty_like(c) for c in columns]
        k = len(columns)
        srcs = (C.c_void_p * k)(*[c.data_ptr() for c in columns])
        dsts = (C.c_void_p * k)(*[o.data_ptr() for o in outs])
        banks = (C.c_uint32 * k)(*[bank_of(c) for c in columns])
        bw = None if bits is None else (C.c_uint32 * k)(*bits)
        self._chk(self.lib.fmcs_gather_columns(self.h, order.numel(), k, srcs, banks, bw, _p(order), dsts))
        return outs

    def extract_group_info(self, sorted_values, in_group=None, out_group=None, seg_offsets=None):
        n = sorted_values.numel()
        if out_group is None:
            out_group = torch.empty(n, dtype=torch.uint32, device=sorted_values.device)
        ng = C.c_uint32(0)
        self._chk(self.lib.fmcs_extract_group_info(self.h, bank_of(sorted_values), n, _p(sorted_values),
                                                   _p(in_group), _p(out_group), _p(seg_offsets), C.byref(ng)))
        return out_group, ng.value

    def sort_pairs(self, keys, oids, keybits):
        self._chk(self.lib.fmcs_sort_pairs(self.h, bank_of(keys), keybits, keys.numel(), _p(keys), _p(oids)))

    def fill_column(self, out, seed, card, zipf, width):
        """Seeded synthetic column into the device tensor `out` (see fmcs_fill_column)."""
        self._chk(self.lib.fmcs_fill_column(self.h, bank_of(out), out.numel(), seed, card, int(bool(zipf)), width,
                                             _p(out)))
        return out

    def sort_pairs_merge(self, keys, oids):
        """Bitonic LDS tiles + merge-path passes; the same result as sort_pairs."""
        self._chk(self.lib.fmcs_sort_pairs_merge(self.h, bank_of(keys), keys.numel(), _p(keys), _p(oids)))

    def segmented_sort(self, keys, oids, seg_offsets, nseg, keybits, mode=MODE_BATCHED, want_info=False):
        info = GroupInfo() if want_info else None
        self._chk(self.lib.fmcs_segmented_sort(self.h, bank_of(keys), keybits, keys.numel(), _p(keys), _p(oids),
                                               _p(seg_offsets), nseg, mode,
                                               C.byref(info) if info is not None else None))
        return info

    def sort_by_group(self, column, order, keybits, in_group=None, out_group=None, last_round=False,
                      mode=MODE_BATCHED, sorted_out=None, want_info=False):
        n = column.numel()
        if not last_round and out_group is None:
            out_group = torch.empty(n, dtype=torch.uint32, device=column.device)
        info = GroupInfo() if want_info else None
        ng = C.c_uint32(0)
        self._chk(self.lib.fmcs_sort_by_group(self.h, bank_of(column), keybits, n, _p(column), _p(order),
                                              _p(in_group), _p(out_group), int(last_round), mode,
                                              _p(sorted_out), C.byref(info) if info is not None else None,
                                              C.byref(ng)))
        return out_group, ng.value, info

    def multi_round_sort(self, cols, bitwidths, order=None, groups=None, mode=MODE_BATCHED, want_count=True):
        """ORDER BY cols (one round per column).  With want_count False the
        group count is not returned (None), which lets the device-driven
        modes run without a host synchronisation."""
        n = cols[0].numel()
        if order is None:
            order = torch.empty(n, dtype=torch.uint32, device=cols[0].device)
        r = len(cols)
        cp = (C.c_void_p * r)(*[c.data_ptr() for c in cols])
        bk = (C.c_uint32 * r)(*[bank_of(c) for c in cols])
        bw = (C.c_uint32 * r)(*bitwidths)
        ng = C.c_uint32(0)
        self._chk(self.lib.fmcs_multi_round_sort(self.h, r, cp, bk, bw, n, _p(order), _p(groups), mode,
                                                 C.byref(ng) if want_count else None))
        return order, (ng.value if want_count else None)

    def set_profiling(self, on):
        self._chk(self.lib.fmcs_ctx_set_profiling(self.h, int(bool(on))))

    def kernel_stats(self, reset=False):
        import json
        buf = C.create_string_buffer(1 << 16)
        self._chk(self.lib.fmcs_kernel_stats(self.h, buf, len(buf)))
        if reset:
            self._chk(self.lib.fmcs_kernel_stats_reset(self.h))
        return json.loads(buf.value.decode())

    def last_breakdown(self):
        a = (C.c_double * 4)()
        self._chk(self.lib.fmcs_last_breakdown(self.h, a))
        return {"gather_ms": a[0], "first_round_ms": a[1], "later_rounds_ms": a[2], "extract_ms": a[3]}

    def hash_round(self, values, order, in_group=None, last_round=False):
        n = values.numel()
        dev = values.device
        gid = torch.empty(n, dtype=torch.uint32, device=dev)
        oo = og = None
        if not last_round:
            oo = torch.empty(n, dtype=torch.uint32, device=dev)
            og = torch.empty(n, dtype=torch.uint32, device=dev)
        ng = C.c_uint32(0)
        self._chk(self.lib.fmcs_hash_round(self.h, bank_of(values), n, _p(values), _p(order), _p(in_group),
                                           _p(gid), _p(oo), _p(og), int(last_round), C.byref(ng)))
        return gid, oo, og, ng.value

    def probe_handoff(self, tasks=65536, grid=1024):
        """Milliseconds of the hand-off probe (fmcs_probe_handoff): modes 0 none,
        1 every wave releases per task, 2 one release after a barrier per
        task, 3 one release per 8 tasks of a block."""
        ms = (C.c_double * 4)()
        self._chk(self.lib.fmcs_probe_handoff(self.h, tasks, grid, ms))
        return list(ms)

    def distinct_count(self, col):
        """Exact number of distinct values of a device column (hash grouping)."""
        c = C.c_uint64(0)
        self._chk(self.lib.fmcs_distinct_count(self.h, bank_of(col), col.numel(), _p(col), C.byref(c)))
        return c.value

    def multi_round_hash(self, cols, hash_type=1):
        n = cols[0].numel()
        dev = cols[0].device
        r = len(cols)
        cp = (C.c_void_p * r)(*[c.data_ptr() for c in cols])
        bk = (C.c_uint32 * r)(*[bank_of(c) for c in cols])
        order = torch.empty(n, dtype=torch.uint32, device=dev)
        gid = torch.empty(n, dtype=torch.uint32, device=dev)
        ng = C.c_uint32(0)
        self._chk(self.lib.fmcs_multi_round_hash(self.h, r, cp, bk, n, hash_type, _p(order), _p(gid),
                                                 C.byref(ng)))
        return order, gid, ng.value


__all__ = ["Context", "bank_bytes", "construct_entries", "bank_of", "dtype_of_bank",
           "MODE_BATCHED", "MODE_PLAIN", "MODE_UNBINNED", "MODE_DEVICE", "MODE_GRAPH"]
