"""ctypes binding of libfmcs.so (the C ABI declared in include/fmcs.h).

The library is the only compute path: if it is missing this module raises,
there is no fallback.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# FMCS_LIB selects another build of the same library (a path to a .so)
LIB_PATH = os.environ.get("FMCS_LIB") or os.path.join(HERE, "libfmcs.so")

FMCS_OK = 0
FMCS_E_ARG = 1
FMCS_E_NOMEM = 2
FMCS_E_HIP = 3
FMCS_E_UNSUPPORTED = 4

MODE_BATCHED = 0
MODE_PLAIN = 1
MODE_UNBINNED = 2
MODE_DEVICE = 3
MODE_GRAPH = 4

# Every exported symbol of include/fmcs.h.
EXPORTS = [
    "fmcs_ctx_create", "fmcs_ctx_destroy", "fmcs_ctx_set_stream", "fmcs_ctx_sync",
    "fmcs_ctx_set_tuning", "fmcs_ctx_get_tuning", "fmcs_ctx_trim", "fmcs_last_error",
    "fmcs_bank_bytes", "fmcs_construct_entries", "fmcs_stitch", "fmcs_stitch_dir", "fmcs_gather",
    "fmcs_gather_columns",
    "fmcs_extract_group_info", "fmcs_sort_pairs", "fmcs_sort_pairs_noalloc", "fmcs_sort_pairs_merge", "fmcs_fill_column", "fmcs_segmented_sort",
    "fmcs_sort_by_group",
    "fmcs_multi_round_sort", "fmcs_last_breakdown", "fmcs_hash_round", "fmcs_multi_round_hash",
    "fmcs_ctx_set_profiling", "fmcs_kernel_stats", "fmcs_kernel_stats_reset", "fmcs_distinct_count",
    "fmcs_classify_memory", "fmcs_pointer_kind", "fmcs_ctx_set_host_staging", "fmcs_probe_handoff",
]

# fmcs_pointer_kind / fmcs_classify_memory
MEM_DEVICE = 0
MEM_PINNED = 1
MEM_MANAGED = 2
MEM_HOST = 3
MEM_OTHER = 4


class Entry(C.Structure):
    _fields_ = [("newcol", C.c_uint32), ("src", C.c_uint32), ("start", C.c_uint32), ("end", C.c_uint32)]


class GroupInfo(C.Structure):
    _fields_ = [("ngroups", C.c_uint32), ("sum_group_size", C.c_uint32),
                ("ngroups_incache", C.c_uint32), ("ngroups_outofcache", C.c_uint32),
                ("ngroups_gtOne", C.c_uint32)]


class Tuning(C.Structure):
    _fields_ = [("small_max", C.c_uint32), ("window", C.c_uint32), ("msd_bits", C.c_uint32),
                ("gather_window_kb", C.c_uint32), ("carry_next", C.c_uint32), ("med_max", C.c_uint32)]


class FmcsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fmcs error {code}: {msg}")
        self.code = code


_lib = None


def load():
    """Load libfmcs.so; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    # torch ships its own HIP runtime; importing it first makes the library
    # bind to that one.  Loaded first, the library would bring in the ROCm
    # install's runtime beside torch's, and one process with two runtimes
    # fails to open the GPU in whichever initialises second.
    import torch  # noqa: F401
    lib = C.CDLL(LIB_PATH)
    vp, u32, u64, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
    P = C.POINTER
    sig = {
        "fmcs_ctx_create": (i32, [i32, P(vp)]),
        "fmcs_ctx_destroy": (i32, [vp]),
        "fmcs_ctx_set_stream": (i32, [vp, vp]),
        "fmcs_ctx_sync": (i32, [vp]),
        "fmcs_ctx_set_tuning": (i32, [vp, P(Tuning)]),
        "fmcs_sort_pairs_merge": (i32, [vp, u32, u64, vp, vp]),
        "fmcs_fill_column": (i32, [vp, u32, u64, u64, u64, i32, u32, vp]),
        "fmcs_ctx_get_tuning": (i32, [vp, P(Tuning)]),
        "fmcs_ctx_trim": (i32, [vp]),
        "fmcs_last_error": (C.c_char_p, [vp]),
        "fmcs_bank_bytes": (u32, [u32, i32]),
        "fmcs_construct_entries": (i32, [u32, P(u32), P(u32), u32, P(u32), P(Entry), u32, P(u32)]),
        "fmcs_stitch": (i32, [vp, u64, u32, P(vp), P(u32), P(u32), u32, P(Entry), u32, P(vp), P(u32)]),
        "fmcs_stitch_dir": (i32, [vp, u64, u32, P(vp), P(u32), P(u32), u32, P(Entry), u32, P(vp), P(u32),
                                  P(C.c_uint8)]),
        "fmcs_gather": (i32, [vp, u32, u64, vp, vp, vp]),
        "fmcs_gather_columns": (i32, [vp, u64, u32, P(vp), P(u32), P(u32), vp, P(vp)]),
        "fmcs_extract_group_info": (i32, [vp, u32, u64, vp, vp, vp, vp, P(u32)]),
        "fmcs_sort_pairs": (i32, [vp, u32, u32, u64, vp, vp]),
        "fmcs_sort_pairs_noalloc": (i32, [vp, u32, u32, u64, vp, vp]),
        "fmcs_segmented_sort": (i32, [vp, u32, u32, u64, vp, vp, vp, u32, i32, P(GroupInfo)]),
        "fmcs_sort_by_group": (i32, [vp, u32, u32, u64, vp, vp, vp, vp, i32, i32, vp, P(GroupInfo), P(u32)]),
        "fmcs_multi_round_sort": (i32, [vp, u32, P(vp), P(u32), P(u32), u64, vp, vp, i32, P(u32)]),
        "fmcs_last_breakdown": (i32, [vp, P(C.c_double)]),
        "fmcs_hash_round": (i32, [vp, u32, u64, vp, vp, vp, vp, vp, vp, i32, P(u32)]),
        "fmcs_multi_round_hash": (i32, [vp, u32, P(vp), P(u32), u64, i32, vp, vp, P(u32)]),
        "fmcs_distinct_count": (i32, [vp, u32, u64, vp, P(u64)]),
        "fmcs_ctx_set_profiling": (i32, [vp, i32]),
        "fmcs_kernel_stats": (i32, [vp, C.c_char_p, u64]),
        "fmcs_kernel_stats_reset": (i32, [vp]),
        "fmcs_classify_memory": (u32, [i32, i32, i32, i32]),
        "fmcs_pointer_kind": (i32, [vp, vp, P(u32)]),
        "fmcs_ctx_set_host_staging": (i32, [vp, i32]),
        "fmcs_probe_handoff": (i32, [vp, u32, u32, P(C.c_double)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    lib = load()
    return [s for s in EXPORTS if hasattr(lib, s)]
