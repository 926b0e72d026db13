"""Column files: the reference's text layout (src/generator.h:43-227) and the
binary column file, through the C ABI of include/fmcs_io.h (host code,
libfmcs.so).  Columns are numpy arrays in their banks."""
import ctypes as C

import numpy as np

from . import _lib

FMCS_E_IO = 5
FMCS_E_FORMAT = 6
MAX_COLS = 64
DESCENDING = 1


class ColfileColumn(C.Structure):
    _fields_ = [("width", C.c_uint32), ("bank_bytes", C.c_uint32), ("cardinality", C.c_uint64),
                ("flags", C.c_uint32), ("name", C.c_char * 16)]


_bound = None


def lib():
    global _bound
    lb = _lib.load()
    if _bound is None:
        u32p, u64p, vpp = C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(C.c_void_p)
        lb.fmcs_io_last_error.restype = C.c_char_p
        lb.fmcs_io_last_error.argtypes = []
        lb.fmcs_text_scan.argtypes = [C.c_char_p, C.c_int, u64p, u32p, u64p]
        lb.fmcs_text_read.argtypes = [C.c_char_p, C.c_int, C.c_uint32, u32p, C.c_uint64, vpp, u64p, C.c_uint32]
        lb.fmcs_text_write.argtypes = [C.c_char_p, C.c_int, C.c_uint32, u32p, C.c_uint64, vpp]
        lb.fmcs_colfile_write.argtypes = [C.c_char_p, C.c_uint64, C.c_uint32, C.POINTER(ColfileColumn), vpp]
        lb.fmcs_colfile_info.argtypes = [C.c_char_p, u64p, u32p, C.POINTER(ColfileColumn), C.c_uint32]
        lb.fmcs_colfile_read.argtypes = [C.c_char_p, C.c_uint32, C.c_uint64, C.c_uint64, C.c_void_p]
        lb.fmcs_colfile_verify.argtypes = [C.c_char_p]
        _bound = True
    return lb


EXPORTS = ["fmcs_io_last_error", "fmcs_text_scan", "fmcs_text_read", "fmcs_text_write", "fmcs_colfile_write",
           "fmcs_colfile_info", "fmcs_colfile_read", "fmcs_colfile_verify"]


class ColumnFileError(_lib.FmcsError):
    pass


def _chk(rc, what):
    if rc != 0:
        raise ColumnFileError(rc, f"{what}: {lib().fmcs_io_last_error().decode()}")


def text_dtype(width):
    """Bank of a text column (ReadFromFile, src/generator.h:77-97)."""
    if not 1 <= width <= 64:
        raise ValueError("width outside 1..64")
    return np.uint16 if width <= 16 else (np.uint32 if width <= 32 else np.uint64)


def _u32(xs):
    return (C.c_uint32 * len(xs))(*[int(x) for x in xs])


def _path(p):
    return str(p).encode()


def text_scan(path, tpch=False):
    """(nrows, ncols, nvalues): the header of a TPC-H layout file (0, 0 for a
    plain one) and the number of values after it."""
    nr, nc, nv = C.c_uint64(0), C.c_uint32(0), C.c_uint64(0)
    _chk(lib().fmcs_text_scan(_path(path), int(bool(tpch)), C.byref(nr), C.byref(nc), C.byref(nv)), "text_scan")
    return nr.value, nc.value, nv.value


def read_text(path, widths, nrows=None, tpch=False, nthreads=0, cardinality=True):
    """DataGenerator::ReadFromFile / ReadFromFile_tpch: returns (columns,
    cardinalities or None).  nrows: rows per column (the header's for a
    TPC-H layout file when None; for a plain file, values / columns)."""
    widths = [int(w) for w in widths]
    k = len(widths)
    if nrows is None:
        hr, _, nv = text_scan(path, tpch)
        nrows = hr if tpch else nv // k
    cols = [np.empty(nrows, dtype=text_dtype(w)) for w in widths]
    ptrs = (C.c_void_p * k)(*[c.ctypes.data for c in cols])
    card = (C.c_uint64 * k)() if cardinality else None
    _chk(lib().fmcs_text_read(_path(path), int(bool(tpch)), k, _u32(widths), nrows, ptrs, card, nthreads),
         "text_read")
    return cols, (list(card) if cardinality else None)


def write_text(path, cols, widths, tpch=False):
    """Columns in the reference's text layout, one value per line (with the
    "nrows ncols" header line for the TPC-H layout)."""
    widths = [int(w) for w in widths]
    cols = [np.ascontiguousarray(c, dtype=text_dtype(w)) for c, w in zip(cols, widths)]
    n = len(cols[0])
    if any(len(c) != n for c in cols):
        raise ValueError("columns differ in length")
    ptrs = (C.c_void_p * len(cols))(*[c.ctypes.data for c in cols])
    _chk(lib().fmcs_text_write(_path(path), int(bool(tpch)), len(cols), _u32(widths), n, ptrs), "text_write")


def _bank_bytes(a):
    return np.dtype(a.dtype).itemsize


def write_columns(path, cols, widths=None, names=None, cardinalities=None, descending=None):
    """Binary column file.  cols: unsigned numpy arrays (their dtype is the
    bank); widths default to the bank's bits."""
    k = len(cols)
    if not 1 <= k <= MAX_COLS:
        raise ValueError("1..64 columns")
    cols = [np.ascontiguousarray(c) for c in cols]
    n = len(cols[0])
    if any(len(c) != n for c in cols):
        raise ValueError("columns differ in length")
    meta = (ColfileColumn * k)()
    for i, c in enumerate(cols):
        if c.dtype not in (np.uint8, np.uint16, np.uint32, np.uint64):
            raise ValueError("columns must be unsigned integer arrays")
        meta[i].bank_bytes = _bank_bytes(c)
        meta[i].width = int(widths[i]) if widths is not None else 8 * _bank_bytes(c)
        meta[i].cardinality = int(cardinalities[i]) if cardinalities is not None else 0
        meta[i].flags = DESCENDING if descending is not None and descending[i] else 0
        nm = (names[i] if names is not None else f"c{i}").encode()
        if len(nm) > 15:
            raise ValueError("column names hold at most 15 bytes")
        meta[i].name = nm
    ptrs = (C.c_void_p * k)(*[c.ctypes.data for c in cols])
    _chk(lib().fmcs_colfile_write(_path(path), n, k, meta, ptrs), "colfile_write")


_DT = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}


def columns_info(path):
    """{'nrows', 'columns': [{'name', 'width', 'bank_bytes', 'cardinality', 'descending'}]}"""
    nr, nc = C.c_uint64(0), C.c_uint32(0)
    meta = (ColfileColumn * MAX_COLS)()
    _chk(lib().fmcs_colfile_info(_path(path), C.byref(nr), C.byref(nc), meta, MAX_COLS), "colfile_info")
    cols = [{"name": meta[i].name.decode(), "width": meta[i].width, "bank_bytes": meta[i].bank_bytes,
             "cardinality": meta[i].cardinality, "descending": bool(meta[i].flags & DESCENDING)}
            for i in range(nc.value)]
    return {"nrows": nr.value, "columns": cols}


def read_columns(path, which=None, row0=0, nrows=None, out=None):
    """Columns of a binary column file (all, or the indices / names in
    `which`), rows row0 .. row0 + nrows - 1.  out: optional arrays to fill
    (e.g. pinned host tensors' numpy views)."""
    info = columns_info(path)
    names = [c["name"] for c in info["columns"]]
    if which is None:
        which = list(range(len(names)))
    idx = [names.index(w) if isinstance(w, str) else int(w) for w in which]
    if nrows is None:
        nrows = info["nrows"] - row0
    res = []
    for j, i in enumerate(idx):
        dt = _DT[info["columns"][i]["bank_bytes"]]
        a = out[j] if out is not None else np.empty(nrows, dtype=dt)
        if a.dtype != dt or len(a) < nrows or not a.flags["C_CONTIGUOUS"]:
            raise ValueError(f"output {j} must be a contiguous {np.dtype(dt).name} array of {nrows} rows")
        _chk(lib().fmcs_colfile_read(_path(path), i, row0, nrows, C.c_void_p(a.ctypes.data)), "colfile_read")
        res.append(a[:nrows])
    return res


def verify(path):
    """Checks every column against its digest (raises on a mismatch)."""
    _chk(lib().fmcs_colfile_verify(_path(path)), "colfile_verify")
