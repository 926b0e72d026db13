/*
 * fmcs_io.h -- C ABI of the column file I/O around the sorting path (host
 * code; no GPU is needed).
 *
 * Two formats:
 *
 *  1. The reference's text files (src/generator.h:43-136): unsigned decimal
 *     values separated by white space, column after column (ncols columns of
 *     nrows values each).  The TPC-H layout (ReadFromFile_tpch, :101-136)
 *     starts with two more values, "nrows ncols".  A column's bank follows
 *     its bit width as in ReadFromFile (:77-97): width <= 16 -> uint16,
 *     <= 32 -> uint32, else uint64.
 *
 *  2. A binary column file (no counterpart in the reference, which has only
 *     the text files): a 4 KiB-aligned header, then every column as its
 *     little-endian bank array at a 4 KiB-aligned offset, so a column or a
 *     row range of it is read with one positioned read (straight into
 *     pinned host memory for the upload).  Header, 32 bytes + 64 per column:
 *       char magic[8] = "FMCSCOLS"; uint32 version = 1; uint32 ncols;
 *       uint64 nrows; uint64 header_bytes;
 *       per column: uint32 width; uint32 bank_bytes; uint64 cardinality;
 *                   uint64 offset; uint64 bytes; uint64 digest; uint32 flags;
 *                   uint32 reserved; char name[16].
 *     digest = sum over the column's 8-byte words w_i (zero-padded) of
 *     mix64(w_i + i * 0x9E3779B97F4A7C15) mod 2^64, mix64 being SplitMix64's
 *     finaliser; flags bit 0: the column sorts descending.
 *
 * Conventions: host buffers owned by the caller; every function returns
 * FMCS_OK or an error code (FMCS_E_ARG, FMCS_E_NOMEM, FMCS_E_IO: the file
 * cannot be opened, read or written; FMCS_E_FORMAT: its contents do not
 * parse) and never exits.  The reference's readers print and exit on a
 * missing file and do not check the values they read.
 */
#ifndef FMCS_IO_H
#define FMCS_IO_H

#include <stdint.h>

#include "fmcs.h"

#ifdef __cplusplus
extern "C" {
#endif

#define FMCS_COLFILE_MAX_COLS 64
#define FMCS_COLFILE_DESCENDING 1u

/* Message of the calling thread's last failed call of this header. */
const char* fmcs_io_last_error(void);

/* Counts the values of a text file.  With tpch != 0, *nrows and *ncols get
 * its header (ReadFromFile_tpch, src/generator.h:109-111) and *nvalues the
 * values after it; otherwise *nrows = *ncols = 0.  Any pointer may be null. */
int fmcs_text_scan(const char* path, int tpch, uint64_t* nrows, uint32_t* ncols, uint64_t* nvalues);

/* DataGenerator::ReadFromFile / ReadFromFile_tpch (src/generator.h:67-136)
 * with ReadValues (:45-66): column c gets nrows values into cols[c] (host
 * memory of nrows elements of the bank that widths[c] selects).  With tpch,
 * nrows must equal the file's header.  Values past the last column are
 * ignored, as the reference's reader ignores them.  A token that is not an
 * unsigned decimal, a value beyond the column's bank, or fewer than
 * ncols * nrows values give FMCS_E_FORMAT.  cardinality (optional) gets
 * each column's distinct count (ReadValues counts with a std::set).
 * nthreads: parser threads (0: all cores, at most 64). */
int fmcs_text_read(const char* path, int tpch, uint32_t ncols, const uint32_t* widths, uint64_t nrows,
                   void* const* cols, uint64_t* cardinality, uint32_t nthreads);

/* Writes columns in the same text layout, one value per line (the layout of
 * DataGenerator::GenAndSaveAsFile, src/generator.h:138-227); with tpch the
 * "nrows ncols" header line comes first. */
int fmcs_text_write(const char* path, int tpch, uint32_t ncols, const uint32_t* widths, uint64_t nrows,
                    const void* const* cols);

/* One column of a binary column file. */
typedef struct {
  uint32_t width;        /* key bits, 1..64 */
  uint32_t bank_bytes;   /* 1, 2, 4 or 8; 0 on write: from width (<= 16 -> 2, <= 32 -> 4, else 8) */
  uint64_t cardinality;  /* distinct values, 0 = unknown */
  uint32_t flags;        /* FMCS_COLFILE_DESCENDING */
  char name[16];         /* NUL-terminated */
} fmcs_colfile_column;

/* Writes nrows rows of ncols columns (cols[c]: host memory in the column's
 * bank) to a binary column file; the digests are computed here.  Values
 * wider than their width give FMCS_E_ARG. */
int fmcs_colfile_write(const char* path, uint64_t nrows, uint32_t ncols, const fmcs_colfile_column* meta,
                       const void* const* cols);

/* Reads the header: *nrows, *ncols and up to cap column records. */
int fmcs_colfile_info(const char* path, uint64_t* nrows, uint32_t* ncols, fmcs_colfile_column* meta, uint32_t cap);

/* Reads rows row0 .. row0 + nrows - 1 of column col into out (host memory,
 * the column's bank). */
int fmcs_colfile_read(const char* path, uint32_t col, uint64_t row0, uint64_t nrows, void* out);

/* Recomputes every column's digest: FMCS_OK when all match, FMCS_E_FORMAT
 * otherwise. */
int fmcs_colfile_verify(const char* path);

#ifdef __cplusplus
}
#endif

#endif
