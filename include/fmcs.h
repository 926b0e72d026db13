/*
 * fmcs.h -- C ABI of the MI355X-native multi-column sorting path.
 *
 * The reference (FastMultiColSort) is a C++ class library with no C ABI;
 * each entry point below replaces one of its hot-path seams, cited as
 * file:line in the reference tree.  Conventions:
 *   - every array argument named *_d is owned by the caller and is memory
 *     the GPU can read (device, pinned or managed); pageable host memory is
 *     rejected with FMCS_E_ARG before anything runs, unless the context
 *     stages host buffers (fmcs_ctx_set_host_staging, below); host arrays
 *     (lists of pointers, widths, entries) are plain host memory;
 *   - results go to caller-named outputs; nothing is swapped or freed;
 *   - calls are ordered on the context's stream; functions that return a
 *     count to the host synchronize that stream;
 *   - every function returns FMCS_OK or an error code and never exits;
 *     fmcs_last_error() gives the message of the last failure;
 *   - a context is not shared between threads.
 * A bank is the 1/2/4/8-byte container of a column; its value uses the low
 * `bitwidth` bits.  Oids are uint32 (src/params.h:9-13).
 */
#ifndef FMCS_H
#define FMCS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FMCS_OK 0
#define FMCS_E_ARG 1
#define FMCS_E_NOMEM 2
#define FMCS_E_HIP 3
#define FMCS_E_UNSUPPORTED 4
#define FMCS_E_IO 5     /* a file cannot be opened, read or written (fmcs_io.h) */
#define FMCS_E_FORMAT 6 /* a file's contents do not parse (fmcs_io.h) */

/* Segmented-sort forms for rounds > 1. */
#define FMCS_MODE_BATCHED 0  /* size-binned: packed LDS tiles + segmented MSD radix */
#define FMCS_MODE_PLAIN 1    /* one sort dispatch per group (src/composer.cpp:1285-1323) */
#define FMCS_MODE_UNBINNED 2 /* one LSD radix sort of (group id, key) composites */
/* The batched form, device-driven: counts stay on the GPU and every kernel
 * reads its sizes there, so a call issues no host synchronisation (unless it
 * returns a count).  Same results as FMCS_MODE_BATCHED. */
#define FMCS_MODE_DEVICE 3
/* FMCS_MODE_DEVICE captured in a hipGraph: fmcs_multi_round_sort replays the
 * context's cached graph when called again with the same columns, widths,
 * n, output pointers and tuning, else runs the call and captures it. */
#define FMCS_MODE_GRAPH 4

typedef struct fmcs_ctx fmcs_ctx;

/* One unit of bit movement into a new column: construct_entry_t
 * (src/types.h:42-46) plus the index of the new column it builds.
 * [start, end] is 1-based from the source column's MSB, inclusive. */
typedef struct {
  uint32_t newcol;
  uint32_t src;
  uint32_t start;
  uint32_t end;
} fmcs_entry;

/* group_info_t (src/types.h:188-195). */
typedef struct {
  uint32_t ngroups;
  uint32_t sum_group_size;
  uint32_t ngroups_incache;
  uint32_t ngroups_outofcache;
  uint32_t ngroups_gtOne;
} fmcs_group_info;

/* Tunables of the batched segmented sort and of the oid gather. */
typedef struct {
  uint32_t small_max;        /* segments up to this many rows go to packed LDS tiles (64..1024; tiles of
                                512 / 1024 / 2048 rows for up to 256 / 512 / 1024; the eager batched mode
                                uses at most 512) */
  uint32_t window;           /* rows per packing window (<= small_max + 1) */
  uint32_t msd_bits;         /* digit width of one MSD level (1..8) */
  uint32_t gather_window_kb; /* 0: one gather pass; else passes over source windows of this many KiB */
  uint32_t carry_next;       /* device-driven modes, how round 2's column reaches round-1 order: 1 = the
                                first round's radix passes carry it along, packed with the first column
                                into one u64 composite key when the first is at most 32 bits, the second
                                is in a u64 bank and the two widths add up to at most 64 (this form also
                                reports second-column values above their width); 2 = carried as a
                                separate column; 0 = gathered after round 1 */
  uint32_t med_max;          /* device-driven modes: 0 = every radix level sorts its own windows (the
                                level-synchronous form); else 1024..16384: the levels split only segments
                                above med_max rows and one persistent launch then sorts every level's
                                windows and single segments, those above 1024 rows each by one workgroup
                                (an MSD level of its own, then its child windows).  The persistent form
                                uses 1024-row window tiles (small_max at most 512). */
} fmcs_tuning;

/* ---- context -------------------------------------------------------- */
int fmcs_ctx_create(int device, fmcs_ctx** out);
int fmcs_ctx_destroy(fmcs_ctx* ctx);
/* hipStream_t as void*; NULL selects the default stream. */
int fmcs_ctx_set_stream(fmcs_ctx* ctx, void* stream);
int fmcs_ctx_sync(fmcs_ctx* ctx);
int fmcs_ctx_set_tuning(fmcs_ctx* ctx, const fmcs_tuning* t);
int fmcs_ctx_get_tuning(fmcs_ctx* ctx, fmcs_tuning* t);
/* Frees the context's cached workspace. */
int fmcs_ctx_trim(fmcs_ctx* ctx);
const char* fmcs_last_error(fmcs_ctx* ctx);

/* ---- buffer arguments ------------------------------------------------- */

/* Every entry point classifies its non-null *_d pointers before it launches
 * anything.  Device memory of the context's GPU, pinned host memory
 * (hipHostMalloc) and managed memory are used in place.  Pageable host
 * memory -- the reference's malloc_aligned buffers (src/common.h:19-31) --
 * is rejected with FMCS_E_ARG, unless the context stages host buffers:
 * then each such argument is copied into a device buffer of the context
 * before the call's work and, if it is an output, copied back and the
 * stream synchronized before the call returns.  Another GPU's memory is
 * rejected with FMCS_E_ARG. */
#define FMCS_MEM_DEVICE 0  /* device memory of the context's GPU */
#define FMCS_MEM_PINNED 1  /* host memory mapped for the GPU (hipHostMalloc) */
#define FMCS_MEM_MANAGED 2 /* managed / unified memory */
#define FMCS_MEM_HOST 3    /* pageable host memory (malloc): the GPU cannot read it */
#define FMCS_MEM_OTHER 4   /* another GPU's memory, or an array */

/* The classification itself, from what hipPointerGetAttributes reported:
 * query_status (its hipError_t; non-zero means the runtime does not know
 * the pointer, i.e. pageable host memory), memory_type (hipMemoryType),
 * owner_device (the attribute's device) and the context's device. */
uint32_t fmcs_classify_memory(int query_status, int memory_type, int owner_device, int ctx_device);
/* The kind of the memory p points into, for this context's device. */
int fmcs_pointer_kind(fmcs_ctx* ctx, const void* p, uint32_t* kind);
/* on != 0: stage pageable host buffers through device memory (off by
 * default).  Results equal those of the same call on device buffers. */
int fmcs_ctx_set_host_staging(fmcs_ctx* ctx, int on);

/* ---- host-side plan helpers ----------------------------------------- */

/* Bank bytes for a width: <=16 -> 2, <=32 -> 4, else 8; allow_u8 adds 1 for
 * <=8 (src/stitch_composer.cpp:622-650, hash path :883-890). */
uint32_t fmcs_bank_bytes(uint32_t width, int allow_u8);

/* Step a of StitchComposer::run_single_plan (src/stitch_composer.cpp:
 * 571-612): maps a column order and a split (new round widths) to the
 * construct entries, MSB first.  Returns the entry count in *n_out. */
int fmcs_construct_entries(uint32_t ncols, const uint32_t* bw, const uint32_t* col_order, uint32_t nsplit,
                           const uint32_t* split, fmcs_entry* entries, uint32_t max_entries, uint32_t* n_out);

/* ---- kernels -------------------------------------------------------- */

/* Step b2 of run_single_plan (src/stitch_composer.cpp:652-703):
 * SpecialiedColumn<T>::ShiftLeftBatch (src/column.h:105-180) and
 * AssembleFromColumn (src/column.h:208-1059), fused: every destination bank
 * is written once from zero.  src_d[i]: bank src_bank[i] holding
 * src_bw[i]-bit values; dst_d[c]: bank dst_bank[c].  Entries must be grouped
 * by newcol in ascending order, as fmcs_construct_entries emits them. */
int fmcs_stitch(fmcs_ctx* ctx, uint64_t n, uint32_t nsrc, const void* const* src_d, const uint32_t* src_bank,
                const uint32_t* src_bw, uint32_t nent, const fmcs_entry* entries, uint32_t ndst,
                void* const* dst_d, const uint32_t* dst_bank);

/* fmcs_stitch with a sort direction per source column (src_desc[i] != 0:
 * descending; NULL: all ascending).  A descending column's value bits are
 * inverted within its width while it is stitched, so sorting the new banks
 * ascending orders it from the largest value; ties keep the ascending oid
 * order.  The reference parses a direction but sorts ascending only
 * (src/composer.cpp:561-563). */
int fmcs_stitch_dir(fmcs_ctx* ctx, uint64_t n, uint32_t nsrc, const void* const* src_d, const uint32_t* src_bank,
                    const uint32_t* src_bw, uint32_t nent, const fmcs_entry* entries, uint32_t ndst,
                    void* const* dst_d, const uint32_t* dst_bank, const uint8_t* src_desc);

/* Composer::BATproject<T> (src/composer.cpp:608-628): out[i] = values[order[i]]. */
int fmcs_gather(fmcs_ctx* ctx, uint32_t bank, uint64_t n, const void* values_d, const uint32_t* order_d,
                void* out_d);

/* Composer::BATproject<T> (src/composer.cpp:608-628) for several payload
 * columns by one order, as the reference's tuple reconstruction projects
 * every output column by the same oids, with packed row elements as in its
 * disabled ChainComposer::BATproject_pack (src/composer.cpp:591-606):
 * outs_d[k][i] = cols_d[k][order_d[i]]
 * for k < ncols (1..32; banks[k] 2, 4 or 8 bytes).  Columns are grouped,
 * widest first, into row records of at most 16 bytes; a group of two or more
 * 16-byte-aligned columns is packed in one sequential pass and gathered with
 * one random record read a row, so a row's random reads touch one memory
 * line per group instead of one per column (the workspace holds n records of
 * 8 or 16 bytes).  Other columns take the column gather.  bits (NULL, or
 * each column's declared width, 1 .. 8 x bank) lets columns whose widths sum
 * to at most 64 share an 8-byte record at bit offsets, taken when that needs
 * no more records a row; a value wider than its declared bits then returns
 * FMCS_E_ARG before any output is written (one host synchronisation).
 * Every order_d[i] must be below n. */
int fmcs_gather_columns(fmcs_ctx* ctx, uint64_t n, uint32_t ncols, const void* const* cols_d,
                        const uint32_t* banks, const uint32_t* bits, const uint32_t* order_d,
                        void* const* outs_d);

/* Composer::extractGroupInfo<T> (src/composer.cpp:630-678): dense monotone
 * group ids of sorted values (and the previous groups, if in_group_d is not
 * NULL).  seg_offsets_d (optional, ngroups+1 entries, up to n+1) receives the
 * first row of every group and n.  *ngroups (host) receives the count. */
int fmcs_extract_group_info(fmcs_ctx* ctx, uint32_t bank, uint64_t n, const void* sorted_d,
                            const uint32_t* in_group_d, uint32_t* out_group_d, uint32_t* seg_offsets_d,
                            uint32_t* ngroups);

/* Mergesort::avxmergesort_anytype<T> (src/mergesort.h:62-64, :2878-3096)
 * as used by Composer::subsort (src/composer.cpp:490-589): sorts the pairs
 * (keys_d[i], oids_d[i]) by the keybits-bit key, in place.  Ties keep their
 * incoming order (LSD radix sort, stable).  A key with a bit at or above
 * keybits is refused: FMCS_E_ARG and nothing is sorted. */
int fmcs_sort_pairs(fmcs_ctx* ctx, uint32_t bank, uint32_t keybits, uint64_t n, void* keys_d, uint32_t* oids_d);

/* fmcs_sort_pairs without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, kernels and result as fmcs_sort_pairs, for
 * n <= FMCS_NOALLOC_MAX_ROWS (larger n: FMCS_E_ARG).  Its buffers live in a
 * static device array of FMCS_NOALLOC_SCRATCH_BYTES that each GPU holds
 * until the process ends, shared by every call on that GPU: calls that
 * overlap on one GPU (two streams, two threads) race and must not be issued.
 * Like fmcs_sort_pairs it synchronizes the context's stream once (the digit
 * histogram is read back to skip passes in which every key has the same
 * digit). */
#define FMCS_NOALLOC_MAX_ROWS (1u << 20)
#define FMCS_NOALLOC_SCRATCH_BYTES (16u << 20)
int fmcs_sort_pairs_noalloc(fmcs_ctx* ctx, uint32_t bank, uint32_t keybits, uint64_t n, void* keys_d,
                            uint32_t* oids_d);

/* The same sort by comparisons, the structure of avxmergesort_anytype
 * (src/mergesort.h:2878-3096: sorting networks over blocks, then a 2-way
 * merge tree): bitonic networks over 2048-row LDS tiles, then merge-path
 * passes.  Pairs are ordered by (key, oid), i.e. the stable order; the
 * result equals fmcs_sort_pairs' for keys within keybits. */
int fmcs_sort_pairs_merge(fmcs_ctx* ctx, uint32_t bank, uint64_t n, void* keys_d, uint32_t* oids_d);

/* Seeded synthetic column on the device, in place of the reference's
 * generator (src/generator.h:139-227, src/zipf.h:33-69; time-seeded there):
 * row i draws u = SplitMix64(seed, i); rank = u mod card, or with zipf
 * floor(card^(u/2^64)) - 1 (density ~ 1/(rank+1)); value = (rank * step) &
 * (2^width - 1), step = max(1, 2^width / card). */
int fmcs_fill_column(fmcs_ctx* ctx, uint32_t bank, uint64_t n, uint64_t seed, uint64_t card, int zipf,
                     uint32_t width, void* out_d);

/* The per-group loop of Composer::sortByGroup (src/composer.cpp:1280-1335):
 * sorts (keys, oids) within each segment [off[s], off[s+1]), in place, by
 * (key, oid).  mode: FMCS_MODE_*.  info (host, optional) receives the group
 * statistics the reference accumulates. */
int fmcs_segmented_sort(fmcs_ctx* ctx, uint32_t bank, uint32_t keybits, uint64_t n, void* keys_d,
                        uint32_t* oids_d, const uint32_t* seg_offsets_d, uint32_t nseg, int mode,
                        fmcs_group_info* info);

/* Composer::sortByGroup<T> (src/composer.cpp:1222-1408; declared
 * src/composer.h:85-88).  column_d: the round's bank indexed by oid.
 * order_d: in = order after the previous round (round 1: identity),
 * out = order after this round.  in_group_d: NULL in round 1.
 * out_group_d: written unless last_round.  sorted_d (optional): the round's
 * sorted values.  *ngroups (host, optional) receives the new group count
 * (0 on the last round). */
int fmcs_sort_by_group(fmcs_ctx* ctx, uint32_t bank, uint32_t keybits, uint64_t n, const void* column_d,
                       uint32_t* order_d, const uint32_t* in_group_d, uint32_t* out_group_d, int last_round,
                       int mode, void* sorted_d, fmcs_group_info* info, uint32_t* ngroups);

/* Composer::multiRoundSorting (src/composer.cpp:48-135) with its results
 * kept: order_d receives the final oid permutation; groups_d (optional)
 * receives the group ids after the last round (this also computes the last
 * round's groups, which the reference skips); *ngroups (optional) the final
 * group count (groups of the full key when groups_d is given, else of the
 * second-to-last round). */
int fmcs_multi_round_sort(fmcs_ctx* ctx, uint32_t nrounds, const void* const* cols_d, const uint32_t* banks,
                          const uint32_t* bitwidths, uint64_t n, uint32_t* order_d, uint32_t* groups_d, int mode,
                          uint32_t* ngroups);

/* Per-phase times of the last fmcs_multi_round_sort / fmcs_multi_round_hash
 * call in milliseconds, measured with HIP events on the context's stream:
 * [0] gather (tuple reconstruct), [1] first-round sort, [2] later-round
 * sorts, [3] group extraction.  Mirrors setting_t's timing fields
 * (src/types.h:206-221). */
int fmcs_last_breakdown(fmcs_ctx* ctx, double* ms4);

/* Per-kernel timing with HIP events on the context's stream (no
 * counterpart in the reference, whose HybridTimer brackets whole phases,
 * src/hybrid_timer.h:16-57).  When on, every major kernel launch is
 * bracketed by two events; fmcs_kernel_stats writes a JSON object
 * {kernel: {"launches": L, "ms": total, "bytes": algorithmic bytes}}. */
int fmcs_ctx_set_profiling(fmcs_ctx* ctx, int on);
int fmcs_kernel_stats(fmcs_ctx* ctx, char* buf, uint64_t cap);
int fmcs_kernel_stats_reset(fmcs_ctx* ctx);

/* ---- hash grouping (src/composer.cpp:137-227, :814-976;
 *      src/hashing.h:289-385; src/partition.h:44-161) ------------------ */

/* One round of hashByGroup_reordered: within each segment of the previous
 * round's groups (in_group_d NULL in round 1) give dense ids in order of
 * first appearance (Hashing::subgroup_range), numbered across segments in
 * segment order.  gid_d receives the ids per row; unless last_round the
 * rows are clustered stably by id (Partition::histogram_based_reordering):
 * out_order_d and out_group_d receive the clustered oids and ids.
 * values_d is the round's bank indexed by oid; order_d the current order
 * (in round 1, in_group_d NULL, values are read in row order: the order is
 * the identity there, as in the reference). */
int fmcs_hash_round(fmcs_ctx* ctx, uint32_t bank, uint64_t n, const void* values_d, const uint32_t* order_d,
                    const uint32_t* in_group_d, uint32_t* gid_d, uint32_t* out_order_d, uint32_t* out_group_d,
                    int last_round, uint32_t* ngroups);

/* Composer::multiRoundHashing_reordered / _non_reordered
 * (src/composer.cpp:137-227, :321-381; chosen by compose_params_t.hash_type,
 * src/chain_composer.cpp:16-25) with the results kept.  hash_type 1 or 2:
 * order_d = order the last round hashed in, gid_d = the last round's ids
 * per row of that order.  hash_type 0: no reordering; gid_d = first-
 * appearance ids of the full tuple per original row, order_d = identity.
 * *ngroups = number of groups of the full tuple. */
int fmcs_multi_round_hash(fmcs_ctx* ctx, uint32_t nrounds, const void* const* cols_d, const uint32_t* banks,
                          uint64_t n, int hash_type, uint32_t* order_d, uint32_t* gid_d, uint32_t* ngroups);

/* Column::GetCardinality (src/column.h:18, :61, :81-88): the number of
 * distinct values of a column, which the planner's card_per_bit statistics
 * start from.  The reference stores it with the column when generating the
 * data; here it is counted exactly from the data on the device (hash
 * grouping, bank 1, 2, 4 or 8 bytes). */
int fmcs_distinct_count(fmcs_ctx* ctx, uint32_t bank, uint64_t n, const void* col_d, uint64_t* count);

/* Measurement probe with no counterpart in the reference: the cost of a
 * hand-off between workgroups inside one launch, the step a single-launch
 * work queue pays whenever one workgroup's output becomes another's input.
 * Each of `grid` blocks copies tasks of 4096 u64 rows (tasks in all, at most
 * 2^17); ms4[m] is the best of 5 launches of mode m: 0 no release, 1 every
 * wave releases its stores at agent scope after each task, 2 one release
 * after a barrier per task, 3 one release per 8 tasks of a block.  Uses
 * three of the context's scratch buffers (2 x tasks x 32 KiB). */
int fmcs_probe_handoff(fmcs_ctx* ctx, uint32_t tasks, uint32_t grid, double* ms4);

#ifdef __cplusplus
}
#endif

#endif /* FMCS_H */
