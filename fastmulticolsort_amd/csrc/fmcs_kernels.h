// fmcs_kernels.h -- kernel declarations and argument blocks.
#pragma once
#include "fmcs_common.h"

#include <type_traits>

namespace fmcs {

constexpr int STITCH_MAX_SRC = 16;
constexpr int STITCH_MAX_DST = 16;
constexpr int STITCH_MAX_ENT = 64;

// Kernel-argument block of the fused stitch kernel (passed by value; lives
// in the kernarg segment, read with scalar loads).
struct StitchArgs {
  const void* src[STITCH_MAX_SRC];
  void* dst[STITCH_MAX_DST];
  u32 src_bank[STITCH_MAX_SRC];
  u32 dst_bank[STITCH_MAX_DST];
  u32 ent_begin[STITCH_MAX_DST + 1];
  uint8_t e_src[STITCH_MAX_ENT];
  uint8_t e_shift[STITCH_MAX_ENT];  // bw_src - end
  uint8_t e_width[STITCH_MAX_ENT];  // end - start + 1
  u32 ndst;
  u32 invert;  // bit s: source s sorts descending (its value bits are inverted)
  u64 n;
};

template <class K>
struct TileSortArgs {
  const K* kin[2];   // source banks by level parity
  const u32* vin[2];
  K* kout;           // may be null (keys not needed)
  u32* vout;
  const u32* tile_lo;
  const u32* tile_hi;
  const u32* tile_s0;
  const u32* tile_s1;
  const u32* seg_start;
  u32 src_sel;
  u32 kb;       // key bits still to sort (low bits of the key)
  u32 pad_seg;  // sentinel segment index, above any real one
  u32 direct;   // 1: one tile given by d_lo/d_hi (plain per-segment form)
  u32 d_lo, d_hi;
  const u32* only;   // if set, only tiles w with only[w] != 0 are sorted
  const u32* wlist;  // if set, block b sorts window wlist[b]
  const u32* wcount;  // if set: persistent loop over wlist[0 .. *wcount)
  u32* fblist;        // bucket tiles refused -> appended here (with fbcount)
  u32* fbcount;       // radix tile sort: persistent loop over fblist[0 .. *fbcount)
  u32 size_lo, size_hi;  // bucket sort: only windows of size_lo + 1 .. size_hi rows (size_hi 0: all)
  u32* rklist;        // radix tile sort: windows too wide to pack -> appended here
  u32* rkcount;       //   (with rkcount), for the comparison rank
  u64 bad;            // bucket sort, if err is set: key bits above the declared width;
  u32* err;           //   a window holding such a key sets *err
};

constexpr u32 XCDS = 8;  // accelerator dies of an MI355X; workgroup b runs on die b % XCDS
constexpr int ONESWEEP_ITEMS = 16;
constexpr int ONESWEEP_TILE = 256 * ONESWEEP_ITEMS;
// Taller tiles for the u64 composite passes of the packed first round and the
// segmented MSD levels: 20 rows a thread, 5120-row tiles, so each tile's 256
// digit runs are 20 rows long on average (fewer partial lines at run ends).
// Measured on config 3: composite passes 1.683 -> 1.625 ms a step, MSD
// scatter 1.215 -> 1.164; the fused first pass is faster at 16 (1.084
// against 1.198); 12 and 24 rows were slower everywhere.
constexpr int WIDE_ITEMS = 20;
constexpr int WIDE_TILE = 256 * WIDE_ITEMS;
// Digit split of an LSD sort over keybits bits in npasses = ceil(keybits / RB)
// passes: widths as even as possible, the narrower digits first.  A pass's
// time grows steeply with its number of digit values (each tile's output
// runs get shorter than a few cache lines), so 20 bits sort as 6 + 7 + 7
// rather than 8 + 8 + 4.  Pass p sorts bits digit_shift .. + digit_width - 1.
__host__ __device__ inline u32 digit_width(u32 keybits, u32 npasses, u32 p) {
  return keybits / npasses + (p >= npasses - keybits % npasses ? 1u : 0u);
}
__host__ __device__ inline u32 digit_shift(u32 keybits, u32 npasses, u32 p) {
  const u32 narrow = npasses - keybits % npasses;
  return p * (keybits / npasses) + (p > narrow ? p - narrow : 0u);
}
constexpr int BOUNDARY_ITEMS = 32;
constexpr int BOUNDARY_TILE = 256 * BOUNDARY_ITEMS;
constexpr int TILESORT_E = 4;
constexpr u32 BUCKET_LIMIT = 24;  // largest bin the bucket tile sort takes (else the radix tile sort)
constexpr int TILESORT_T = 256 * TILESORT_E;
constexpr int TILESORT_TMAX = 2048;  // largest tile of the device-driven path (256 threads x 8 rows)
constexpr int SCAN_ITEMS = 8;
constexpr u32 TSCAN_CT = 32;  // tiles per chunk of the per-segment tile-count scan
constexpr u32 MERGE_TILE = 2048;  // rows per bitonic tile / per merge-pass block
constexpr int SCAN_TILE = 256 * SCAN_ITEMS;

__global__ void k_stitch(StitchArgs a);
template <class T>
__global__ void k_gather(const T*, const u32*, T*, u64);
template <class T>
__global__ void k_gather_win(const T*, const u32*, T*, u64, u32, u32);
// Up to eight payload columns of 2, 4 or 8 bytes packed side by side into
// R-byte row records (R = 8 or 16; each column at an offset aligned to its
// width), so that a gather by an order touches one memory line a row
// instead of one per column.
constexpr int REC_ROWS = 4;  // rows a thread of the record gather (a multiple of 4)
struct RecCols {
  const void* src[8];
  void* dst[8];
  u32 width[8];
  u32 off[8];   // byte offset (byte records) or bit offset (bit records)
  u32 bits[8];  // bit records: the column's declared width
  u32 ncol;
};
template <int R>
__global__ void k_rec_pack(RecCols, void*, u64);
template <int R>
__global__ void k_rec_gather(RecCols, const void*, const u32*, u64);
// Bit records: the columns' values at bit offsets of one u64 (their
// declared widths sum to at most 64); a value wider than declared sets *err.
__global__ void k_bitrec_pack(RecCols, u64*, u64, u32*);
__global__ void k_bitrec_gather(RecCols, const u64*, const u32*, u64);
template <class K, int ITEMS>
__global__ void k_bcount(const K*, const u32*, u64, u32, u32*, u32*);
template <class K, int ITEMS>
__global__ void k_bwrite(const K*, const u32*, u64, u32, const u32*, u32*, u32*, u32*, const u32*);
template <class K>
__global__ void k_hist_all(const K*, u64, u32, u32, u64, u32*);
__global__ void k_digit_base(const u32*, const u32*, u32*, u32*, u32*);
// P = NoPay: no payload column
struct NoPay {};
// payload output of a scatter pass: the payload, or the u64 composite
// (key << fshift) | payload when fused
template <class P, bool FUSE>
using PayOut = typename std::conditional<FUSE, u64, P>::type;
template <class K, int ITEMS, int RB, class P, bool FUSE, bool LSD, class KO = K>
__global__ void k_onesweep(const K*, const u32*, KO*, u32*, const u32*, const u32*, const u32*, u32, const u32*,
                           const u32*, const u32*, u32, u32, const u32*, const u32*, const P*, PayOut<P, FUSE>*, u32,
                           u64, u32*, u32*, uint8_t*, u32, const u32*);
template <class K, int E>
__global__ void k_tile_sort(TileSortArgs<K>);
// the window descriptor arrays (tile_lo, tile_hi, tile_s0, tile_s1, wlist of
// the argument block) are passed again as restrict pointers
template <class K, int E>
__global__ void k_tile_bucket(TileSortArgs<K>, u32*, const u32*, const u32*, const u32*, const u32*, const u32*);
template <class K, int E>
__global__ void k_tile_rank(TileSortArgs<K>);
__global__ void k_windows(const u32*, const u32*, u32, u32, u32, const u32*, const u32*, u32, u32*, u32*, u32*, u32*,
                          u32*, u32*);
__global__ void k_offsets_to_list(const u32*, u32, u32*, u32*);
__global__ void k_large_flags(const u32*, u32, u32, u32*, u32*);
__global__ void k_large_compact(const u32*, const u32*, u32, u32, const u32*, u32, u32*, u32*, u32*);
template <class K>
__global__ void k_copy_segs(const K*, const u32*, K*, u32*, const u32*, const u32*, const u32*, u32, u32);
template <int ITEMS, bool HEADS>
__global__ void k_scan_excl(const u32*, u32*, u64, u32*, u32*, u32*);
template <bool HEADS>
__global__ void k_scan_reduce(const u32*, u64, u64, u32*);
__global__ void k_scan_sums(u32*, u32, u32*);
template <bool HEADS>
__global__ void k_scan_rescan(const u32*, u32*, u64, u64, const u32*);
__global__ void k_fill_u32(u32*, u64, u32);
__global__ void k_iota_u32(u32*, u64);
template <class K>
__global__ void k_composite(const K*, const u32*, u64, u32, u64*, u32*);
template <class K>
__global__ void k_key_check(const K*, u64, u64, u32*);
template <class K>
__global__ void k_strip_composite(const u64*, u64, u64, K*);
template <class K>
__global__ void k_hash_insert(const K*, const u32*, u64, u32*, u32, u32*, u32*);
__global__ void k_hash_heads(const u32*, const u32*, u64, u32*);
// hash grouping of composites that fit 64 bits (fmcs_hash.hip)
template <class K>
__global__ void k_hash_prep(const K*, const u32*, u64, u32, u64*, u32*);
template <class K>
__global__ void k_hash_local(const K*, const u32*, u64, u32, u64*, u32*, u64, u32*, u32*, u32, u32, u32*);
template <class K>
__global__ void k_hash_local_first(const K*, const u32*, u64, u32, const u64*, const u32*, u64, const u32*, u32*);
__global__ void k_hash_pstart(const u64*, u32, u32, u32*);
__global__ void k_hash_window(const u64*, const u32*, const u32*, u32, u32, u32*, u32*, u32*);
__global__ void k_hash_large(const u64*, const u32*, const u32*, const u32*, u32, u64*, u32*, u64, u32*, u32*);
__global__ void k_hash_large_first(const u64*, const u32*, const u32*, const u32*, u32, const u64*, const u32*, u64,
                                   const u32*, u32*);
__global__ void k_hash_assign_f(const u32*, const u32*, u64, u32*);
__global__ void k_hash_rank_slots(const u32*, const u32*, const u32*, u64, u32*);
__global__ void k_hash_assign(const u32*, const u32*, u64, u32*);
template <class K>
__global__ void k_fill_random(K*, u64, u64, u32);
template <class T>
__global__ void k_fill_column(T*, u64, u64, u64, u32, double, u64, u64);
__global__ void k_uniform_offsets(u32*, u32, u32, u64);


// Comparison sort (fmcs_merge.hip).
template <class K>
__global__ void k_bitonic_tile(const K*, const u32*, K*, u32*, u64);
template <class K>
__global__ void k_merge_pass(const K*, const u32*, K*, u32*, u64, u64, const u32*);
template <class K>
__global__ void k_merge_splits(const K*, const u32*, u64, u64, u32, u32*);

// Device-driven list kernels (fmcs_dd.hip): sizes come from device words.
__global__ void kd_set4(u32*, u32, u32, u32, u32);
__global__ void kd_offsets_to_list(const u32*, const u32*, u32*, u32*);
__global__ void kd_windows(const u32*, const u32*, const u32*, u32, u32, const u32*, const u32*, u32, u32*, u32*,
                           u32*, u32*, u32*, u32*, u32, u32*, u32*, u32*, u32*, u32*, u32*);
template <int ITEMS, bool FLAGS>
__global__ void kd_scan_excl(const u32*, u32*, const u32*, u32*, u32*, u32*, u32, u32*, const u32*, u32, u32*, u32*,
                             u32*, u32*);
template <class K, int ITEMS, int RB>
__global__ void kd_seg_hist(const K*, const u32*, const u32*, const u32*, const u32*, const u32*, u32, u32, u32*,
                            u32*, u64, u32*, u32*);
template <int RB>
__global__ void kd_tscan_local(u32*, const u32*, const u32*, const u32*, u32*, u32*, const u32*);
template <int RB>
__global__ void kd_tscan_carry(u32*, const u32*, const u32*);
template <int RB>
__global__ void kd_digit_base(const u32*, const u32*, u32*, u32*, u32*, const u32*, u32*);
template <class K>
__global__ void kd_copy_segs(const K*, const u32*, K*, u32*, const u32*, const u32*, const u32*, const u32*,
                             const u32*, u32);
__global__ void kd_copy1(const u32*, u32*);
template <class K>
__global__ void kd_copy_pairs(const K*, const u32*, K*, u32*, u64);
template <class T>
__global__ void kd_copy_col(const T*, T*, u64);

// Level records of the device-driven segmented sort: DD_LVL_WORDS words per
// level, these at their places.
constexpr u32 DD_LVL_WORDS = 16;
enum { LV_NSEG, LV_WCOUNT, LV_ROWS_SMALL, LV_ROWS_LARGE, LV_NLARGE, LV_NTILES, LV_SCAN_A, LV_SCAN_B, LV_OS, LV_FB,
       LV_NSINGLE, LV_FB2, LV_RK, LV_RK2 };

// The persistent size-binned launch (fmcs_bins.hip).  After the radix levels
// have split every segment above med_max rows, one launch sorts the work that
// every level left behind: single segments of small_max + 1 .. med_max rows
// (those above one window tile sorted by the block alone: an MSD level of
// its own plus its child windows) and the windows of packed small segments.
// Window tiles are BIN_T rows.
constexpr int BIN_E = 4;
constexpr u32 BIN_T = 256 * BIN_E;
constexpr u32 BIN_LVL_MAX = 12;
constexpr u32 BIN_MED_MAX = 16384;   // largest medium segment (bounds the block's stack)
constexpr u32 BIN_STACK = 128;       // segments above BIN_T rows pending in one block
constexpr u32 BIN_SCRATCH = 2048;    // global words per block: child starts, stack, window list
template <class K>
struct BinArgs {
  const K* kin[2];   // rows of a level by its parity (0: the caller's buffers)
  const u32* vin[2];
  K* kbuf[2];        // the same buffers, written by a medium segment's MSD level
  u32* vbuf[2];
  K* kout;           // final rows (kout null: sorted keys not wanted)
  u32* vout;
  const u32* starts;  // every level's segment starts, at loff[level]
  const u32* tlo;     // window slots, at woff[level]
  const u32* thi;
  const u32* ts0;
  const u32* ts1;
  const u32* wlist;   // compact window ids, at woff[level]
  const u32* slo;     // single segments, at soff[level]
  const u32* shi;
  const u32* sseg;
  const u32* lv;      // level records (window and single counts)
  u32* scratch;       // BIN_SCRATCH words per block
  u32* err;           // level-0 rows: a key with a bit in bad sets *err |= 1; a full stack *err |= 2
  u64 bad;
  // windows the bucket sort refuses, by source buffer: list b at b * fbcap of
  // fblo/fbhi/fbs0/fbs1, counted in fbcount[b]; their segment starts in
  // fbpool (row-indexed, n words)
  u32* fbpool;
  u32* fblo;
  u32* fbhi;
  u32* fbs0;
  u32* fbs1;
  u32* fbcount;
  u32 fbcap;
  u32 nlev, W, small_max;
  u32 loff[BIN_LVL_MAX], woff[BIN_LVL_MAX], soff[BIN_LVL_MAX], kbs[BIN_LVL_MAX];
};
template <class K>
__global__ void k_seg_bins(BinArgs<K>);
constexpr u32 PROBE_ROWS = 4096;  // u64 rows a task of k_handoff_probe (32 KiB)
template <int MODE>
__global__ void k_handoff_probe(const u64*, u64*, u32, u32*);

}  // namespace fmcs
