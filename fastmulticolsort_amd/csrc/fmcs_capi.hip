// fmcs_capi.hip -- context, workspace and the C-ABI entry points (fmcs.h).
//
// Host orchestration of the kernels in fmcs_kernels.hip.  All launches go to
// the context's stream; temporary device memory comes from a per-context
// workspace that grows on demand and is reused across calls.
#include <algorithm>
#include <cmath>
#include <map>
#include <tuple>
#include <utility>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fmcs.h"
#include "fmcs_internal.h"
#include "fmcs_kernels.h"

using namespace fmcs;

namespace {

enum Slot {
  WS_KA, WS_KB, WS_VB, WS_T0K, WS_T0V, WS_T1K, WS_T1V, WS_STATUS, WS_HIST, WS_DBASE, WS_SEGOFF, WS_SEGOFF2,
  WS_ST, WS_SZ, WS_BST, WS_BSZ, WS_FLAG, WS_POS, WS_LLO, WS_LHI, WS_LTILES, WS_TBASE, WS_TLO, WS_THI, WS_TS0,
  WS_TS1, WS_COMP, WS_COMP2, WS_GROUPTMP, WS_MISC, WS_HASH, WS_HASH2, WS_HASH3, WS_HT_TABLE, WS_HT_FIRST,
  WS_HT_RANK, WS_HT_SLOT, WS_HT_HEAD, WS_HT_PREFIX, WS_FALLBACK, WS_WLIST, WS_DD, WS_SCANST, WS_FBLIST, WS_TCOUNT, WS_CARRY, WS_BCNT, WS_T0P, WS_T1P, WS_PAY, WS_SGL, WS_FBLIST2, WS_RKL, WS_RKL2, WS_DNEXT, WS_TSEG, WS_HG_KEY, WS_HG_POS, WS_HG_SKEY, WS_HG_SPOS, WS_HG_F, WS_HG_AUX, WS_HG_PSTART,
  WS_HG_TKEY, WS_HG_TFIRST, WS_HG_CHUNK, WS_PLIST, WS_BINSC, WS_FBWIN, WS_FBPOOL, WS_UNB0, WS_UNB1, WS_UNBK, WS_DIG0, WS_BBITS, WS_REC, WS_RECERR, WS_USER0,
  WS_STAGE0 = WS_USER0 + 24,  // device copies of staged host buffers (one slot per argument)
  WS_COUNT = WS_STAGE0 + 72
};

struct Buf {
  void* p = nullptr;
  size_t bytes = 0;
};

}  // namespace

constexpr u32 PIN_WORDS = 4096;

struct fmcs_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  Buf ws[WS_COUNT];
  u32* small_d = nullptr;   // 64 words of device scratch (counters, totals, consts)
  u32* host_pin = nullptr;  // PIN_WORDS words pinned for readbacks
  fmcs_tuning tune = {512, 512, 8, 0, 1, 0};
  hipEvent_t ev[8] = {};
  double breakdown[4] = {0, 0, 0, 0};
  // per-kernel event timing (off unless enabled)
  struct Rec {
    const char* name;
    hipEvent_t e0, e1;
    double bytes;
    const u32* dcount;  // if set: bytes += per_unit * (*dcount), read when the records are flushed
    double per_unit;
    double extra;       // side outputs a later kernel of this design reads (not the pass's own bytes)
  };
  struct Agg {
    u64 launches = 0;
    double ms = 0, bytes = 0, extra = 0;
  };
  // extra bytes of the next profiled launch (set just before it)
  double extra_next = 0;
  bool prof = false;
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  std::map<std::string, Agg> agg;
  // device-driven path: workspace reallocation epoch (a captured graph is
  // valid only while no buffer it names moved), capture state, cached graph
  u64 ws_epoch = 0;
  bool capturing = false;
  hipStream_t cap_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<u64> gkey;
  u32 ncu = 0;
  std::map<const void*, u32> resident;
  bool dd_pending = false;  // a device-driven call's key-error word is not yet checked
  // Arena mode (the allocation-free entry point): while arena is set, every
  // workspace slot is carved out of [arena, arena + arena_cap) for the
  // current call instead of being allocated; a slot asked for again with no
  // more bytes gets the same memory.
  char* arena = nullptr;
  size_t arena_cap = 0, arena_used = 0;
  size_t arena_off[WS_COUNT] = {}, arena_len[WS_COUNT] = {};
  // this device's copy of the static scratch array (fmcs_noalloc.hip), or
  // null with the reason in scratch_err
  void* scratch_ptr = nullptr;
  std::string scratch_err;
  // stage pageable host buffers through device memory (else they are
  // rejected with FMCS_E_ARG)
  bool host_staging = false;
};

#define HIPCHK(x)                                                                          \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) return fail(ctx, FMCS_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define HIPCHK_C(c, x)                                                                      \
  do {                                                                                      \
    hipError_t e_ = (x);                                                                    \
    if (e_ != hipSuccess) return fail(c, FMCS_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define TRY(x)              \
  do {                      \
    int rc_ = (x);          \
    if (rc_ != FMCS_OK) return rc_; \
  } while (0)

namespace {

int fail(fmcs_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

int getbuf(fmcs_ctx* ctx, Slot s, size_t bytes, void** out) {
  bytes = std::max<size_t>(bytes, 256);
  if (ctx->arena) {
    if (ctx->arena_len[s] < bytes) {
      const size_t off = (ctx->arena_used + 255) & ~(size_t)255;
      if (off > ctx->arena_cap || bytes > ctx->arena_cap - off)
        return fail(ctx, FMCS_E_ARG, "scratch of " + std::to_string(ctx->arena_cap) + " bytes exhausted");
      ctx->arena_off[s] = off;
      ctx->arena_len[s] = bytes;
      ctx->arena_used = off + bytes;
    }
    *out = ctx->arena + ctx->arena_off[s];
    return FMCS_OK;
  }
  Buf& b = ctx->ws[s];
  if (b.bytes < bytes) {
    if (ctx->capturing) return fail(ctx, FMCS_E_HIP, "workspace growth during graph capture");
    ctx->ws_epoch++;
    if (b.p) {
      hipError_t e = hipStreamSynchronize(ctx->stream);
      if (e != hipSuccess) return fail(ctx, FMCS_E_HIP, hipGetErrorString(e));
      (void)hipFree(b.p);
      b.p = nullptr;
      b.bytes = 0;
    }
    size_t want = bytes + bytes / 8;
    if (hipMalloc(&b.p, want) != hipSuccess) {
      (void)hipGetLastError();
      b.p = nullptr;
      return fail(ctx, FMCS_E_NOMEM, "workspace allocation of " + std::to_string(want) + " bytes failed");
    }
    b.bytes = want;
  }
  *out = b.p;
  return FMCS_OK;
}

template <class T>
int getbuf_t(fmcs_ctx* ctx, Slot s, size_t count, T** out) {
  void* p;
  TRY(getbuf(ctx, s, count * sizeof(T), &p));
  *out = (T*)p;
  return FMCS_OK;
}

int launch_check(fmcs_ctx* ctx, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(ctx, FMCS_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return FMCS_OK;
}

// Blocks of `fn` resident on the whole device at 256 threads (cached).
u32 resident(fmcs_ctx* ctx, const void* fn) {
  auto it = ctx->resident.find(fn);
  if (it != ctx->resident.end()) return it->second;
  if (ctx->ncu == 0) {
    hipDeviceProp_t prop;
    ctx->ncu = (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess && prop.multiProcessorCount > 0)
                   ? (u32)prop.multiProcessorCount
                   : 256u;
  }
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, 256, 0) != hipSuccess || nb <= 0) {
    (void)hipGetLastError();
    nb = 1;
  }
  u32 r = (u32)nb * ctx->ncu;
  ctx->resident[fn] = r;
  return r;
}

u64 cap_grid(u64 want, u64 cap) { return std::max<u64>(1, std::min<u64>(want, cap)); }

// Grid of a persistent kernel (blocks loop over work items taken from an
// atomic counter): twice the resident blocks, so a block that finishes early
// is replaced at once and the CUs stay full; extra blocks find no work.
u64 persist_grid(fmcs_ctx* ctx, const void* fn, u64 want) { return cap_grid(want, 2ull * resident(ctx, fn)); }
// Grid of the bucket tile sort's window loops: four times the resident
// blocks (windows differ in cost; the later blocks even out the end of the
// launch: 1.129 -> 1.112 ms a step against twice)
u64 bucket_grid(fmcs_ctx* ctx, const void* fn, u64 want) {
  const u64 g = cap_grid(want, 4ull * resident(ctx, fn));
  return (g + XCDS - 1) / XCDS * XCDS;
}

// Grid of k_onesweep: persistent, a multiple of XCDS (its XCD-aware tile walk).
u64 onesweep_grid(fmcs_ctx* ctx, const void* fn, u64 ntiles) {
  const u64 g = std::min<u64>(std::max<u64>(ntiles, 1), resident(ctx, fn));
  return (g + XCDS - 1) / XCDS * XCDS;
}

int fill(fmcs_ctx* ctx, u32* p, u64 n, u32 v) {
  if (n == 0) return FMCS_OK;
  u32 blocks = (u32)std::min<u64>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, ctx->stream, p, n, v);
  return launch_check(ctx, "k_fill_u32");
}

int readback(fmcs_ctx* ctx, const u32* d, u32 count, u32* h) {
  if (count > PIN_WORDS) return fail(ctx, FMCS_E_ARG, "readback too large");
  hipError_t e = hipMemcpyAsync(ctx->host_pin, d, count * sizeof(u32), hipMemcpyDeviceToHost, ctx->stream);
  if (e != hipSuccess) return fail(ctx, FMCS_E_HIP, hipGetErrorString(e));
  e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return fail(ctx, FMCS_E_HIP, hipGetErrorString(e));
  std::memcpy(h, ctx->host_pin, count * sizeof(u32));
  return FMCS_OK;
}

int upload(fmcs_ctx* ctx, u32* d, const u32* h, u32 count) {
  // small host->device constant upload through the pinned buffer
  if (count > 64) return fail(ctx, FMCS_E_ARG, "upload too large");
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return fail(ctx, FMCS_E_HIP, hipGetErrorString(e));
  std::memcpy(ctx->host_pin, h, count * sizeof(u32));
  e = hipMemcpyAsync(d, ctx->host_pin, count * sizeof(u32), hipMemcpyHostToDevice, ctx->stream);
  if (e != hipSuccess) return fail(ctx, FMCS_E_HIP, hipGetErrorString(e));
  e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return fail(ctx, FMCS_E_HIP, hipGetErrorString(e));
  return FMCS_OK;
}

bool bank_ok(u32 b) { return b == 2 || b == 4 || b == 8; }

hipEvent_t pool_get(fmcs_ctx* ctx) {
  if (!ctx->pool.empty()) {
    hipEvent_t e = ctx->pool.back();
    ctx->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

int prof_flush(fmcs_ctx* ctx) {
  if (ctx->recs.empty()) return FMCS_OK;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (auto& r : ctx->recs) {
    float ms = 0;
    if (r.e0 && r.e1 && hipEventElapsedTime(&ms, r.e0, r.e1) == hipSuccess) {
      double bytes = r.bytes;
      if (r.dcount) {
        u32 c = 0;
        HIPCHK(hipMemcpy(&c, r.dcount, sizeof(u32), hipMemcpyDeviceToHost));
        bytes += r.per_unit * c;
      }
      auto& a = ctx->agg[r.name];
      a.launches++;
      a.ms += ms;
      a.bytes += bytes;
      a.extra += r.extra;
    }
    if (r.e0) ctx->pool.push_back(r.e0);
    if (r.e1) ctx->pool.push_back(r.e1);
  }
  ctx->recs.clear();
  return FMCS_OK;
}

// Brackets one launch with events on the context's stream when profiling.
struct Prof {
  fmcs_ctx* ctx;
  int idx = -1;
  Prof(fmcs_ctx* c, const char* name, double bytes, const u32* dcount = nullptr, double per_unit = 0) : ctx(c) {
    const double extra = ctx->extra_next;
    ctx->extra_next = 0;
    if (!ctx->prof || ctx->capturing || ctx->arena) return;
    if (ctx->recs.size() > 8192) prof_flush(ctx);
    fmcs_ctx::Rec r{name, pool_get(ctx), pool_get(ctx), bytes, dcount, per_unit, extra};
    if (r.e0) (void)hipEventRecord(r.e0, ctx->stream);
    ctx->recs.push_back(r);
    idx = (int)ctx->recs.size() - 1;
  }
  void set_bytes(double b) {
    if (idx >= 0) ctx->recs[idx].bytes = b;
  }
  ~Prof() {
    if (idx >= 0 && ctx->recs[idx].e1) (void)hipEventRecord(ctx->recs[idx].e1, ctx->stream);
  }
};

template <typename... KA, size_t... I>
hipError_t launch_tuple(const void* fn, u32 grid, hipStream_t st, std::tuple<KA...>& t, std::index_sequence<I...>) {
  void* argv[sizeof...(KA) + 1] = {(void*)&std::get<I>(t)...};
  return hipLaunchKernel(fn, dim3(grid), dim3(256), argv, 0, st);
}

// Launches kern<<<grid, 256>>> on the context's stream with per-kernel
// timing when profiling; bytes = static bytes + per_unit * (*dcount).
template <typename... KA, typename... A>
int klaunch(fmcs_ctx* ctx, const char* name, void (*kern)(KA...), u64 grid, double bytes, const u32* dcount,
            double per_unit, A... args) {
  if (grid == 0) return FMCS_OK;
  std::tuple<KA...> t(args...);
  Prof pf(ctx, name, bytes, dcount, per_unit);
  hipError_t e = launch_tuple((const void*)kern, (u32)grid, ctx->stream, t, std::index_sequence_for<KA...>{});
  if (e != hipSuccess) return fail(ctx, FMCS_E_HIP, std::string(name) + ": " + hipGetErrorString(e));
  return launch_check(ctx, name);
}

// Segmented tile-offset scan of the per-tile digit counts (written by
// kd_seg_hist into tcount over the same tiling) and the scatter pass.
// A payload column (pin -> pout) moves with the pairs unless P is NoPay;
// FUSE writes (key << fshift) | payload composites to pout instead of keys.
// kside (no payload): also write every key >> fshift as a u32.  LSD: the
// pass covers one segment (the kernel skips the tile -> segment search).
template <class K, int RB = RADIX_BITS, class P = NoPay, bool FUSE = false, bool LSD = false,
          int ITEMS = ONESWEEP_ITEMS, class KO = K>
int scatter_pass(fmcs_ctx* ctx, const char* name, const K* kin, const u32* vin, KO* kout, u32* vout, const u32* lo,
                 const u32* hi, const u32* tbase, const u32* nsegs_d, const u32* ntiles_d, u64 ntiles_max,
                 const u32* dbase, u32* tcount, u32 shift, u32 dbits, double bytes, const u32* dcount, double per,
                 const P* pin = nullptr, PayOut<P, FUSE>* pout = nullptr, u32 fshift = 0, u64 pbad = 0,
                 u32* err = nullptr, u32* kside = nullptr, uint8_t* dnext = nullptr, u32 nshift = 0,
                 const u32* tseg = nullptr) {
  constexpr u32 R = 1u << RB;
  const double extra = ctx->extra_next;  // belongs to the scatter launch, not the scans before it
  ctx->extra_next = 0;
  const u64 nch = std::max<u64>(1, (ntiles_max + TSCAN_CT - 1) / TSCAN_CT);
  u32* carry;
  TRY(getbuf_t(ctx, WS_CARRY, nch * (R + 1), &carry));
  u32* cflag = carry + nch * R;
  TRY(klaunch(ctx, "k_tile_scan", kd_tscan_local<RB>, cap_grid(nch, 2048), 0.0, nullptr, 0.0, tcount, tbase, nsegs_d,
              ntiles_d, carry, cflag, tseg));
  TRY(klaunch(ctx, "k_tile_scan", kd_tscan_carry<RB>, R, 0.0, nullptr, 0.0, carry, (const u32*)cflag, ntiles_d));
  auto kern = k_onesweep<K, ITEMS, RB, P, FUSE, LSD, KO>;
  ctx->extra_next = extra;
  return klaunch(ctx, name, kern, onesweep_grid(ctx, (const void*)kern, ntiles_max), bytes, dcount, per, kin, vin,
                 kout, vout, lo, hi, tbase, 0u, dbase, (const u32*)tcount, (const u32*)carry, shift, dbits, nsegs_d,
                 ntiles_d, pin, pout, fshift, pbad, err, kside, dnext, nshift, tseg);
}

// exclusive scan of n u32 values -> out, the sum to *total_d (device) if
// non-null; heads: scan the head flags (in[i] == i) of a first-position
// array instead of its values
int scan_excl(fmcs_ctx* ctx, const u32* in, u32* out, u64 n, u32* total_d, bool heads = false) {
  if (n == 0) return FMCS_OK;
  u64 ntiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  u32* status;
  TRY(getbuf_t(ctx, WS_STATUS, ntiles, &status));
  if (n >= (1ull << 21)) {
    // long arrays: reduce-then-scan over about 1024 chunks of C rows
    const u64 C = ((n + 1023) / 1024 + 1023) / 1024 * 1024;
    const u64 G = (n + C - 1) / C;  // <= 1024 <= ntiles
    u32* sums = status;
    {
      Prof pf(ctx, "k_scan_reduce", 4.0 * n);
      hipLaunchKernelGGL((heads ? k_scan_reduce<true> : k_scan_reduce<false>), dim3((u32)G), dim3(256), 0,
                         ctx->stream, in, n, C, sums);
      TRY(launch_check(ctx, "k_scan_reduce"));
    }
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(256), 0, ctx->stream, sums, (u32)G, total_d);
    TRY(launch_check(ctx, "k_scan_sums"));
    Prof pf(ctx, "k_scan_rescan", 8.0 * n);
    hipLaunchKernelGGL((heads ? k_scan_rescan<true> : k_scan_rescan<false>), dim3((u32)G), dim3(256), 0,
                       ctx->stream, in, out, n, C, (const u32*)sums);
    return launch_check(ctx, "k_scan_rescan");
  }
  TRY(fill(ctx, status, ntiles, 0));
  TRY(fill(ctx, ctx->small_d + 8, 1, 0));
  Prof pf(ctx, "k_scan_excl", 8.0 * n);
  hipLaunchKernelGGL((heads ? k_scan_excl<SCAN_ITEMS, true> : k_scan_excl<SCAN_ITEMS, false>), dim3((u32)ntiles),
                     dim3(256), 0, ctx->stream, in, out, n, status, ctx->small_d + 8, total_d);
  return launch_check(ctx, "k_scan_excl");
}

// ------------------------------------------------------------------------
// LSD radix sort of (kin, vin) -> (kout, vout) over the low keybits bits.
// Ping-pongs through two workspace pairs; never writes a buffer a pass reads.
// strict: a key with a bit at or above keybits is refused (FMCS_E_ARG);
// otherwise the bits above keybits are ignored (a partition on low bits).
// ------------------------------------------------------------------------
template <class K>
int radix_sort(fmcs_ctx* ctx, const K* kin, const u32* vin, K* kout, u32* vout, u64 n, u32 keybits,
               bool strict = true) {
  if (n == 0) return FMCS_OK;
  if (n >= (1ull << 30)) return fail(ctx, FMCS_E_UNSUPPORTED, "radix sort supports n < 2^30");
  keybits = std::min<u32>(keybits, 8 * sizeof(K));
  u32 npasses = (keybits + RADIX_BITS - 1) / RADIX_BITS;
  if (n == 1 || npasses == 0) {
    if ((const void*)kout != (const void*)kin) HIPCHK(hipMemcpyAsync(kout, kin, n * sizeof(K), hipMemcpyDeviceToDevice, ctx->stream));
    if (vout != vin) HIPCHK(hipMemcpyAsync(vout, vin, n * sizeof(u32), hipMemcpyDeviceToDevice, ctx->stream));
    return FMCS_OK;
  }
  u32 *hist, *dbase, *tcount;
  // every pass's digit counts, then the flag of keys above keybits
  TRY(getbuf_t(ctx, WS_HIST, (u64)npasses * RADIX + 1, &hist));
  TRY(getbuf_t(ctx, WS_DBASE, (u64)npasses * RADIX, &dbase));
  // 64-bit keys on the taller tiles (longer digit runs)
  constexpr int IT = sizeof(K) == 8 ? WIDE_ITEMS : ONESWEEP_ITEMS;
  constexpr u64 TILE = 256 * IT;
  u64 ntiles = (n + TILE - 1) / TILE;
  TRY(getbuf_t(ctx, WS_TCOUNT, ntiles * RADIX, &tcount));
  TRY(fill(ctx, hist, (u64)npasses * RADIX + 1, 0));
  {
    u32 blocks = (u32)std::min<u64>((n + 255) / 256, 2048);
    Prof pf(ctx, "k_hist_all", (double)n * sizeof(K));
    const u64 bad = strict && keybits < 64 ? ~((1ull << keybits) - 1) : 0ull;
    hipLaunchKernelGGL((k_hist_all<K>), dim3(blocks), dim3(256), 0, ctx->stream, kin, n, npasses, keybits, bad, hist);
    TRY(launch_check(ctx, "k_hist_all"));
  }
  // per-pass trivial detection (one digit holds every row) and digit bases
  std::vector<u32> h((size_t)npasses * RADIX + 1);
  TRY(readback(ctx, hist, npasses * RADIX + 1, h.data()));
  if (h[(size_t)npasses * RADIX]) return fail(ctx, FMCS_E_ARG, "keys exceed keybits");
  hipLaunchKernelGGL(k_digit_base, dim3(npasses), dim3(256), 0, ctx->stream, hist, (const u32*)nullptr, dbase,
                     (u32*)nullptr, (u32*)nullptr);
  TRY(launch_check(ctx, "k_digit_base"));
  std::vector<u32> passes;
  for (u32 p = 0; p < npasses; ++p) {
    u64 tot = 0, mx = 0;
    for (u32 d = 0; d < RADIX; ++d) {
      tot += h[p * RADIX + d];
      mx = std::max<u64>(mx, h[p * RADIX + d]);
    }
    if (mx != tot) passes.push_back(p);
  }
  K *t0k, *t1k;
  u32 *t0v, *t1v;
  TRY(getbuf_t(ctx, WS_T0K, n, &t0k));
  TRY(getbuf_t(ctx, WS_T0V, n, &t0v));
  if (ctx->arena && (const void*)kin == (const void*)kout && vin == vout) {
    // in place from scratch: the caller's arrays are the second buffer
    // (odd passes write them, even passes the scratch pair; after an odd
    // number of passes the result is copied home below)
    t1k = kout;
    t1v = vout;
  } else {
    TRY(getbuf_t(ctx, WS_T1K, n, &t1k));
    TRY(getbuf_t(ctx, WS_T1V, n, &t1v));
  }
  // constants: seg_lo=0, seg_hi=n, tile_base={0, ntiles}, nsegs = 1
  u32 consts[5] = {0u, (u32)n, 0u, (u32)ntiles, 1u};
  TRY(upload(ctx, ctx->small_d + 16, consts, 5));
  const u32* cst = ctx->small_d + 16;
  const K* sk = kin;
  const u32* sv = vin;
  for (size_t i = 0; i < passes.size(); ++i) {
    u32 p = passes[i];
    bool last = (i + 1 == passes.size());
    K* dk;
    u32* dv;
    if (last && (const void*)kout != (const void*)sk && vout != sv) {
      dk = kout;
      dv = vout;
    } else {
      dk = (i & 1) ? t1k : t0k;
      dv = (i & 1) ? t1v : t0v;
    }
    const u32 dsh = digit_shift(keybits, npasses, p), dw = digit_width(keybits, npasses, p);
    TRY(klaunch(ctx, "k_tile_hist", kd_seg_hist<K, IT, RADIX_BITS>, cap_grid(ntiles, 2048),
                (double)n * sizeof(K), nullptr, 0.0, sk, cst, cst + 1, cst + 2, cst + 4, cst + 3, dsh, dw,
                (u32*)nullptr, tcount, (u64)0, (u32*)nullptr, (u32*)nullptr));
    TRY((scatter_pass<K, RADIX_BITS, NoPay, false, true, IT>)(ctx, "k_onesweep", sk, sv, dk, dv, cst, cst + 1, cst + 2,
                                                         cst + 4, cst + 3, ntiles,
                        dbase + (u64)p * RADIX, tcount, dsh, dw, 2.0 * n * (sizeof(K) + 4),
                        nullptr, 0.0));
    sk = dk;
    sv = dv;
  }
  if ((const void*)sk != (const void*)kout) HIPCHK(hipMemcpyAsync(kout, sk, n * sizeof(K), hipMemcpyDeviceToDevice, ctx->stream));
  if (sv != vout) HIPCHK(hipMemcpyAsync(vout, sv, n * sizeof(u32), hipMemcpyDeviceToDevice, ctx->stream));
  return FMCS_OK;
}

// Comparison sort of (kin, vin) -> (kout, vout): 2048-row tiles (register
// bitonic networks, merge path across waves), then merge-path passes
// doubling the sorted run length through the workspace.
template <class K>
int merge_sort(fmcs_ctx* ctx, const K* kin, const u32* vin, K* kout, u32* vout, u64 n) {
  if (n == 0) return FMCS_OK;
  if (n >= (1ull << 30)) return fail(ctx, FMCS_E_UNSUPPORTED, "merge sort supports n < 2^30");
  K *t0k, *t1k;
  u32 *t0v, *t1v;
  TRY(getbuf_t(ctx, WS_T0K, n, &t0k));
  TRY(getbuf_t(ctx, WS_T0V, n, &t0v));
  TRY(getbuf_t(ctx, WS_T1K, n, &t1k));
  TRY(getbuf_t(ctx, WS_T1V, n, &t1v));
  const u64 grid = (n + MERGE_TILE - 1) / MERGE_TILE;
  const double pass_bytes = 2.0 * n * (sizeof(K) + 4);
  u32 npass = 0;
  for (u64 run = MERGE_TILE; run < n; run *= 2) ++npass;
  // the tile sort writes the output itself when no merge pass follows (in
  // place is safe: a block reads its whole tile before writing it); else
  // the passes alternate between two work buffers and the last one writes
  // the output
  TRY(klaunch(ctx, "k_bitonic_tile", k_bitonic_tile<K>, grid, pass_bytes, nullptr, 0.0, kin, vin,
              npass ? t0k : kout, npass ? t0v : vout, n));
  if (!npass) return FMCS_OK;
  u32* split;
  TRY(getbuf_t(ctx, WS_STATUS, grid + 1, &split));
  K* sk = t0k;
  u32* sv = t0v;
  K* dk = t1k;
  u32* dv = t1v;
  u32 pass = 0;
  for (u64 run = MERGE_TILE; run < n; run *= 2, ++pass) {
    const bool last = pass + 1 == npass;
    TRY(klaunch(ctx, "k_merge_splits", k_merge_splits<K>, (grid + 255) / 256, 0.0, nullptr, 0.0, (const K*)sk,
                (const u32*)sv, n, run, (u32)grid, split));
    TRY(klaunch(ctx, "k_merge_pass", k_merge_pass<K>, grid, pass_bytes, nullptr, 0.0, (const K*)sk, (const u32*)sv,
                last ? kout : dk, last ? vout : dv, n, run, (const u32*)split));
    std::swap(sk, dk);
    std::swap(sv, dv);
  }
  return FMCS_OK;
}

// Group boundaries: head counts per tile, their exclusive scan, then ids
// and offsets; the group count goes to *count_d (device).  Without per-row
// ids the count pass also writes every row's head bit (n / 8 bytes) and the
// write pass reads those instead of the keys (and group ids) again.
template <class K>
int boundary_launch(fmcs_ctx* ctx, const K* v, const u32* in_group, u64 n, u32* out_group, u32* seg_off,
                    u32* count_d, u32 shift = 0) {
  const u64 ntiles = (n + BOUNDARY_TILE - 1) / BOUNDARY_TILE;
  u32* tc;
  TRY(getbuf_t(ctx, WS_BCNT, 2 * ntiles, &tc));
  u32* hb = nullptr;
  if (!out_group) TRY(getbuf_t(ctx, WS_BBITS, ntiles * 256, &hb));
  const double rd = (double)n * (sizeof(K) + (in_group ? 4 : 0));
  TRY(klaunch(ctx, "k_bcount", k_bcount<K, BOUNDARY_ITEMS>, ntiles, rd, nullptr, 0.0, v, in_group, n, shift, tc,
              hb));
  TRY(scan_excl(ctx, tc, tc + ntiles, ntiles, nullptr));
  return klaunch(ctx, "k_bwrite", k_bwrite<K, BOUNDARY_ITEMS>, ntiles,
                 hb ? n / 8.0 : rd + (out_group ? 4.0 * n : 0.0), nullptr, 0.0, v, in_group, n, shift,
                 (const u32*)(tc + ntiles), out_group, seg_off, count_d, (const u32*)hb);
}

// boundary: dense ids + optional segment offsets; returns count on host
template <class K>
int boundary(fmcs_ctx* ctx, const K* v, const u32* in_group, u64 n, u32* out_group, u32* seg_off, u32* ng) {
  if (n == 0) {
    *ng = 0;
    if (seg_off) TRY(fill(ctx, seg_off, 1, 0));
    return FMCS_OK;
  }
  if (n >= (1ull << 30)) return fail(ctx, FMCS_E_UNSUPPORTED, "boundary supports n < 2^30");
  TRY(boundary_launch<K>(ctx, v, in_group, n, out_group, seg_off, ctx->small_d + 41));
  return readback(ctx, ctx->small_d + 41, 1, ng);
}

// Tiles hold 256 * E rows: E = 2 (512 rows) when small segments are at most
// 256 rows, else E = 4 (1024 rows).  Smaller tiles need half the LDS, so
// twice as many are resident per CU to hide their load/scan latencies.
template <class K>
int tile_sort_launch(fmcs_ctx* ctx, TileSortArgs<K> a, u32 grid, double rows, u32 T = TILESORT_T) {
  if (grid == 0) return FMCS_OK;
  Prof pf(ctx, "k_tile_sort", 2.0 * rows * (sizeof(K) + 4));
  if (T <= 512)
    hipLaunchKernelGGL((k_tile_sort<K, 2>), dim3(grid), dim3(256), 0, ctx->stream, a);
  else
    hipLaunchKernelGGL((k_tile_sort<K, 4>), dim3(grid), dim3(256), 0, ctx->stream, a);
  return launch_check(ctx, "k_tile_sort");
}

template <class K>
int tile_bucket_launch(fmcs_ctx* ctx, TileSortArgs<K> a, u32* fb, u32 grid, double rows, u32 T) {
  if (grid == 0) return FMCS_OK;
  Prof pf(ctx, "k_tile_bucket", 2.0 * rows * (sizeof(K) + 4));
  if (T <= 512)
    hipLaunchKernelGGL((k_tile_bucket<K, 2>), dim3(grid), dim3(256), 0, ctx->stream, a, fb, a.tile_lo, a.tile_hi,
                       a.tile_s0, a.tile_s1, a.wlist);
  else
    hipLaunchKernelGGL((k_tile_bucket<K, 4>), dim3(grid), dim3(256), 0, ctx->stream, a, fb, a.tile_lo, a.tile_hi,
                       a.tile_s0, a.tile_s1, a.wlist);
  return launch_check(ctx, "k_tile_bucket");
}

constexpr u32 TS_SEGBITS = 11;  // local segment index of a tile of up to TILESORT_TMAX = 2048 rows

// ------------------------------------------------------------------------
// Batched, size-binned segmented sort of (keys, vals) within segments.
// Level 0 data in (keys, vals); results in place.
// ------------------------------------------------------------------------
template <class K>
int seg_sort_batched(fmcs_ctx* ctx, K* keys, u32* vals, u64 n, u32 keybits, const u32* seg_off, u32 nseg,
                     u32* err = nullptr) {
  if (n <= 1 || nseg == 0) return FMCS_OK;
  const u32 small_max = std::min<u32>(ctx->tune.small_max, (u32)TILESORT_T / 2);
  const u32 T = small_max <= 256 ? 512u : (u32)TILESORT_T;
  u32 W = std::max<u32>(1, std::min<u32>(ctx->tune.window, small_max + 1));
  if (W + small_max > T) W = T - small_max;
  const u32 msd_bits = std::max<u32>(1, std::min<u32>(ctx->tune.msd_bits, RADIX_BITS));
  keybits = std::min<u32>(keybits, 8 * sizeof(K));
  K* kb_;
  u32* vb_;
  TRY(getbuf_t(ctx, WS_KB, n, &kb_));
  TRY(getbuf_t(ctx, WS_VB, n, &vb_));
  // two segment-list slots: the current level's list and the next one
  const Slot lst[2] = {WS_ST, WS_BST}, lsz[2] = {WS_SZ, WS_BSZ};
  int cur = 0;
  u32 *cst, *csz;
  TRY(getbuf_t(ctx, lst[cur], (u64)nseg + 1, &cst));
  TRY(getbuf_t(ctx, lsz[cur], (u64)nseg + 1, &csz));
  hipLaunchKernelGGL(k_offsets_to_list, dim3((nseg + 255) / 256), dim3(256), 0, ctx->stream, seg_off, nseg, cst, csz);
  TRY(launch_check(ctx, "k_offsets_to_list"));
  u32 cur_n = nseg;
  u32 kb = keybits;
  u32 level = 0;
  // parents of the current bucket list (level >= 1): the previous level's
  // large segments and their tile bases
  const u32* plo = nullptr;
  const u32* ptb = nullptr;
  u32 pntiles = 0;
  for (;;) {
    u32 sm = small_max;
    u32 tkb = kb;
    // keys too wide to pack a segment index beside them: refused tiles go
    // to the comparison-rank fallback instead of the radix tile sort
    const bool packable = kb + TS_SEGBITS <= 64;
    const u32 Wl = std::min<u32>(W, sm + 1);
    const u32 wpt = (ONESWEEP_TILE + Wl - 1) / Wl;
    const u64 nw = plo ? (u64)pntiles * wpt : (n + Wl - 1) / Wl;
    u32 *tlo, *thi, *ts0, *ts1;
    TRY(getbuf_t(ctx, WS_TLO, nw, &tlo));
    TRY(getbuf_t(ctx, WS_THI, nw, &thi));
    TRY(getbuf_t(ctx, WS_TS0, nw, &ts0));
    TRY(getbuf_t(ctx, WS_TS1, nw, &ts1));
    u32* wlist;
    TRY(getbuf_t(ctx, WS_WLIST, nw, &wlist));
    TRY(fill(ctx, tlo, nw, 0));
    TRY(fill(ctx, thi, nw, 0));
    TRY(fill(ctx, ctx->small_d + 51, 1, 0));
    hipLaunchKernelGGL(k_windows, dim3((cur_n + 255) / 256), dim3(256), 0, ctx->stream, cst, csz, cur_n, Wl, sm, plo,
                       ptb, wpt, tlo, thi, ts0, ts1, wlist, ctx->small_d + 51);
    TRY(launch_check(ctx, "k_windows"));
    // large segments of this level, and the row counts of both bins
    u32 *flag, *pos;
    TRY(getbuf_t(ctx, WS_FLAG, (u64)cur_n + 1, &flag));
    TRY(getbuf_t(ctx, WS_POS, (u64)cur_n + 1, &pos));
    TRY(fill(ctx, ctx->small_d + 52, 2, 0));
    hipLaunchKernelGGL(k_large_flags, dim3((cur_n + 255) / 256), dim3(256), 0, ctx->stream, csz, cur_n, sm, flag,
                       ctx->small_d + 52);
    TRY(launch_check(ctx, "k_large_flags"));
    TRY(scan_excl(ctx, flag, pos, cur_n, ctx->small_d + 48));
    u32 cnts[6];
    TRY(readback(ctx, ctx->small_d + 48, 6, cnts));
    const u32 nlarge = cnts[0];
    const u32 nwin = cnts[3];  // windows holding rows (compact list)
    const u64 tile_rows_est = cnts[4];
    const u64 large_rows = cnts[5];
    TileSortArgs<K> a;
    std::memset(&a, 0, sizeof(a));
    a.kin[0] = keys; a.vin[0] = vals; a.kin[1] = kb_; a.vin[1] = vb_;
    a.kout = keys; a.vout = vals;
    a.tile_lo = tlo; a.tile_hi = thi; a.tile_s0 = ts0; a.tile_s1 = ts1; a.seg_start = cst;
    a.src_sel = level & 1; a.kb = tkb; a.pad_seg = T; a.direct = 0; a.wlist = wlist;
    if (level == 0 && err && keybits < 8 * sizeof(K)) {  // every row is a window row or a large-segment row here
      a.bad = ~((1ull << keybits) - 1);
      a.err = err;
    }
    {
      // bucket fast path; tiles it refuses go to the radix tile sort
      u32* fb;
      TRY(getbuf_t(ctx, WS_FALLBACK, nw, &fb));
      TRY(fill(ctx, fb, nw, 0));
      TRY(tile_bucket_launch<K>(ctx, a, fb, nwin, (double)tile_rows_est, T));
      a.only = fb;
      if (packable) {
        TRY(tile_sort_launch<K>(ctx, a, nwin, 0.0, T));
      } else if (nwin) {
        Prof pf(ctx, "k_tile_rank", 0.0);
        if (T <= 512)
          hipLaunchKernelGGL((k_tile_rank<K, 2>), dim3(nwin), dim3(256), 0, ctx->stream, a);
        else
          hipLaunchKernelGGL((k_tile_rank<K, 4>), dim3(nwin), dim3(256), 0, ctx->stream, a);
        TRY(launch_check(ctx, "k_tile_rank"));
      }
    }
    if (nlarge == 0) break;
    u32 *llo, *lhi, *ltiles, *tbase;
    TRY(getbuf_t(ctx, WS_LLO, nlarge, &llo));
    TRY(getbuf_t(ctx, WS_LHI, nlarge, &lhi));
    TRY(getbuf_t(ctx, WS_LTILES, (u64)nlarge + 1, &ltiles));
    TRY(getbuf_t(ctx, WS_TBASE, (u64)nlarge + 1, &tbase));
    hipLaunchKernelGGL(k_large_compact, dim3((cur_n + 255) / 256), dim3(256), 0, ctx->stream, cst, csz, cur_n, sm, pos,
                       (u32)ONESWEEP_TILE, llo, lhi, ltiles);
    TRY(launch_check(ctx, "k_large_compact"));
    TRY(scan_excl(ctx, ltiles, tbase, nlarge, ctx->small_d + 49));
    u32 ntiles;
    TRY(readback(ctx, ctx->small_d + 49, 1, &ntiles));
    const K* skey = (level & 1) ? kb_ : keys;
    const u32* sval = (level & 1) ? vb_ : vals;
    K* dkey = (level & 1) ? keys : kb_;
    u32* dval = (level & 1) ? vals : vb_;
    if (kb == 0) {
      // every key bit consumed: rows are already in final order
      if (level & 1) {
        hipLaunchKernelGGL((k_copy_segs<K>), dim3(ntiles), dim3(256), 0, ctx->stream, skey, sval, keys, vals, llo, lhi,
                           tbase, nlarge, (u32)ONESWEEP_TILE);
        TRY(launch_check(ctx, "k_copy_segs"));
      }
      break;
    }
    const u32 D = std::min<u32>(msd_bits, kb);
    const u32 shift = kb - D;
    u32 *hist, *dbase, *tcount;
    TRY(getbuf_t(ctx, WS_HIST, (u64)nlarge * RADIX, &hist));
    TRY(getbuf_t(ctx, WS_DBASE, (u64)nlarge * RADIX, &dbase));
    TRY(getbuf_t(ctx, WS_TCOUNT, (u64)ntiles * RADIX, &tcount));
    TRY(fill(ctx, hist, (u64)nlarge * RADIX, 0));
    const u32* nlarge_d = ctx->small_d + 48;
    const u32* ntiles_d = ctx->small_d + 49;
    TRY(klaunch(ctx, "k_seg_hist", kd_seg_hist<K, ONESWEEP_ITEMS, RADIX_BITS>, cap_grid(ntiles, 2048),
                (double)large_rows * sizeof(K), nullptr, 0.0, skey, (const u32*)llo, (const u32*)lhi,
                (const u32*)tbase, nlarge_d, ntiles_d, shift, D, hist, tcount, a.bad, a.err, (u32*)nullptr));
    const int nxt = cur ^ 1;
    u32 *nst, *nsz;
    TRY(getbuf_t(ctx, lst[nxt], (u64)nlarge * RADIX + 1, &nst));
    TRY(getbuf_t(ctx, lsz[nxt], (u64)nlarge * RADIX + 1, &nsz));
    hipLaunchKernelGGL(k_digit_base, dim3(nlarge), dim3(256), 0, ctx->stream, hist, llo, dbase, nst, nsz);
    TRY(launch_check(ctx, "k_digit_base"));
    TRY(scatter_pass<K>(ctx, "k_onesweep_msd", skey, sval, dkey, dval, llo, lhi, tbase, nlarge_d, ntiles_d, ntiles,
                        dbase, tcount, shift, D, 2.0 * large_rows * (sizeof(K) + 4), nullptr, 0.0));
    cur = nxt;
    cst = nst;
    csz = nsz;
    cur_n = nlarge * RADIX;
    plo = llo;
    ptb = tbase;
    pntiles = ntiles;
    kb -= D;
    ++level;
  }
  return FMCS_OK;
}

// plain per-round GPU form: one sort dispatch per group, in group order
template <class K>
int seg_sort_plain(fmcs_ctx* ctx, K* keys, u32* vals, u64 n, u32 keybits, const u32* seg_off, u32 nseg) {
  std::vector<u32> off((size_t)nseg + 1);
  HIPCHK(hipMemcpyAsync(off.data(), seg_off, ((size_t)nseg + 1) * 4, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (u32 s = 0; s < nseg; ++s) {
    u32 lo = off[s], hi = off[s + 1];
    if (hi < lo || hi > n) return fail(ctx, FMCS_E_ARG, "segment offsets out of range");
    u32 sz = hi - lo;
    if (sz <= 1) continue;
    if (sz <= TILESORT_T && keybits + TS_SEGBITS <= 64) {
      TileSortArgs<K> a;
      std::memset(&a, 0, sizeof(a));
      a.kin[0] = keys; a.vin[0] = vals; a.kout = keys; a.vout = vals;
      a.kb = keybits; a.pad_seg = TILESORT_T; a.direct = 1; a.d_lo = lo; a.d_hi = hi;
      TRY(tile_sort_launch<K>(ctx, a, 1, (double)sz));
    } else {
      TRY(radix_sort<K>(ctx, keys + lo, vals + lo, keys + lo, vals + lo, sz, keybits));
    }
  }
  return FMCS_OK;
}

template <class K>
int gather_launch(fmcs_ctx* ctx, const K* column, const u32* order, K* out, u64 n);

// The unbinned form when (group id, key) does not fit 64 bits: the same LSD
// order in two stable sorts, by the key (carrying each row's position), then
// by the group id of those positions; rows keep their incoming order among
// equal (group, key) pairs, as in the composite form.
template <class K>
int seg_sort_unbinned_wide(fmcs_ctx* ctx, K* keys, u32* vals, u64 n, u32 keybits, const u32* group, u32 gbits,
                           u32* err) {
  const u64 grid = cap_grid((n + 255) / 256, 8192);
  if (err)
    TRY(klaunch(ctx, "k_key_check", k_key_check<K>, grid, (double)n * sizeof(K), nullptr, 0.0, (const K*)keys, n,
                ~((1ull << keybits) - 1), err));
  K* ka;
  u32 *pos, *u0, *u1;
  u64 *c, *c2;
  TRY(getbuf_t(ctx, WS_UNBK, n, &ka));  // the caller's keys may live in WS_KA
  TRY(getbuf_t(ctx, WS_UNB0, n, &pos));
  TRY(getbuf_t(ctx, WS_UNB1, n, &u0));
  TRY(getbuf_t(ctx, WS_COMP, n + 2, &c));
  TRY(getbuf_t(ctx, WS_COMP2, n + 2, &c2));
  const u64 half = (n + 3) & ~3ull;  // the second array of each buffer 16-byte aligned (gather orders)
  u32* gA = (u32*)c;
  u32* iA = (u32*)c + half;
  u32* jB = (u32*)c2;
  u1 = (u32*)c2 + half;
  // 1. by key: ka = sorted keys, iA = their rows
  TRY(klaunch(ctx, "k_iota", k_iota_u32, grid, 0.0, nullptr, 0.0, pos, n));
  TRY(radix_sort<K>(ctx, keys, pos, ka, iA, n, keybits));
  // 2. by group, stably: jB = positions in the key order
  TRY(gather_launch<u32>(ctx, group, iA, gA, n));
  TRY(klaunch(ctx, "k_iota", k_iota_u32, grid, 0.0, nullptr, 0.0, pos, n));
  TRY(radix_sort<u32>(ctx, gA, pos, u0, jB, n, gbits));
  // 3. rows out: keys from the key order, oids through the original rows
  TRY(gather_launch<K>(ctx, ka, jB, keys, n));
  TRY(gather_launch<u32>(ctx, iA, jB, u0, n));
  TRY(gather_launch<u32>(ctx, vals, u0, u1, n));
  HIPCHK(hipMemcpyAsync(vals, u1, n * sizeof(u32), hipMemcpyDeviceToDevice, ctx->stream));
  return FMCS_OK;
}

// single unbinned segmented sort: LSD radix over (group id, key) composites
template <class K>
int seg_sort_unbinned(fmcs_ctx* ctx, K* keys, u32* vals, u64 n, u32 keybits, const u32* group, u32 ngroups,
                      u32* err) {
  u32 gbits = 0;
  while (gbits < 32 && (1ull << gbits) < ngroups) ++gbits;
  if (gbits + keybits > 64) return seg_sort_unbinned_wide<K>(ctx, keys, vals, n, keybits, group, gbits, err);
  u64 *c, *c2;
  TRY(getbuf_t(ctx, WS_COMP, n, &c));
  TRY(getbuf_t(ctx, WS_COMP2, n, &c2));
  u32 blocks = (u32)std::min<u64>((n + 255) / 256, 8192);
  hipLaunchKernelGGL((k_composite<K>), dim3(blocks), dim3(256), 0, ctx->stream, keys, group, n, keybits, c, err);
  TRY(launch_check(ctx, "k_composite"));
  TRY(radix_sort<u64>(ctx, c, vals, c2, vals, n, gbits + keybits));
  u64 km = keybits >= 64 ? ~0ull : ((1ull << keybits) - 1);
  hipLaunchKernelGGL((k_strip_composite<K>), dim3(blocks), dim3(256), 0, ctx->stream, c2, n, km, keys);
  return launch_check(ctx, "k_strip_composite");
}

template <class K>
int seg_sort_dd_host(fmcs_ctx* ctx, K* keys, u32* vals, u64 n, u32 keybits, const u32* seg_off, u32 nseg);

// Keys with bits above keybits are an error in every mode: the eager modes
// return FMCS_E_ARG (the plain form before it sorts anything); the
// device-driven ones set the context's error word, reported by the next
// call that returns a count or by fmcs_ctx_sync.
template <class K>
int seg_sort(fmcs_ctx* ctx, K* keys, u32* vals, u64 n, u32 keybits, const u32* seg_off, u32 nseg, int mode,
             const u32* group) {
  if (mode == FMCS_MODE_DEVICE || mode == FMCS_MODE_GRAPH)
    return seg_sort_dd_host<K>(ctx, keys, vals, n, keybits, seg_off, nseg);
  if (mode != FMCS_MODE_BATCHED && mode != FMCS_MODE_PLAIN && mode != FMCS_MODE_UNBINNED)
    return fail(ctx, FMCS_E_ARG, "unknown mode");
  if (mode == FMCS_MODE_UNBINNED && !group) return fail(ctx, FMCS_E_ARG, "unbinned mode needs group ids");
  if (n <= 1 || nseg == 0) return FMCS_OK;
  keybits = std::min<u32>(keybits, 8 * sizeof(K));
  const bool check = keybits < 8 * sizeof(K);
  u32* err = ctx->small_d + 60;
  if (check) TRY(fill(ctx, err, 1, 0));
  u32 e = 0;
  switch (mode) {
    case FMCS_MODE_BATCHED: TRY(seg_sort_batched<K>(ctx, keys, vals, n, keybits, seg_off, nseg, check ? err : nullptr)); break;
    case FMCS_MODE_PLAIN:
      if (check) {
        TRY(klaunch(ctx, "k_key_check", k_key_check<K>, cap_grid((n + 255) / 256, 2048), (double)n * sizeof(K),
                    nullptr, 0.0, (const K*)keys, n, ~((1ull << keybits) - 1), err));
        TRY(readback(ctx, err, 1, &e));
        if (e) return fail(ctx, FMCS_E_ARG, "keys exceed keybits");
      }
      return seg_sort_plain<K>(ctx, keys, vals, n, keybits, seg_off, nseg);
    default: TRY(seg_sort_unbinned<K>(ctx, keys, vals, n, keybits, group, nseg, err)); break;
  }
  if (check) {
    TRY(readback(ctx, err, 1, &e));
    if (e) return fail(ctx, FMCS_E_ARG, "keys exceed keybits");
  }
  return FMCS_OK;
}

int group_info_from_offsets(fmcs_ctx* ctx, u32 bank, const u32* seg_off, u32 nseg, fmcs_group_info* info) {
  std::vector<u32> off((size_t)nseg + 1);
  HIPCHK(hipMemcpyAsync(off.data(), seg_off, ((size_t)nseg + 1) * 4, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  u32 bs = (256u * 1024u) / (4u * std::max<u32>(bank, 4));  // BLOCKSIZE<T>() (src/params.h:141-143)
  fmcs_group_info g = {0, 0, 0, 0, 0};
  for (u32 s = 0; s < nseg; ++s) {
    u32 sz = off[s + 1] - off[s];
    g.ngroups++;
    g.sum_group_size += sz;
    if (sz < bs) g.ngroups_incache++; else g.ngroups_outofcache++;
    g.ngroups_gtOne += (sz > 1);
  }
  *info = g;
  return FMCS_OK;
}

template <class K>
int gather_launch(fmcs_ctx* ctx, const K* column, const u32* order, K* out, u64 n);

template <class K>
int sort_by_group_t(fmcs_ctx* ctx, u32 keybits, u64 n, const K* column, u32* order, const u32* in_group,
                    u32* out_group, int last_round, int mode, K* sorted, fmcs_group_info* info, u32* ngroups,
                    const u32* in_seg_off, u32 in_nseg, u32* out_seg_off, double* t_sort, double* t_gather,
                    double* t_extract) {
  K* ka = sorted;
  if (!ka) TRY(getbuf_t(ctx, WS_KA, n, &ka));
  hipEvent_t* ev = ctx->ev;
  HIPCHK(hipEventRecord(ev[0], ctx->stream));
  if (!in_group) {
    HIPCHK(hipEventRecord(ev[1], ctx->stream));
    TRY(radix_sort<K>(ctx, column, order, ka, order, n, keybits));
  } else {
    if (n) TRY(gather_launch<K>(ctx, column, order, ka, n));
    HIPCHK(hipEventRecord(ev[1], ctx->stream));
    const u32* so = in_seg_off;
    u32 nseg = in_nseg;
    if (!so) {
      u32* tmp;
      TRY(getbuf_t(ctx, WS_SEGOFF, n + 1, &tmp));
      TRY(boundary<u32>(ctx, in_group, nullptr, n, nullptr, tmp, &nseg));
      so = tmp;
    }
    if (info) TRY(group_info_from_offsets(ctx, sizeof(K), so, nseg, info));
    TRY(seg_sort<K>(ctx, ka, order, n, keybits, so, nseg, mode, in_group));
  }
  HIPCHK(hipEventRecord(ev[2], ctx->stream));
  u32 ng = 0;
  if (!last_round) TRY(boundary<K>(ctx, ka, in_group, n, out_group, out_seg_off, &ng));
  HIPCHK(hipEventRecord(ev[3], ctx->stream));
  HIPCHK(hipEventSynchronize(ev[3]));
  float a = 0, b = 0, c = 0;
  HIPCHK(hipEventElapsedTime(&a, ev[0], ev[1]));
  HIPCHK(hipEventElapsedTime(&b, ev[1], ev[2]));
  HIPCHK(hipEventElapsedTime(&c, ev[2], ev[3]));
  if (t_gather) *t_gather += a;
  if (t_sort) *t_sort += b;
  if (t_extract) *t_extract += c;
  if (ngroups) *ngroups = ng;
  return FMCS_OK;
}

// ========================================================================
// Device-driven batched path.  Counts stay in device memory; every kernel
// reads its sizes from there and runs grid-stride or persistent over them,
// so a multi-round sort issues no host synchronisation and can be captured
// in a hipGraph.  Buffers are sized from host-side bounds: a level's large
// segments each hold more than small_max rows, so at most n / (small_max+1)
// of them exist; their tiles number at most n / TILE + that.
// ========================================================================
constexpr u32 DD_LVL_MAX = 72;
constexpr u32 DD_ROUNDS_MAX = 64;
constexpr u32 DD_HDR = 256;  // [0, 64) group count per round, [72] key error, [76, 81) LSD consts, [88] segment count,
                             // [96, 100) LSD consts of WIDE_TILE passes
constexpr u32 DD_ERR = 72;
constexpr u32 DD_CONST = 76;
constexpr u32 DD_CONST_W = 96;  // one-segment constants of the WIDE_TILE passes
constexpr u64 DD_WORDS = DD_HDR + (u64)DD_ROUNDS_MAX * DD_LVL_MAX * DD_LVL_WORDS;

// The device-driven word area.  Its key-error word dd[DD_ERR] is sticky:
// cleared once when the area is allocated and again only when an error
// has been read back and reported (dd_finish), so an unsynchronised call's
// error survives later calls until a sync or a count readback reports it.
int dd_get(fmcs_ctx* ctx, u32** dd) {
  const bool fresh = ctx->ws[WS_DD].p == nullptr;
  TRY(getbuf_t(ctx, WS_DD, DD_WORDS, dd));
  if (fresh) TRY(fill(ctx, *dd + DD_ERR, 1, 0));
  return FMCS_OK;
}

// Clears the first `words` words of the area, all but the error word.
int dd_clear(fmcs_ctx* ctx, u32* dd, u64 words) {
  TRY(fill(ctx, dd, DD_ERR, 0));
  return fill(ctx, dd + DD_ERR + 1, words - DD_ERR - 1, 0);
}

// Exclusive scan of a device-sized list.  st holds ceil(n_max / SCAN_TILE)
// look-back status words that the kernel before it cleared (the first
// ceil(*n_d / SCAN_TILE) of them are enough).  flags: in holds the sizes of
// segments starting at seg_start, the scan is over (size > small_max), the
// large segments' (lo, hi, tiles) go to their places in llo, lhi, ltiles with
// their digit counts in hist (if set) zeroed, and rows (if set) [0..1] get
// the small and large rows.
int dd_scan(fmcs_ctx* ctx, const u32* in, u32* out, const u32* n_d, u64 n_max, u32* st, u32* counter, u32* total,
            bool flags = false, u32 small_max = 0, u32* rows = nullptr, const u32* seg_start = nullptr,
            u32* llo = nullptr, u32* lhi = nullptr, u32* ltiles = nullptr, u32* hist = nullptr,
            u32 tile_rows = ONESWEEP_TILE) {
  const u64 tiles_max = std::max<u64>(1, (n_max + SCAN_TILE - 1) / SCAN_TILE);
  // blocks take tiles from the counter, so a few hundred cover any list:
  // a larger grid only adds launch time (most lists are far below their bound)
  auto fn = flags ? kd_scan_excl<SCAN_ITEMS, true> : kd_scan_excl<SCAN_ITEMS, false>;
  return klaunch(ctx, "k_scan", fn, cap_grid(tiles_max, std::min<u64>(512, resident(ctx, (const void*)fn))), 0.0,
                 nullptr, 0.0, in, out, n_d, st, counter, total, small_max, rows, seg_start, tile_rows, llo,
                 lhi, ltiles, hist);
}

// LSD radix sort over every pass (no host-side trivial-pass skipping).
// Keys above keybits set the error word dd[DD_ERR].  vin == nullptr stands
// for the identity order (values = row indices), generated in the first pass.
// Unless P is NoPay, the payload column pin moves with the rows into pout.
template <class K, int RB, class P = NoPay>
int radix_sort_dd_t(fmcs_ctx* ctx, const K* kin, const u32* vin, K* kout, u32* vout, u64 n, u32 keybits, u32* dd,
                    const P* pin = nullptr, P* pout = nullptr) {
  constexpr u32 R = 1u << RB;
  constexpr bool HASP = !std::is_same<P, NoPay>::value;
  if (n == 0) return FMCS_OK;
  keybits = std::min<u32>(keybits, 8 * sizeof(K));
  const u32 npasses = (keybits + RB - 1) / RB;
  if (n == 1 || npasses == 0) {
    if (HASP && (const void*)pin != (const void*)pout)
      TRY(klaunch(ctx, "kd_copy_col", kd_copy_col<P>, cap_grid((n + 255) / 256, 4096), 2.0 * n * sizeof(P), nullptr, 0.0,
                  pin, pout, n));
    if ((const void*)kout != (const void*)kin || vout != vin)
      return klaunch(ctx, "kd_copy_pairs", kd_copy_pairs<K>, 1, 0.0, nullptr, 0.0, kin, vin, kout, vout, n);
    return FMCS_OK;
  }
  // 64-bit keys without a payload on the taller tiles (longer digit runs)
  constexpr int IT = (sizeof(K) == 8 && !HASP) ? WIDE_ITEMS : ONESWEEP_ITEMS;
  const u64 ntiles = (n + 256 * IT - 1) / (256 * IT);
  u32 *hist, *dbase, *tcount;
  TRY(getbuf_t(ctx, WS_HIST, (u64)npasses * R, &hist));
  TRY(getbuf_t(ctx, WS_DBASE, (u64)npasses * R, &dbase));
  TRY(getbuf_t(ctx, WS_TCOUNT, ntiles * R, &tcount));
  K *t0k, *t1k;
  u32 *t0v, *t1v;
  TRY(getbuf_t(ctx, WS_T0K, n, &t0k));
  TRY(getbuf_t(ctx, WS_T0V, n, &t0v));
  TRY(getbuf_t(ctx, WS_T1K, n, &t1k));
  TRY(getbuf_t(ctx, WS_T1V, n, &t1v));
  P *t0p = nullptr, *t1p = nullptr;
  if (HASP) {
    if (npasses > 1) TRY(getbuf_t(ctx, WS_T0P, n, &t0p));
    if (npasses > 2) TRY(getbuf_t(ctx, WS_T1P, n, &t1p));
  }
  TRY(fill(ctx, hist, (u64)npasses * R, 0));
  // one segment [0, n) of ntiles tiles: seg_lo, seg_hi, tile_base[0], tile_base[1] (= ntiles); nsegs = 1
  u32* cst = dd + DD_CONST;
  TRY(klaunch(ctx, "k_set", kd_set4, 1, 0.0, nullptr, 0.0, cst, 0u, (u32)n, 0u, (u32)ntiles));
  TRY(klaunch(ctx, "k_set", kd_set4, 1, 0.0, nullptr, 0.0, cst + 4, 1u, 0u, 0u, 0u));
  const u32* one = cst + 4;
  const K* sk = kin;
  const u32* sv = vin;
  const P* sp = pin;
  for (u32 p = 0; p < npasses; ++p) {
    const bool last = (p + 1 == npasses);
    K* dk;
    u32* dv;
    P* dp = nullptr;
    if (last && (const void*)kout != (const void*)sk && vout != sv &&
        (!HASP || (const void*)pout != (const void*)sp)) {
      dk = kout;
      dv = vout;
      dp = pout;
    } else {
      dk = (p & 1) ? t1k : t0k;
      dv = (p & 1) ? t1v : t0v;
      if (HASP) {
        // a last pass that cannot write pout directly needs a second buffer
        if (!t0p) TRY(getbuf_t(ctx, WS_T0P, n, &t0p));
        if ((p & 1) && !t1p) TRY(getbuf_t(ctx, WS_T1P, n, &t1p));
        dp = (p & 1) ? t1p : t0p;
      }
    }
    u32* hp = hist + (u64)p * R;
    // the first pass also flags keys with bits above keybits
    const u64 bad = (p == 0 && keybits < 8 * sizeof(K)) ? ~((1ull << keybits) - 1) : 0ull;
    const u32 dsh = digit_shift(keybits, npasses, p), dw = digit_width(keybits, npasses, p);
    TRY(klaunch(ctx, "k_tile_hist", kd_seg_hist<K, IT, RB>, cap_grid(ntiles, 2048), (double)n * sizeof(K),
                nullptr, 0.0, sk, (const u32*)cst, (const u32*)cst + 1, (const u32*)cst + 2, one, (const u32*)cst + 3,
                dsh, dw, hp, tcount, bad, dd + DD_ERR, (u32*)nullptr));
    TRY(klaunch(ctx, "k_digit_base", kd_digit_base<RB>, 1, 0.0, nullptr, 0.0, (const u32*)hp, (const u32*)cst,
                dbase + (u64)p * R, (u32*)nullptr, (u32*)nullptr, one, (u32*)nullptr));
    const double pb = HASP ? 2.0 * n * sizeof(P) : 0.0;
    const int rc = scatter_pass<K, RB, P, false, true, IT>(ctx, "k_onesweep", sk, sv, dk, dv, cst, cst + 1, cst + 2, one, cst + 3,
                                          ntiles, dbase + (u64)p * R, tcount, dsh, dw,
                                          2.0 * n * (sizeof(K) + 4) + pb, nullptr, 0.0, sp, dp);
    if (rc != FMCS_OK) return rc;
    sk = dk;
    sv = dv;
    sp = dp;
  }
  if ((const void*)sk != (const void*)kout || sv != vout)
    TRY(klaunch(ctx, "kd_copy_pairs", kd_copy_pairs<K>, cap_grid((n + 255) / 256, 4096), 2.0 * n * (sizeof(K) + 4), nullptr,
                0.0, sk, sv, kout, vout, n));
  if (HASP && (const void*)sp != (const void*)pout)
    TRY(klaunch(ctx, "kd_copy_col", kd_copy_col<P>, cap_grid((n + 255) / 256, 4096), 2.0 * n * sizeof(P), nullptr, 0.0,
                sp, pout, n));
  return FMCS_OK;
}

// 8-bit digits.  (10-bit digits halve a 20-bit key's passes to two, but a
// 1024-digit scatter pass measured 1.75x an 8-bit one here: 4-row digit
// runs per 4096-row tile and two resident tiles per CU instead of three.)
template <class K>
int radix_sort_dd(fmcs_ctx* ctx, const K* kin, const u32* vin, K* kout, u32* vout, u64 n, u32 keybits, u32* dd) {
  return radix_sort_dd_t<K, RADIX_BITS>(ctx, kin, vin, kout, vout, n, keybits, dd);
}

template <class K>
int boundary_dd(fmcs_ctx* ctx, const K* v, const u32* in_group, u64 n, u32* out_group, u32* seg_off, u32* count_d,
                u32 shift = 0) {
  if (n == 0) return FMCS_OK;
  return boundary_launch<K>(ctx, v, in_group, n, out_group, seg_off, count_d, shift);
}

// Round 1 with round 2's u64 column packed beside it: one LSD sort of
// (c1 << pbits) | c2 composites on their c1 bits.  The first pass reads
// both columns and writes the composites; the later passes move u64
// composites and oids (12 bytes a row against 16 for a key, an oid and a
// carried u64).  With more than one pass, the last one also writes the sorted
// c1 keys to side (u32), which a boundary pass reads at half the bytes of
// the composites.  Keys above keybits, and c2 values above pbits, set
// dd[DD_ERR].  Needs n >= 2, keybits >= 1 and keybits + pbits <= 64.  With
// more than one pass and dig set, the last pass also writes every row's
// composite >> dig_shift as a byte there (the digit the second round's first
// MSD level sorts on, which its histogram then reads instead of the keys).
// The key bits are split over the passes by digit_width / digit_shift.
template <class K>
int radix_sort_dd_fused(fmcs_ctx* ctx, const K* kin, const u64* pin, u32 keybits, u32 pbits, u64* cout, u32* vout,
                        u32* side, u64 n, u32* dd, uint8_t* dig = nullptr, u32 dig_shift = 0) {
  constexpr u32 RB = RADIX_BITS, R = 1u << RB;
  const u32 npasses = (keybits + RB - 1) / RB;
  // pass p sorts c1's bits dshift(p) .. dshift(p) + dwidth(p) - 1
  auto dwidth = [&](u32 p) { return digit_width(keybits, npasses, p); };
  auto dshift = [&](u32 p) { return digit_shift(keybits, npasses, p); };
  // the fused first pass on ONESWEEP_TILE-row tiles, the composite passes
  // on WIDE_TILE-row tiles (each with its own one-segment constants)
  const u64 ntiles = (n + ONESWEEP_TILE - 1) / ONESWEEP_TILE;
  const u64 ntiles_w = (n + WIDE_TILE - 1) / WIDE_TILE;
  u32 *hist, *dbase, *tcount, *t0v, *t1v;
  u64 *t0c = nullptr, *t1c = nullptr;
  TRY(getbuf_t(ctx, WS_HIST, (u64)npasses * R, &hist));
  TRY(getbuf_t(ctx, WS_DBASE, (u64)npasses * R, &dbase));
  TRY(getbuf_t(ctx, WS_TCOUNT, ntiles * R, &tcount));
  TRY(getbuf_t(ctx, WS_T0V, n, &t0v));
  TRY(getbuf_t(ctx, WS_T1V, n, &t1v));
  if (npasses > 1) TRY(getbuf_t(ctx, WS_T0P, n, &t0c));
  if (npasses > 2) TRY(getbuf_t(ctx, WS_T1P, n, &t1c));
  // every pass but the last writes each row's digit of the next pass as a
  // byte beside it, so the next pass's tile histogram reads 1 byte a row
  // instead of the 8-byte composite (read before the next scatter rewrites
  // the bytes, so one buffer serves every pass)
  uint8_t* dnx = nullptr;
  if (npasses > 1) TRY(getbuf_t(ctx, WS_DNEXT, n, &dnx));
  TRY(fill(ctx, hist, (u64)npasses * R, 0));
  u32* cst = dd + DD_CONST;
  u32* csw = dd + DD_CONST_W;
  TRY(klaunch(ctx, "k_set", kd_set4, 1, 0.0, nullptr, 0.0, cst, 0u, (u32)n, 0u, (u32)ntiles));
  TRY(klaunch(ctx, "k_set", kd_set4, 1, 0.0, nullptr, 0.0, cst + 4, 1u, 0u, 0u, 0u));
  if (npasses > 1) TRY(klaunch(ctx, "k_set", kd_set4, 1, 0.0, nullptr, 0.0, csw, 0u, (u32)n, 0u, (u32)ntiles_w));
  const u32* one = cst + 4;
  const u64 kbad = keybits < 8 * sizeof(K) ? ~((1ull << keybits) - 1) : 0ull;
  const u64 pbad = pbits < 64 ? ~((1ull << pbits) - 1) : 0ull;
  const u64* sc = nullptr;
  const u32* sv = nullptr;
  for (u32 p = 0; p < npasses; ++p) {
    const bool last = (p + 1 == npasses);
    u64* dc = last ? cout : ((p & 1) ? t1c : t0c);
    u32* dv = last ? vout : ((p & 1) ? t1v : t0v);
    const u32 dbits = dwidth(p);
    u32* hp = hist + (u64)p * R;
    if (p == 0) {
      TRY(klaunch(ctx, "k_tile_hist", kd_seg_hist<K, ONESWEEP_ITEMS, RB>, cap_grid(ntiles, 2048),
                  (double)n * sizeof(K), nullptr, 0.0, kin, (const u32*)cst, (const u32*)cst + 1,
                  (const u32*)cst + 2, one, (const u32*)cst + 3, 0u, dbits, hp, tcount, kbad, dd + DD_ERR,
                  (u32*)nullptr));
    } else {
      TRY(klaunch(ctx, "k_tile_hist", kd_seg_hist<uint8_t, WIDE_ITEMS, RB>, cap_grid(ntiles_w, 2048),
                  (double)n, nullptr, 0.0, (const uint8_t*)dnx, (const u32*)csw, (const u32*)csw + 1,
                  (const u32*)csw + 2, one, (const u32*)csw + 3, 0u, dbits, hp, tcount, 0ull, (u32*)nullptr,
                  (u32*)nullptr));
    }
    TRY(klaunch(ctx, "k_digit_base", kd_digit_base<RB>, 1, 0.0, nullptr, 0.0, (const u32*)hp, (const u32*)cst,
                dbase + (u64)p * R, (u32*)nullptr, (u32*)nullptr, one, (u32*)nullptr));
    if (p == 0) {
      // the next digit: c1's bits from dshift(1) (a byte a row beside the
      // pass's own bytes: c1 and c2 read, composite and oid written)
      ctx->extra_next = last ? 0.0 : 1.0 * n;
      TRY((scatter_pass<K, RB, u64, true, true>(ctx, "k_onesweep_fuse", kin, (const u32*)nullptr, (K*)nullptr, dv, cst, cst + 1,
                                          cst + 2, one, cst + 3, ntiles, dbase, tcount, 0u, dbits,
                                          (double)n * (sizeof(K) + 8 + 8 + 4), nullptr, 0.0, pin, dc,
                                          pbits, pbad, dd + DD_ERR, (u32*)nullptr, last ? (uint8_t*)nullptr : dnx,
                                          last ? 0u : dshift(1))));
    } else {
      // 2(w + 4) a row for the composite pass; the next digit (a byte a
      // row) or, on the last pass, the sorted c1 (4 bytes a row) besides
      ctx->extra_next = last ? (dig ? 5.0 : 4.0) * n : 1.0 * n;
      TRY((scatter_pass<u64, RB, NoPay, false, true, WIDE_ITEMS>)(ctx, "k_onesweep", sc, sv, dc, dv, csw, csw + 1, csw + 2, one, csw + 3, ntiles_w,
                            dbase + (u64)p * R, tcount, pbits + dshift(p), dbits,
                            2.0 * n * (8 + 4), nullptr, 0.0, (const NoPay*)nullptr,
                            (NoPay*)nullptr, pbits, 0ull, (u32*)nullptr, last ? side : (u32*)nullptr,
                            last ? dig : dnx, last ? dig_shift : pbits + dshift(p + 1)));
    }
    sc = dc;
    sv = dv;
  }
  return FMCS_OK;
}

struct DdLevel {
  u32 kb, sm, tkb, Wl, wpt, D, shift, single_max;
  bool last, packable;
  u64 nseg_max, nw_max, nlarge_max, ntiles_max, nsingle_max;
};

// Batched segmented sort of (keys, vals) within the *nseg_d segments given
// by seg_off; results in place.  lv: DD_LVL_MAX level-state records, zero.
// keys_out false: only vals are wanted; the tile sorts do not write the
// sorted keys back (keys ends unspecified).
template <class K>
int seg_sort_dd(fmcs_ctx* ctx, K* keys, u32* vals, u64 n, u32 keybits, const u32* seg_off, const u32* nseg_d,
                u64 nseg_max, u32* lv, bool keys_out = true, u32* err = nullptr, const uint8_t* dig0 = nullptr) {
  if (n <= 1 || nseg_max == 0) return FMCS_OK;
  // The persistent form (tune.med_max set): the levels only split segments
  // above med_max rows; one k_seg_bins launch then sorts every level's
  // windows (BIN_T-row tiles) and single segments of up to med_max rows.
  bool persist = ctx->tune.med_max != 0;
  const u32 med_max = std::min<u32>(std::max<u32>(ctx->tune.med_max, BIN_T), BIN_MED_MAX);
  // tiles of 512, 1024 or 2048 rows (2, 4 or 8 per thread) for segments of
  // up to 256, 512 or 1024 rows
  const u32 small_max = std::min<u32>(ctx->tune.small_max, persist ? BIN_T / 2 : (u32)TILESORT_TMAX / 2);
  const u32 T = persist ? BIN_T : (small_max <= 256 ? 512u : (small_max <= 512 ? 1024u : 2048u));
  u32 W = std::max<u32>(1, std::min<u32>(ctx->tune.window, small_max + 1));
  if (W + small_max > T) W = T - small_max;
  const u32 msd_bits = std::max<u32>(1, std::min<u32>(ctx->tune.msd_bits, RADIX_BITS));
  // static level schedule with buffer bounds
  std::vector<DdLevel> L;
  for (;;) {
    L.clear();
    u32 kb = std::min<u32>(keybits, 8 * sizeof(K));
    u64 cur_max = nseg_max, prev_tiles = 0;
    for (;;) {
      DdLevel d{};
      d.kb = kb;
      d.sm = small_max;
      d.tkb = kb;
      d.packable = kb + TS_SEGBITS <= 64;
      d.Wl = std::min<u32>(W, d.sm + 1);
      d.wpt = (WIDE_TILE + d.Wl - 1) / d.Wl;
      d.nseg_max = cur_max;
      d.nw_max = L.empty() ? (n + d.Wl - 1) / d.Wl : prev_tiles * d.wpt;
      // segments of sm + 1 .. 2T rows (at most the largest tile) sort alone
      // in an LDS tile of T or 2T rows ("single" windows); larger ones take
      // an MSD level.  Persistent: up to med_max rows, sorted by one block
      d.single_max = persist ? med_max : std::min<u32>(2 * T, (u32)TILESORT_TMAX);
      d.nsingle_max = std::min<u64>(cur_max, n / ((u64)d.sm + 1));
      d.nlarge_max = std::min<u64>(cur_max, n / ((u64)d.single_max + 1));
      d.ntiles_max = std::min<u64>(n / WIDE_TILE + d.nlarge_max, n);
      d.last = (kb == 0);
      if (!d.last) {
        d.D = std::min<u32>(msd_bits, kb);
        d.shift = kb - d.D;
      }
      L.push_back(d);
      if (d.last || d.nlarge_max == 0) break;
      if (L.size() >= DD_LVL_MAX - 1) return fail(ctx, FMCS_E_UNSUPPORTED, "too many segmented-sort levels");
      cur_max = d.nlarge_max * RADIX;
      prev_tiles = d.ntiles_max;
      kb -= d.D;
    }
    // the persistent launch takes at most BIN_LVL_MAX levels (narrow MSD
    // digits on wide keys need more): the level-synchronous form then
    if (!persist || L.size() <= BIN_LVL_MAX) break;
    persist = false;
  }
  u64 m_list[2] = {0, 0}, m_nw = 1, m_flag = 1, m_large = 1, m_tiles = 1, m_single = 1;
  for (size_t l = 0; l < L.size(); ++l) {
    m_single = std::max<u64>(m_single, L[l].nsingle_max);
    m_list[l & 1] = std::max<u64>(m_list[l & 1], L[l].nseg_max + 1);
    m_nw = std::max<u64>(m_nw, L[l].nw_max);
    m_flag = std::max<u64>(m_flag, L[l].nseg_max + 1);
    m_large = std::max<u64>(m_large, L[l].nlarge_max + 1);
    m_tiles = std::max<u64>(m_tiles, L[l].ntiles_max + 1);
  }
  // persistent form: every level's segment starts, window slots and single
  // segments stay until the one launch after the levels, each level in a
  // region of its own
  std::vector<u64> loff(L.size() + 1, 0), woff(L.size() + 1, 0), soff(L.size() + 1, 0);
  for (size_t l = 0; l < L.size(); ++l) {
    loff[l + 1] = loff[l] + L[l].nseg_max + 1;
    woff[l + 1] = woff[l] + L[l].nw_max;
    soff[l + 1] = soff[l] + L[l].nsingle_max;
  }
  if (persist) {
    if (loff[L.size()] >= (1ull << 32) || woff[L.size()] >= (1ull << 32) || soff[L.size()] >= (1ull << 32))
      return fail(ctx, FMCS_E_UNSUPPORTED, "segmented-sort lists exceed 2^32 entries");
    m_nw = std::max<u64>(woff[L.size()], 1);
    m_single = std::max<u64>(soff[L.size()], 1);
  }
  u32* plist = nullptr;
  if (persist) TRY(getbuf_t(ctx, WS_PLIST, loff[L.size()], &plist));
  K* kb_;
  u32* vb_;
  TRY(getbuf_t(ctx, WS_KB, n + 2, &kb_));  // + 2: two aligned u32 arrays when the keys narrow
  TRY(getbuf_t(ctx, WS_VB, n, &vb_));
  u32 *lst[2], *lsz[2];
  TRY(getbuf_t(ctx, WS_ST, std::max<u64>(m_list[0], 1), &lst[0]));
  TRY(getbuf_t(ctx, WS_SZ, std::max<u64>(m_list[0], 1), &lsz[0]));
  TRY(getbuf_t(ctx, WS_BST, std::max<u64>(m_list[1], 1), &lst[1]));
  TRY(getbuf_t(ctx, WS_BSZ, std::max<u64>(m_list[1], 1), &lsz[1]));
  u32 *tlo, *thi, *ts0, *ts1, *wlist, *fb, *fblist, *llo, *lhi, *ltiles, *tbase, *hist, *dbase, *tcount;
  TRY(getbuf_t(ctx, WS_TLO, m_nw, &tlo));
  TRY(getbuf_t(ctx, WS_THI, m_nw, &thi));
  TRY(getbuf_t(ctx, WS_TS0, m_nw, &ts0));
  TRY(getbuf_t(ctx, WS_TS1, m_nw, &ts1));
  TRY(getbuf_t(ctx, WS_WLIST, m_nw, &wlist));
  TRY(getbuf_t(ctx, WS_FALLBACK, m_nw, &fb));
  TRY(getbuf_t(ctx, WS_FBLIST, m_nw, &fblist));
  TRY(getbuf_t(ctx, WS_LLO, m_large, &llo));
  TRY(getbuf_t(ctx, WS_LHI, m_large, &lhi));
  TRY(getbuf_t(ctx, WS_LTILES, m_large, &ltiles));
  TRY(getbuf_t(ctx, WS_TBASE, m_large, &tbase));
  TRY(getbuf_t(ctx, WS_HIST, m_large * RADIX, &hist));
  TRY(getbuf_t(ctx, WS_DBASE, m_large * RADIX, &dbase));
  TRY(getbuf_t(ctx, WS_TCOUNT, m_tiles * RADIX, &tcount));
  // every MSD tile's segment: written by the histogram, read by the tile scan and the scatter
  u32* tseg;
  TRY(getbuf_t(ctx, WS_TSEG, m_tiles, &tseg));
  u32 *sgl, *fblist2;
  TRY(getbuf_t(ctx, WS_SGL, 3 * m_single, &sgl));
  TRY(getbuf_t(ctx, WS_FBLIST2, m_single, &fblist2));
  // a region for each of a level's two scans, both cleared by kd_windows
  const u64 st_words = (m_flag + SCAN_TILE - 1) / SCAN_TILE;
  u32* scan_st;
  TRY(getbuf_t(ctx, WS_SCANST, 2 * st_words, &scan_st));
  u32* scan_st2 = scan_st + st_words;
  u32 *slo = sgl, *shi = sgl + m_single, *sseg = sgl + 2 * m_single;
  u32 *rkl, *rkl2;
  TRY(getbuf_t(ctx, WS_RKL, m_nw, &rkl));
  TRY(getbuf_t(ctx, WS_RKL2, m_single, &rkl2));
  // Narrowed keys: in a round whose sorted keys are not wanted, with 64-bit
  // keys of which at most 32 bits are left to sort after the first MSD level,
  // every later level moves u32 keys (the bits still to sort): the later MSD
  // passes and the windows after the first level read and write 4 key bytes
  // a row instead of 8.  The first level scatters its large segments' keys
  // into the first u32 array; later levels alternate between two u32 arrays
  // carved from the level buffer.
  const bool narrow = !keys_out && !persist && sizeof(K) == 8 && L.size() >= 2 && !L[0].last && L[1].kb <= 32;
  u32* k32[2] = {nullptr, nullptr};  // the keys of levels l >= 1 are in k32[l & 1]
  if (narrow) {
    const u64 half = (n + 3) & ~3ull;  // the second array 16-byte aligned
    k32[1] = (u32*)kb_;
    k32[0] = (u32*)kb_ + half;
  }
  // The windows of level l (rows in the level's buffers kin0 / kin1 by
  // parity, keys of type KL): the bucket sort, the refused windows to the
  // radix tile sort (and the comparison rank), then the single-segment
  // windows.
  auto level_tiles = [&](auto* kin0, auto* kin1, size_t l, const DdLevel& d, u32* S, u32* cst, u64 bad,
                         u32* errp) -> int {
    using KL = typename std::remove_pointer<decltype(kin0)>::type;
    auto fn_bucket = T <= 512 ? k_tile_bucket<KL, 2> : (T <= 1024 ? k_tile_bucket<KL, 4> : k_tile_bucket<KL, 8>);
    auto fn_tsort = T <= 512 ? k_tile_sort<KL, 2> : (T <= 1024 ? k_tile_sort<KL, 4> : k_tile_sort<KL, 8>);
    auto fn_trank = T <= 512 ? k_tile_rank<KL, 2> : (T <= 1024 ? k_tile_rank<KL, 4> : k_tile_rank<KL, 8>);
    // single windows above T rows: the next tile size
    const u32 T2 = std::min<u32>(2 * T, (u32)TILESORT_TMAX);
    auto fn_bucket2 = T2 <= 512 ? k_tile_bucket<KL, 2> : (T2 <= 1024 ? k_tile_bucket<KL, 4> : k_tile_bucket<KL, 8>);
    auto fn_tsort2 = T2 <= 512 ? k_tile_sort<KL, 2> : (T2 <= 1024 ? k_tile_sort<KL, 4> : k_tile_sort<KL, 8>);
    auto fn_trank2 = T2 <= 512 ? k_tile_rank<KL, 2> : (T2 <= 1024 ? k_tile_rank<KL, 4> : k_tile_rank<KL, 8>);
    TileSortArgs<KL> a;
    std::memset(&a, 0, sizeof(a));
    a.kin[0] = kin0; a.vin[0] = vals; a.kin[1] = kin1; a.vin[1] = vb_;
    if constexpr (std::is_same<KL, K>::value) a.kout = keys_out ? keys : nullptr;
    a.vout = vals;
    a.tile_lo = tlo; a.tile_hi = thi; a.tile_s0 = ts0; a.tile_s1 = ts1; a.seg_start = cst;
    a.src_sel = (u32)(l & 1); a.kb = d.tkb; a.pad_seg = T; a.direct = 0; a.wlist = wlist;
    a.wcount = S + LV_WCOUNT; a.fblist = fblist; a.fbcount = S + LV_FB;
    a.size_lo = 0; a.size_hi = ~0u;
    a.bad = bad; a.err = errp;
    {
      const u64 g = bucket_grid(ctx, (const void*)fn_bucket, d.nw_max);
      // rows read (key, oid) and written (oid, and the key when wanted)
      const double per = keys_out ? 2.0 * (sizeof(KL) + 4) : sizeof(KL) + 8.0;
      TRY(klaunch(ctx, "k_tile_bucket", fn_bucket, g, 0.0, S + LV_ROWS_SMALL, per, a, fb, a.tile_lo, a.tile_hi,
                  a.tile_s0, a.tile_s1, a.wlist));
    }
    {
      // refused windows: the radix tile sort over the bits that vary inside
      // each; where even those leave no room for the segment index (keys
      // wider than 53 bits), the comparison rank
      TileSortArgs<KL> b = a;
      b.wcount = nullptr;
      b.rklist = rkl; b.rkcount = S + LV_RK;
      const u64 g = persist_grid(ctx, (const void*)fn_tsort, d.nw_max);
      TRY(klaunch(ctx, "k_tile_sort", fn_tsort, g, 0.0, nullptr, 0.0, b));
      if (!d.packable) {
        b.fblist = rkl; b.fbcount = S + LV_RK;
        TRY(klaunch(ctx, "k_tile_rank", fn_trank, g, 0.0, nullptr, 0.0, b));
      }
    }
    if (d.nsingle_max) {
      // the single-segment windows (in segment order): up to T rows in
      // T-row tiles, then the larger ones in 2T-row tiles; refused windows of
      // either size go to the 2T-row radix tile sort
      TileSortArgs<KL> c = a;
      c.tile_lo = slo; c.tile_hi = shi; c.tile_s0 = sseg; c.tile_s1 = sseg;
      c.wlist = nullptr; c.wcount = S + LV_NSINGLE; c.fblist = fblist2; c.fbcount = S + LV_FB2;
      c.size_lo = 0; c.size_hi = T;
      const u64 g = bucket_grid(ctx, (const void*)fn_bucket, d.nsingle_max);
      TRY(klaunch(ctx, "k_tile_bucket", fn_bucket, g, 0.0, nullptr, 0.0, c, fb, c.tile_lo, c.tile_hi, c.tile_s0,
                  c.tile_s1, c.wlist));
      if (d.single_max > T) {
        c.size_lo = T; c.size_hi = d.single_max;
        const u64 gb = bucket_grid(ctx, (const void*)fn_bucket2, d.nsingle_max);
        TRY(klaunch(ctx, "k_tile_bucket_2t", fn_bucket2, gb, 0.0, nullptr, 0.0, c, fb, c.tile_lo, c.tile_hi,
                    c.tile_s0, c.tile_s1, c.wlist));
      }
      c.wcount = nullptr;
      c.rklist = rkl2; c.rkcount = S + LV_RK2;
      const u64 g2 = persist_grid(ctx, (const void*)fn_tsort2, d.nsingle_max);
      TRY(klaunch(ctx, "k_tile_sort", fn_tsort2, g2, 0.0, nullptr, 0.0, c));
      if (!d.packable) {
        c.fblist = rkl2; c.fbcount = S + LV_RK2;
        TRY(klaunch(ctx, "k_tile_rank", fn_trank2, g2, 0.0, nullptr, 0.0, c));
      }
    }
    return FMCS_OK;
  };
  // The MSD level of level l over its large segments: keys of type KI from
  // skey into dkey (type KO: the keys' low bits when narrower), oids from
  // sval into dval; on an odd last level (every key bit consumed) the large
  // segments only move back to the final rows.
  auto level_msd = [&](const auto* skey, auto* dkey, const u32* sval, u32* dval, size_t l, const DdLevel& d, u32* S,
                       int cur, u64 bad, u32* errp) -> int {
    using KI = typename std::remove_const<typename std::remove_pointer<decltype(skey)>::type>::type;
    using KO = typename std::remove_pointer<decltype(dkey)>::type;
    if (d.last) {
      KI* kfin = nullptr;
      if constexpr (std::is_same<KI, K>::value) kfin = keys_out ? keys : nullptr;
      return klaunch(ctx, "kd_copy_segs", kd_copy_segs<KI>, cap_grid(d.ntiles_max, 2048), 0.0, S + LV_ROWS_LARGE,
                     2.0 * (sizeof(KI) + 4), skey, sval, kfin, vals, (const u32*)llo, (const u32*)lhi,
                     (const u32*)tbase, (const u32*)(S + LV_NLARGE), (const u32*)(S + LV_NTILES), (u32)WIDE_TILE);
    }
    if (l == 0 && dig0)
      // the first level's digits as bytes (written by the previous round's
      // last radix pass), read instead of the keys
      TRY(klaunch(ctx, "k_seg_hist", kd_seg_hist<uint8_t, WIDE_ITEMS, RADIX_BITS>, cap_grid(d.ntiles_max, 2048), 0.0,
                  S + LV_ROWS_LARGE, 1.0, dig0, (const u32*)llo, (const u32*)lhi, (const u32*)tbase,
                  (const u32*)(S + LV_NLARGE), (const u32*)(S + LV_NTILES), 0u, d.D, hist, tcount, 0ull,
                  (u32*)nullptr, tseg));
    else
      TRY(klaunch(ctx, "k_seg_hist", kd_seg_hist<KI, WIDE_ITEMS, RADIX_BITS>, cap_grid(d.ntiles_max, 2048), 0.0,
                  S + LV_ROWS_LARGE, (double)sizeof(KI), skey, (const u32*)llo, (const u32*)lhi,
                  (const u32*)tbase, (const u32*)(S + LV_NLARGE), (const u32*)(S + LV_NTILES), d.shift, d.D, hist,
                  tcount, bad, errp, tseg));
    const int nxt = cur ^ 1;
    TRY(klaunch(ctx, "k_digit_base", kd_digit_base<RADIX_BITS>, cap_grid(d.nlarge_max, 2048), 0.0, nullptr, 0.0,
                (const u32*)hist, (const u32*)llo, dbase, persist ? plist + loff[l + 1] : lst[nxt], lsz[nxt],
                (const u32*)(S + LV_NLARGE), l + 1 < L.size() ? S + DD_LVL_WORDS + LV_NSEG : nullptr));
    return scatter_pass<KI, RADIX_BITS, NoPay, false, false, WIDE_ITEMS, KO>(
        ctx, "k_onesweep_msd", skey, sval, dkey, dval, llo, lhi, tbase, S + LV_NLARGE, S + LV_NTILES, d.ntiles_max,
        dbase, tcount, d.shift, d.D, 0.0, S + LV_ROWS_LARGE, sizeof(KI) + 8.0 + sizeof(KO), (const NoPay*)nullptr,
        (NoPay*)nullptr, 0u, 0ull, (u32*)nullptr, (u32*)nullptr, (uint8_t*)nullptr, 0u, tseg);
  };
  TRY(klaunch(ctx, "k_offsets_to_list", kd_offsets_to_list, cap_grid((nseg_max + 255) / 256, 2048), 0.0, nullptr, 0.0,
              seg_off, nseg_d, persist ? plist : lst[0], lsz[0]));
  const u32* plo = nullptr;
  const u32* ptb = nullptr;
  for (size_t l = 0; l < L.size(); ++l) {
    const DdLevel& d = L[l];
    u32* S = lv + l * DD_LVL_WORDS;
    const u32* nseg_p = (l == 0) ? nseg_d : S + LV_NSEG;
    const int cur = (int)(l & 1);
    u32* cst = persist ? plist + loff[l] : lst[cur];
    u32* csz = lsz[cur];
    const u64 wo = persist ? woff[l] : 0, so = persist ? soff[l] : 0;
    TRY(klaunch(ctx, "k_windows", kd_windows, cap_grid((d.nseg_max + 2047) / 2048, 2048), 0.0, nullptr, 0.0,
                (const u32*)cst, (const u32*)csz, nseg_p, d.Wl, d.sm, plo, ptb, d.wpt, tlo + wo, thi + wo, ts0 + wo,
                ts1 + wo, wlist + wo, S + LV_WCOUNT, d.single_max, slo + so, shi + so, sseg + so, S + LV_NSINGLE,
                scan_st, scan_st2));
    // the large segments (above single_max rows) compacted, with their digit
    // counts zeroed on the levels that scatter; the row counts only feed the
    // per-kernel byte figures of a profiled run
    TRY(dd_scan(ctx, csz, nullptr, nseg_p, d.nseg_max, scan_st, S + LV_SCAN_A, S + LV_NLARGE, true, d.single_max,
                (ctx->prof && !ctx->capturing) ? S + LV_ROWS_SMALL : nullptr, cst, llo, lhi, ltiles,
                d.last ? nullptr : hist, WIDE_TILE));
    // level 0 reads every row (window rows in the tile sorts, large-segment
    // rows in its histogram), so it checks the declared width of every key
    const bool check0 = l == 0 && err && keybits < 8 * sizeof(K);
    const u64 bad = check0 ? ~((1ull << keybits) - 1) : 0ull;
    u32* errp = check0 ? err : nullptr;
    if (!persist) {
      const int rc = (narrow && l > 0) ? level_tiles(k32[0], k32[1], l, d, S, cst, bad, errp)
                                       : level_tiles(keys, kb_, l, d, S, cst, bad, errp);
      if (rc != FMCS_OK) return rc;
    }
    if (d.nlarge_max == 0) break;
    const u32* sval = (l & 1) ? vb_ : vals;
    u32* dval = (l & 1) ? vals : vb_;
    if (d.last && !(l & 1)) break;  // every key bit consumed; large segments already in place
    TRY(dd_scan(ctx, ltiles, tbase, S + LV_NLARGE, d.nlarge_max, scan_st2, S + LV_SCAN_B, S + LV_NTILES));
    // the level's MSD pass over its large segments: keys from the level's
    // buffer into the next level's (u32 from the first level on when narrowed)
    int rc;
    if (!narrow)
      rc = level_msd((l & 1) ? (const K*)kb_ : (const K*)keys, (l & 1) ? keys : kb_, sval, dval, l, d, S, cur, bad,
                     errp);
    else if (l == 0)
      rc = level_msd((const K*)keys, k32[1], sval, dval, l, d, S, cur, bad, errp);
    else
      rc = level_msd((const u32*)k32[l & 1], k32[(l & 1) ^ 1], sval, dval, l, d, S, cur, bad, errp);
    if (rc != FMCS_OK) return rc;
    if (d.last) break;
    plo = llo;
    ptb = tbase;
  }
  if (!persist) return FMCS_OK;
  // one launch for every level's single segments and windows
  auto fn_bins = k_seg_bins<K>;
  const u64 gbins = std::max<u64>(resident(ctx, (const void*)fn_bins) / XCDS, 1) * XCDS;
  u32* binsc;
  TRY(getbuf_t(ctx, WS_BINSC, gbins * BIN_SCRATCH, &binsc));
  BinArgs<K> b;
  std::memset(&b, 0, sizeof(b));
  b.kin[0] = keys; b.kin[1] = kb_; b.vin[0] = vals; b.vin[1] = vb_;
  b.kbuf[0] = keys; b.kbuf[1] = kb_; b.vbuf[0] = vals; b.vbuf[1] = vb_;
  b.kout = keys_out ? keys : nullptr; b.vout = vals;
  b.starts = plist; b.tlo = tlo; b.thi = thi; b.ts0 = ts0; b.ts1 = ts1; b.wlist = wlist;
  b.slo = slo; b.shi = shi; b.sseg = sseg; b.lv = lv; b.scratch = binsc;
  b.err = err;
  b.bad = (err && keybits < 8 * sizeof(K)) ? ~((1ull << keybits) - 1) : 0ull;
  b.nlev = (u32)L.size(); b.W = W; b.small_max = small_max;
  for (size_t l = 0; l < L.size(); ++l) {
    b.loff[l] = (u32)loff[l]; b.woff[l] = (u32)woff[l]; b.soff[l] = (u32)soff[l]; b.kbs[l] = L[l].tkb;
  }
  // refused windows (a bin above BUCKET_LIMIT rows, so at least
  // BUCKET_LIMIT + 1 rows each): two lists by source buffer, their counters
  // in the zeroed level record after the last level
  const u64 fbcap = n / (BUCKET_LIMIT + 1) + 1;
  u32 *fbw, *fbpool;
  TRY(getbuf_t(ctx, WS_FBWIN, 10 * fbcap, &fbw));
  TRY(getbuf_t(ctx, WS_FBPOOL, n, &fbpool));
  u32* fbcnt = lv + L.size() * DD_LVL_WORDS;  // [0, 2) refused, [2, 4) too wide for the radix tile sort
  b.fbpool = fbpool; b.fblo = fbw; b.fbhi = fbw + 2 * fbcap; b.fbs0 = fbw + 4 * fbcap; b.fbs1 = fbw + 6 * fbcap;
  b.fbcount = fbcnt; b.fbcap = (u32)fbcap;
  // every row ends in one window or single segment: read (key, oid), written
  // (oid, and the key when wanted)
  const double per = keys_out ? 2.0 * (sizeof(K) + 4) : sizeof(K) + 8.0;
  TRY(klaunch(ctx, "k_seg_bins", fn_bins, gbins, (double)n * per, nullptr, 0.0, b));
  // the refused windows: radix tile sort over the key bits that vary inside
  // each, the comparison rank where those leave no room for a segment index
  const u32 kbits = std::min<u32>(keybits, 8 * sizeof(K));
  for (u32 src = 0; src < 2; ++src) {
    TileSortArgs<K> f;
    std::memset(&f, 0, sizeof(f));
    f.kin[0] = keys; f.vin[0] = vals; f.kin[1] = kb_; f.vin[1] = vb_;
    f.kout = keys_out ? keys : nullptr; f.vout = vals;
    f.tile_lo = b.fblo + src * fbcap; f.tile_hi = b.fbhi + src * fbcap;
    f.tile_s0 = b.fbs0 + src * fbcap; f.tile_s1 = b.fbs1 + src * fbcap;
    f.seg_start = fbpool; f.src_sel = src; f.kb = kbits; f.pad_seg = BIN_T;
    f.fbcount = fbcnt + src;
    f.rklist = fbw + 8 * fbcap + src * fbcap; f.rkcount = fbcnt + 2 + src;
    const u64 g = persist_grid(ctx, (const void*)k_tile_sort<K, BIN_E>, fbcap);
    TRY(klaunch(ctx, "k_tile_sort", k_tile_sort<K, BIN_E>, g, 0.0, nullptr, 0.0, f));
    if (kbits + TS_SEGBITS > 64) {
      f.fblist = f.rklist; f.fbcount = f.rkcount;
      TRY(klaunch(ctx, "k_tile_rank", k_tile_rank<K, BIN_E>, g, 0.0, nullptr, 0.0, f));
    }
  }
  return FMCS_OK;
}

// Device-driven segmented sort with a host-known segment count (one level
// area; the count is written on the device, no synchronisation).
template <class K>
int seg_sort_dd_host(fmcs_ctx* ctx, K* keys, u32* vals, u64 n, u32 keybits, const u32* seg_off, u32 nseg) {
  if (n <= 1 || nseg == 0) return FMCS_OK;
  u32* dd;
  TRY(dd_get(ctx, &dd));
  TRY(dd_clear(ctx, dd, DD_HDR + (u64)DD_LVL_MAX * DD_LVL_WORDS));
  TRY(klaunch(ctx, "k_set", kd_set4, 1, 0.0, nullptr, 0.0, dd + 88, nseg, 0u, 0u, 0u));
  TRY(seg_sort_dd<K>(ctx, keys, vals, n, keybits, seg_off, dd + 88, nseg, dd + DD_HDR, true, dd + DD_ERR));
  if (ctx->prof) TRY(prof_flush(ctx));
  return FMCS_OK;
}

// out[i] = column[order[i]] (BATproject, src/composer.cpp:94-99): one pass,
// or passes over source windows of tune.gather_window_kb KiB.
template <class K>
int gather_launch(fmcs_ctx* ctx, const K* column, const u32* order, K* out, u64 n) {
  if (((uintptr_t)order & 15) != 0) return fail(ctx, FMCS_E_ARG, "order must be 16-byte aligned");
  const u64 grid = (n + 1023) / 1024;
  const u64 wrows = (u64)ctx->tune.gather_window_kb * 1024 / sizeof(K);
  if (wrows == 0 || wrows >= n)
    return klaunch(ctx, "k_gather", k_gather<K>, grid, (double)n * (4 + 2 * sizeof(K)), nullptr, 0.0, column, order,
                   out, n);
  for (u64 lo = 0; lo < n; lo += wrows) {
    const u64 len = std::min<u64>(wrows, n - lo);
    // each pass reads the whole order; the value reads and writes add up to one
    TRY(klaunch(ctx, "k_gather", k_gather_win<K>, grid, (double)n * 4 + (double)len * 2 * sizeof(K), nullptr, 0.0,
                column, order, out, n, (u32)lo, (u32)len));
  }
  return FMCS_OK;
}

// One round.  carried: this round's column already in the previous
// round's order (moved by that round's radix passes), sorted in place; else
// the column is gathered into WS_KA.  Round 0 moves pnext (the next round's
// column) into pout unless P is NoPay.
template <class K, class P = NoPay>
int dd_round(fmcs_ctx* ctx, u32 r, u32 keybits, u64 n, const K* column, u32* order, const u32* gin,
             const u32* soff, u32* gout, u32* oout, bool want_groups, u32* dd, K* carried = nullptr,
             const P* pnext = nullptr, P* pout = nullptr, bool check_keys = true, const uint8_t* dig0 = nullptr) {
  K* ka = carried;
  if (!ka) TRY(getbuf_t(ctx, WS_KA, n, &ka));
  if (r == 0) {
    // the identity order of round 1 (src/composer.cpp:60-63) comes from the first pass
    TRY((radix_sort_dd_t<K, RADIX_BITS, P>(ctx, column, nullptr, ka, order, n, keybits, dd, pnext, pout)));
  } else {
    if (!carried) TRY(gather_launch<K>(ctx, column, order, ka, n));
    u32* lv = dd + DD_HDR + (u64)r * DD_LVL_MAX * DD_LVL_WORDS;
    // the sorted keys feed only this round's boundary
    TRY(seg_sort_dd<K>(ctx, ka, order, n, keybits, soff, dd + (r - 1), n, lv, want_groups,
                       check_keys ? dd + DD_ERR : nullptr, dig0));
  }
  if (want_groups) TRY(boundary_dd<K>(ctx, ka, gin, n, gout, oout, dd + r));
  return FMCS_OK;
}

// Round 0 with the next round's column (bank pbank bytes) carried into pout.
template <class K>
int dd_round0_carry(fmcs_ctx* ctx, u32 keybits, u64 n, const K* column, u32* order, u32* gout, u32* oout,
                    bool want_groups, u32* dd, u32 pbank, const void* pnext, void* pout) {
  switch (pbank) {
    case 2:
      return dd_round<K, u16>(ctx, 0, keybits, n, column, order, nullptr, nullptr, gout, oout, want_groups, dd,
                              nullptr, (const u16*)pnext, (u16*)pout);
    case 4:
      return dd_round<K, u32>(ctx, 0, keybits, n, column, order, nullptr, nullptr, gout, oout, want_groups, dd,
                              nullptr, (const u32*)pnext, (u32*)pout);
    default:
      return dd_round<K, u64>(ctx, 0, keybits, n, column, order, nullptr, nullptr, gout, oout, want_groups, dd,
                              nullptr, (const u64*)pnext, (u64*)pout);
  }
}

// The whole multi-round sort, enqueued without a host synchronisation.
int mrs_dd(fmcs_ctx* ctx, u32 nrounds, const void* const* cols, const u32* banks, const u32* bw, u64 n, u32* order,
           u32* groups_d) {
  u32* dd;
  TRY(dd_get(ctx, &dd));
  TRY(dd_clear(ctx, dd, DD_WORDS));
  u32 *ga, *gb, *offa, *offb;
  TRY(getbuf_t(ctx, WS_HASH, n, &ga));
  TRY(getbuf_t(ctx, WS_HASH2, n, &gb));
  TRY(getbuf_t(ctx, WS_SEGOFF, n + 1, &offa));
  TRY(getbuf_t(ctx, WS_SEGOFF2, n + 1, &offb));
  // The first round's scatter passes can carry the second round's column
  // along: it then arrives in round-1 order with no gather (a coalesced
  // read and write per pass instead of one random read per row).  A u64
  // second column that fits beside a first column of at most 32 bits is
  // packed with it into one u64 composite key (carry_next 1).
  void* carried = nullptr;
  const bool fused = ctx->tune.carry_next == 1 && nrounds >= 2 && n > 1 && banks[0] <= 4 && banks[1] == 8 &&
                     bw[0] + bw[1] <= 64;
  if (ctx->tune.carry_next && nrounds >= 2 && n > 1) TRY(getbuf(ctx, WS_PAY, (size_t)n * banks[1], &carried));
  const u32* gin = nullptr;
  const u32* soff = nullptr;
  uint8_t* dig0 = nullptr;  // the packed form's round-2 first digits (see radix_sort_dd_fused)
  u32 dig_shift = 0;
  for (u32 r = 0; r < nrounds; ++r) {
    const bool last = (r + 1 == nrounds);
    const bool want_groups = !last || groups_d;
    u32* gout = last && groups_d ? groups_d : ((gin == ga) ? gb : ga);
    // dense ids of this round feed only the next round's boundary
    if (!last && !(r + 2 < nrounds || groups_d)) gout = nullptr;
    u32* oout = (soff == offa) ? offb : offa;
    void* kc = (r == 1) ? carried : nullptr;
    int rc;
    if (r == 0 && fused) {
      u64* comp = (u64*)carried;
      u32* side;
      TRY(getbuf_t(ctx, WS_KA, n, &side));
      // with two or more round-1 passes the last one also writes the second
      // round's first MSD digit of every row as a byte
      if (bw[0] > RADIX_BITS) {
        const u32 d0 = std::min<u32>(std::max<u32>(1, std::min<u32>(ctx->tune.msd_bits, RADIX_BITS)), bw[1]);
        TRY(getbuf_t(ctx, WS_DIG0, n, &dig0));
        dig_shift = bw[1] - d0;
      }
      rc = banks[0] == 2 ? radix_sort_dd_fused<u16>(ctx, (const u16*)cols[0], (const u64*)cols[1], bw[0], bw[1],
                                                    comp, order, side, n, dd, dig0, dig_shift)
                         : radix_sort_dd_fused<u32>(ctx, (const u32*)cols[0], (const u64*)cols[1], bw[0], bw[1],
                                                    comp, order, side, n, dd, dig0, dig_shift);
      // round-1 groups: the sorted c1 keys (one pass: the composites' c1 part)
      if (rc == FMCS_OK && want_groups)
        rc = bw[0] > RADIX_BITS ? boundary_dd<u32>(ctx, side, nullptr, n, gout, oout, dd + 0)
                                : boundary_dd<u64>(ctx, comp, nullptr, n, gout, oout, dd + 0, bw[1]);
    } else if (r == 0 && carried) {
      switch (banks[0]) {
        case 2: rc = dd_round0_carry<u16>(ctx, bw[0], n, (const u16*)cols[0], order, gout, oout, want_groups, dd, banks[1], cols[1], carried); break;
        case 4: rc = dd_round0_carry<u32>(ctx, bw[0], n, (const u32*)cols[0], order, gout, oout, want_groups, dd, banks[1], cols[1], carried); break;
        default: rc = dd_round0_carry<u64>(ctx, bw[0], n, (const u64*)cols[0], order, gout, oout, want_groups, dd, banks[1], cols[1], carried); break;
      }
    } else {
      // round 2 of the packed form sorts composites on their low bits (the
      // first column above them is no error; the second was checked when packed)
      const bool check = !(fused && r == 1);
      switch (banks[r]) {
        case 2: rc = dd_round<u16>(ctx, r, bw[r], n, (const u16*)cols[r], order, gin, soff, gout, oout, want_groups, dd, (u16*)kc, (const NoPay*)nullptr, (NoPay*)nullptr, check); break;
        case 4: rc = dd_round<u32>(ctx, r, bw[r], n, (const u32*)cols[r], order, gin, soff, gout, oout, want_groups, dd, (u32*)kc, (const NoPay*)nullptr, (NoPay*)nullptr, check); break;
        default: rc = dd_round<u64>(ctx, r, bw[r], n, (const u64*)cols[r], order, gin, soff, gout, oout, want_groups, dd, (u64*)kc, (const NoPay*)nullptr, (NoPay*)nullptr, check, r == 1 ? dig0 : nullptr); break;
      }
    }
    if (rc != FMCS_OK) return rc;
    gin = gout;
    soff = oout;
  }
  return FMCS_OK;
}

// Host view of the device-driven run: the group count and the key error.
int dd_finish(fmcs_ctx* ctx, u32 nrounds, bool have_groups, u32* ngroups) {
  u32* dd;
  TRY(getbuf_t(ctx, WS_DD, DD_WORDS, &dd));
  u32 h[DD_ERR + 1];
  TRY(readback(ctx, dd, DD_ERR + 1, h));
  ctx->dd_pending = false;
  if (h[DD_ERR]) {
    TRY(fill(ctx, dd + DD_ERR, 1, 0));  // reported once
    if (h[DD_ERR] & 2u) return fail(ctx, FMCS_E_HIP, "segmented sort: a block's segment stack overflowed");
    return fail(ctx, FMCS_E_ARG, "keys exceed keybits");
  }
  if (ngroups) *ngroups = have_groups ? h[nrounds - 1] : (nrounds >= 2 ? h[nrounds - 2] : 0u);
  return FMCS_OK;
}

void graph_drop(fmcs_ctx* ctx) {
  if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
  if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
  ctx->gexec = nullptr;
  ctx->graph = nullptr;
  ctx->gkey.clear();
}

// Replays the cached graph of this exact call, or runs the call eagerly and
// captures it for the next one.  The graph names workspace buffers, so the
// key holds the workspace epoch: any reallocation forces a new capture.
int mrs_graph(fmcs_ctx* ctx, u32 nrounds, const void* const* cols, const u32* banks, const u32* bw, u64 n, u32* order,
              u32* groups_d) {
  std::vector<u64> key = {(u64)nrounds, n, (u64)(uintptr_t)order, (u64)(uintptr_t)groups_d, ctx->tune.small_max,
                          ctx->tune.window, ctx->tune.msd_bits, ctx->tune.gather_window_kb, ctx->tune.carry_next,
                          ctx->tune.med_max,
                          (u64)(uintptr_t)ctx->stream};
  for (u32 r = 0; r < nrounds; ++r) {
    key.push_back((u64)(uintptr_t)cols[r]);
    key.push_back(((u64)banks[r] << 32) | bw[r]);
  }
  if (ctx->gexec && ctx->gkey.size() == key.size() + 1 && std::equal(key.begin(), key.end(), ctx->gkey.begin()) &&
      ctx->gkey.back() == ctx->ws_epoch) {
    HIPCHK(hipGraphLaunch(ctx->gexec, ctx->stream));
    return FMCS_OK;
  }
  graph_drop(ctx);
  TRY(mrs_dd(ctx, nrounds, cols, banks, bw, n, order, groups_d));  // allocates, computes this call's result
  if (!ctx->cap_stream) HIPCHK(hipStreamCreateWithFlags(&ctx->cap_stream, hipStreamNonBlocking));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  hipStream_t user = ctx->stream;
  ctx->stream = ctx->cap_stream;
  ctx->capturing = true;
  hipGraph_t g = nullptr;
  int rc = FMCS_OK;
  if (hipStreamBeginCapture(ctx->cap_stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    rc = fail(ctx, FMCS_E_HIP, "hipStreamBeginCapture failed");
  } else {
    rc = mrs_dd(ctx, nrounds, cols, banks, bw, n, order, groups_d);
    hipError_t e = hipStreamEndCapture(ctx->cap_stream, &g);
    if (rc == FMCS_OK && e != hipSuccess) rc = fail(ctx, FMCS_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  }
  ctx->capturing = false;
  ctx->stream = user;
  if (rc != FMCS_OK) {
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    return rc;
  }
  hipGraphExec_t ex = nullptr;
  if (hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGraphDestroy(g);
    return fail(ctx, FMCS_E_HIP, "hipGraphInstantiate failed");
  }
  ctx->graph = g;
  ctx->gexec = ex;
  ctx->gkey = key;
  ctx->gkey.push_back(ctx->ws_epoch);
  return FMCS_OK;
}

// hash_ids for composites (seg << 8 bank) | key of at most 64 bits (every
// round of 1-, 2- and 4-byte banks, and 8-byte banks without segments):
// the first position of every row's group (F) from LDS tables over chunks,
// in the local form when the chunks hold few distinct hashes, else
// partitioned (fmcs_hash.hip); then heads, their scan and the ids.
// The local form.  Its global table has room for 8x the sampled estimate
// of the chunk-local distinct hashes; a fuller table (an estimate too low)
// is reported by pass 1 and the pass runs again with room for every row.
template <class K>
int hash_first_local(fmcs_ctx* ctx, const K* keys, const u32* seg, u64 n, u32 kb, u64 distinct, u32* F) {
  u32* marker = ctx->small_d + 62;
  u32* overflow = ctx->small_d + 63;
  const u64 nch = (n + 2047) / 2048;
  auto kl = k_hash_local<K>;
  auto kf = k_hash_local_first<K>;
  const double rd = (double)n * (sizeof(K) + (seg ? 4 : 0));
  u64 want = std::max<u64>(4096, 8 * distinct);
  for (int attempt = 0; attempt < 2; ++attempt) {
    u64 tcap = 1;
    while (tcap < want) tcap <<= 1;
    u64* tkeys;
    u32* tfirst;
    TRY(getbuf_t(ctx, WS_HG_TKEY, tcap, &tkeys));
    TRY(getbuf_t(ctx, WS_HG_TFIRST, tcap, &tfirst));
    TRY(fill(ctx, (u32*)tkeys, 2 * tcap, 0xffffffffu));
    TRY(fill(ctx, tfirst, tcap, 0xffffffffu));
    TRY(fill(ctx, marker, 1, ~0u));
    TRY(fill(ctx, overflow, 1, 0));
    TRY(klaunch(ctx, "k_hash_local", kl, cap_grid(nch, 2 * resident(ctx, (const void*)kl)), rd, nullptr, 0.0, keys,
                seg, n, kb, tkeys, tfirst, tcap - 1, marker, ctx->small_d + 61, 1u, 1u, overflow));
    u32 over = 0;
    TRY(readback(ctx, overflow, 1, &over));
    if (!over)
      return klaunch(ctx, "k_hash_local_first", kf, cap_grid(nch, 2 * resident(ctx, (const void*)kf)), rd + 4.0 * n,
                     nullptr, 0.0, keys, seg, n, kb, (const u64*)tkeys, (const u32*)tfirst, tcap - 1,
                     (const u32*)marker, F);
    want = 2 * n;  // every row distinct fits
  }
  return fail(ctx, FMCS_E_HIP, "hash table overflow with room for every row");
}

template <class K>
int hash_first_partitioned(fmcs_ctx* ctx, const K* keys, const u32* seg, u64 n, u32 kb, u32* F) {
  // partitions of about 512 .. 1024 rows on average, whole 8-bit digits
  const u32 PB = n <= (1ull << 17) ? 8u : (n <= (1ull << 26) ? 16u : 24u);
  const u32 nparts = 1u << PB;
  const u32 pw = (u32)std::max<u64>(1, ((u64)nparts * 1024) / n);
  const u32 nwin = (nparts + pw - 1) / pw;
  u64 *hk, *hs;
  u32 *pos, *ps, *large, *pstart;
  TRY(getbuf_t(ctx, WS_HG_KEY, n, &hk));
  TRY(getbuf_t(ctx, WS_HG_POS, n, &pos));
  TRY(getbuf_t(ctx, WS_HG_SKEY, n, &hs));
  TRY(getbuf_t(ctx, WS_HG_SPOS, n, &ps));
  TRY(getbuf_t(ctx, WS_HG_PSTART, 2 * (u64)nwin, &large));
  TRY(getbuf_t(ctx, WS_HG_CHUNK, (u64)nparts + 1, &pstart));
  const double rd = (double)n * (sizeof(K) + (seg ? 4 : 0));
  TRY(klaunch(ctx, "k_hash_prep", k_hash_prep<K>, cap_grid((n + 255) / 256, 8192), rd + 12.0 * n, nullptr, 0.0, keys,
              seg, n, kb, hk, pos));
  // stable partition of (hash, position) on the hash's low PB bits
  TRY(radix_sort<u64>(ctx, hk, pos, hs, ps, n, PB, false));
  u32* nlarge_d = ctx->small_d + 61;
  u32* marker = ctx->small_d + 62;
  TRY(fill(ctx, nlarge_d, 1, 0));
  TRY(fill(ctx, marker, 1, ~0u));
  TRY(klaunch(ctx, "k_hash_pstart", k_hash_pstart, cap_grid(((u64)nparts + 256) / 256, 8192), 0.0, nullptr, 0.0,
              (const u64*)hs, (u32)n, nparts - 1, pstart));
  TRY(klaunch(ctx, "k_hash_window", k_hash_window, nwin, (double)n * (8 + 4 + 4), nullptr, 0.0, (const u64*)hs,
              (const u32*)ps, (const u32*)pstart, nparts, pw, F, large, nlarge_d));
  u32 nlarge = 0;
  TRY(readback(ctx, nlarge_d, 1, &nlarge));
  if (nlarge == 0) return FMCS_OK;
  // windows too large for LDS: chunks of their rows through one global table
  std::vector<u32> wl(2 * (size_t)nlarge);
  HIPCHK(hipMemcpyAsync(wl.data(), large, wl.size() * 4, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  std::vector<u32> cl;
  u64 rows = 0;
  for (u32 k = 0; k < nlarge; ++k) {
    const u32 lo = wl[2 * k], hi = wl[2 * k + 1];
    rows += hi - lo;
    for (u32 a = lo; a < hi; a += 2048) cl.push_back(a);
  }
  std::vector<u32> ch(cl.size());
  {
    size_t q = 0;
    for (u32 k = 0; k < nlarge; ++k)
      for (u32 a = wl[2 * k]; a < wl[2 * k + 1]; a += 2048) ch[q++] = std::min<u32>(wl[2 * k + 1], a + 2048);
  }
  const u32 nch = (u32)cl.size();
  u64 tcap = 1;
  while (tcap < 2 * rows) tcap <<= 1;
  u64* tkeys;
  u32 *tfirst, *cd;
  TRY(getbuf_t(ctx, WS_HG_TKEY, tcap, &tkeys));
  TRY(getbuf_t(ctx, WS_HG_TFIRST, tcap, &tfirst));
  TRY(getbuf_t(ctx, WS_HG_AUX, 2 * (u64)nch, &cd));  // F's heads come later; the slot is free until then
  HIPCHK(hipMemcpyAsync(cd, cl.data(), (size_t)nch * 4, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(cd + nch, ch.data(), (size_t)nch * 4, hipMemcpyHostToDevice, ctx->stream));
  TRY(fill(ctx, (u32*)tkeys, 2 * tcap, 0xffffffffu));
  TRY(fill(ctx, tfirst, tcap, 0xffffffffu));
  const u64 g = cap_grid(nch, 2 * resident(ctx, (const void*)k_hash_large));
  u32* overflow = ctx->small_d + 63;  // cannot be set: the table has room for every row
  TRY(fill(ctx, overflow, 1, 0));
  TRY(klaunch(ctx, "k_hash_large", k_hash_large, g, (double)rows * 12, nullptr, 0.0, (const u64*)hs, (const u32*)ps,
              (const u32*)cd, (const u32*)(cd + nch), nch, tkeys, tfirst, tcap - 1, marker, overflow));
  TRY(klaunch(ctx, "k_hash_large_first", k_hash_large_first, g, (double)rows * 16, nullptr, 0.0, (const u64*)hs,
              (const u32*)ps, (const u32*)cd, (const u32*)(cd + nch), nch, (const u64*)tkeys, (const u32*)tfirst,
              tcap - 1, (const u32*)marker, F));
  HIPCHK(hipStreamSynchronize(ctx->stream));  // the chunk lists above are host memory
  return FMCS_OK;
}

template <class K>
int hash_ids_packed(fmcs_ctx* ctx, const K* keys, const u32* seg, u64 n, u32* gid, u32* ng) {
  const u32 kb = 8 * sizeof(K);
  u32* F;
  TRY(getbuf_t(ctx, WS_HG_F, n, &F));
  // a sample of about 256 of the 2048-row chunks: the local form's global
  // work is the chunks' own distinct hashes, so few of them (at most an
  // eighth of the rows: values repeated many times within a chunk) -> the
  // local form, its table sized from the sample's distinct count (a table
  // too small is found and enlarged); else the partitioned one
  const u64 nch = (n + 2047) / 2048;
  const u32 stride = (u32)std::max<u64>(1, nch / 256);
  const u64 sampled = (nch + stride - 1) / stride;
  const u64 srows = std::min<u64>(n, sampled * 2048);
  u64 scap = 1;
  while (scap < 2 * srows) scap <<= 1;
  u64* stab;
  TRY(getbuf_t(ctx, WS_HG_TKEY, scap, &stab));
  TRY(fill(ctx, (u32*)stab, 2 * scap, 0xffffffffu));
  TRY(fill(ctx, ctx->small_d + 61, 2, 0));
  auto kl = k_hash_local<K>;
  TRY(klaunch(ctx, "k_hash_count", kl, cap_grid(sampled, 2 * resident(ctx, (const void*)kl)),
              (double)srows * (sizeof(K) + (seg ? 4 : 0)), nullptr, 0.0, keys, seg, n, kb, stab, (u32*)nullptr,
              scap - 1, (u32*)nullptr, ctx->small_d + 61, 0u, stride, (u32*)nullptr));
  u32 sample[2] = {0, 0};
  TRY(readback(ctx, ctx->small_d + 61, 2, sample));
  const u64 distinct = (u64)sample[1] + 1;
  if ((u64)sample[0] * 8 <= srows)
    TRY(hash_first_local<K>(ctx, keys, seg, n, kb, distinct, F));
  else
    TRY(hash_first_partitioned<K>(ctx, keys, seg, n, kb, F));
  const u64 grid = cap_grid((n + 255) / 256, 8192);
  u32* prefix;
  TRY(getbuf_t(ctx, WS_HG_POS, n, &prefix));  // the partition positions are dead by now
  // heads (F[p] == p) numbered by the scan itself
  TRY(scan_excl(ctx, F, prefix, n, ctx->small_d + 56, true));
  TRY(klaunch(ctx, "k_hash_assign_f", k_hash_assign_f, grid, (double)n * 12, nullptr, 0.0, (const u32*)F,
              (const u32*)prefix, n, gid));
  return readback(ctx, ctx->small_d + 56, 1, ng);
}

// First-appearance group ids of (seg[i], keys[i]) over rows in order,
// numbered 0.. across the whole array (segments contiguous or not).
int hash_ids(fmcs_ctx* ctx, u32 bank, const void* keys, const u32* seg, u64 n, u32* gid, u32* ng) {
  *ng = 0;
  if (n == 0) return FMCS_OK;
  if (!seg || bank <= 4) {
    switch (bank) {
      case 1: return hash_ids_packed<uint8_t>(ctx, (const uint8_t*)keys, seg, n, gid, ng);
      case 2: return hash_ids_packed<u16>(ctx, (const u16*)keys, seg, n, gid, ng);
      case 4: return hash_ids_packed<u32>(ctx, (const u32*)keys, seg, n, gid, ng);
      default: return hash_ids_packed<u64>(ctx, (const u64*)keys, seg, n, gid, ng);
    }
  }
  // wider composites: one open-addressing table over (segment, key)
  u64 cap = 1;
  while (cap < 2 * n) cap <<= 1;
  if (cap > (1ull << 31)) return fail(ctx, FMCS_E_UNSUPPORTED, "hash table too large");
  u32 *table, *first, *rank, *slot, *head, *prefix;
  TRY(getbuf_t(ctx, WS_HT_TABLE, cap, &table));
  TRY(getbuf_t(ctx, WS_HT_FIRST, cap, &first));
  TRY(getbuf_t(ctx, WS_HT_RANK, cap, &rank));
  TRY(getbuf_t(ctx, WS_HT_SLOT, n, &slot));
  TRY(getbuf_t(ctx, WS_HT_HEAD, n, &head));
  TRY(getbuf_t(ctx, WS_HT_PREFIX, n, &prefix));
  TRY(fill(ctx, table, cap, 0xffffffffu));
  TRY(fill(ctx, first, cap, 0xffffffffu));
  u32 blocks = (u32)std::min<u64>((n + 255) / 256, 16384);
  {
    Prof pf(ctx, "k_hash_insert", (double)n * (bank + (seg ? 4 : 0) + 4));
    switch (bank) {
      case 1: hipLaunchKernelGGL((k_hash_insert<uint8_t>), dim3(blocks), dim3(256), 0, ctx->stream, (const uint8_t*)keys, seg, n, table, (u32)(cap - 1), slot, first); break;
      case 2: hipLaunchKernelGGL((k_hash_insert<u16>), dim3(blocks), dim3(256), 0, ctx->stream, (const u16*)keys, seg, n, table, (u32)(cap - 1), slot, first); break;
      case 4: hipLaunchKernelGGL((k_hash_insert<u32>), dim3(blocks), dim3(256), 0, ctx->stream, (const u32*)keys, seg, n, table, (u32)(cap - 1), slot, first); break;
      default: hipLaunchKernelGGL((k_hash_insert<u64>), dim3(blocks), dim3(256), 0, ctx->stream, (const u64*)keys, seg, n, table, (u32)(cap - 1), slot, first); break;
    }
    TRY(launch_check(ctx, "k_hash_insert"));
  }
  hipLaunchKernelGGL(k_hash_heads, dim3(blocks), dim3(256), 0, ctx->stream, slot, first, n, head);
  TRY(launch_check(ctx, "k_hash_heads"));
  TRY(scan_excl(ctx, head, prefix, n, ctx->small_d + 56));
  hipLaunchKernelGGL(k_hash_rank_slots, dim3(blocks), dim3(256), 0, ctx->stream, slot, head, prefix, n, rank);
  TRY(launch_check(ctx, "k_hash_rank_slots"));
  hipLaunchKernelGGL(k_hash_assign, dim3(blocks), dim3(256), 0, ctx->stream, slot, rank, n, gid);
  TRY(launch_check(ctx, "k_hash_assign"));
  return readback(ctx, ctx->small_d + 56, 1, ng);
}

}  // namespace

// =========================================================================
// Services for the host composer layer
// =========================================================================
namespace fmcs {
namespace internal {

hipStream_t stream_of(fmcs_ctx* ctx) { return ctx->stream; }

int set_error(fmcs_ctx* ctx, int code, const char* msg) { return fail(ctx, code, msg); }

int scratch(fmcs_ctx* ctx, int slot, size_t bytes, void** out) {
  if (slot < 0 || slot >= USER_SLOTS) return fail(ctx, FMCS_E_ARG, "bad scratch slot");
  return getbuf(ctx, (Slot)(WS_USER0 + slot), bytes, out);
}

int fill_random(fmcs_ctx* ctx, void* p, uint32_t bank, uint64_t n, uint64_t seed, uint32_t bits) {
  if (n == 0) return FMCS_OK;
  u32 blocks = (u32)std::min<u64>((n + 255) / 256, 8192);
  switch (bank) {
    case 1: hipLaunchKernelGGL((k_fill_random<uint8_t>), dim3(blocks), dim3(256), 0, ctx->stream, (uint8_t*)p, n, seed, bits); break;
    case 2: hipLaunchKernelGGL((k_fill_random<u16>), dim3(blocks), dim3(256), 0, ctx->stream, (u16*)p, n, seed, bits); break;
    case 4: hipLaunchKernelGGL((k_fill_random<u32>), dim3(blocks), dim3(256), 0, ctx->stream, (u32*)p, n, seed, bits); break;
    case 8: hipLaunchKernelGGL((k_fill_random<u64>), dim3(blocks), dim3(256), 0, ctx->stream, (u64*)p, n, seed, bits); break;
    default: return fail(ctx, FMCS_E_ARG, "bad bank");
  }
  return launch_check(ctx, "k_fill_random");
}

int uniform_offsets(fmcs_ctx* ctx, uint32_t* off, uint32_t nseg, uint64_t n, uint32_t size) {
  u64 cnt = (u64)nseg + 1;
  u32 blocks = (u32)std::min<u64>((cnt + 255) / 256, 8192);
  hipLaunchKernelGGL(k_uniform_offsets, dim3(blocks), dim3(256), 0, ctx->stream, off, nseg, size, n);
  return launch_check(ctx, "k_uniform_offsets");
}

int scratch_of(fmcs_ctx* ctx, void** ptr) {
  if (!ctx->scratch_ptr) return fail(ctx, FMCS_E_HIP, "no scratch array on this device: " + ctx->scratch_err);
  *ptr = ctx->scratch_ptr;
  return FMCS_OK;
}

// fmcs_sort_pairs in place with every workspace slot carved out of
// [scratch, scratch + cap): no allocation, the same kernels and passes.
int sort_pairs_in_scratch(fmcs_ctx* ctx, uint32_t bank, uint32_t keybits, uint64_t n, void* keys, uint32_t* oids,
                          void* scratch, size_t cap) {
  ctx->arena = (char*)scratch;
  ctx->arena_cap = cap;
  ctx->arena_used = 0;
  for (int s = 0; s < WS_COUNT; ++s) ctx->arena_len[s] = 0;
  int rc;
  switch (bank) {
    case 2: rc = radix_sort<u16>(ctx, (u16*)keys, oids, (u16*)keys, oids, n, keybits); break;
    case 4: rc = radix_sort<u32>(ctx, (u32*)keys, oids, (u32*)keys, oids, n, keybits); break;
    default: rc = radix_sort<u64>(ctx, (u64*)keys, oids, (u64*)keys, oids, n, keybits); break;
  }
  ctx->arena = nullptr;
  ctx->arena_cap = 0;
  return rc;
}

int Stager::arg_void(void** p, uint64_t bytes, Dir dir, const char* name) {
  if (!*p || bytes == 0) return FMCS_OK;
  uint32_t kind = FMCS_MEM_OTHER;
  TRY(fmcs_pointer_kind(ctx_, *p, &kind));
  if (kind == FMCS_MEM_DEVICE || kind == FMCS_MEM_PINNED || kind == FMCS_MEM_MANAGED) return FMCS_OK;
  if (kind == FMCS_MEM_OTHER)
    return fail(ctx_, FMCS_E_ARG, std::string(name) + " is not memory this context's GPU can use");
  if (!allow_ || !ctx_->host_staging)
    return fail(ctx_, FMCS_E_ARG,
                std::string(name) + " is pageable host memory: pass device memory, or stage it with "
                                    "fmcs_ctx_set_host_staging" + (allow_ ? "" : " (not with this entry point)"));
  if (WS_STAGE0 + next_ >= WS_COUNT) return fail(ctx_, FMCS_E_UNSUPPORTED, "too many staged buffers in one call");
  void* dev;
  TRY(getbuf(ctx_, (Slot)(WS_STAGE0 + next_++), bytes, &dev));
  if (dir & IN) HIPCHK_C(ctx_, hipMemcpyAsync(dev, *p, bytes, hipMemcpyHostToDevice, ctx_->stream));
  if (dir & OUT) outs_.push_back({*p, dev, bytes});
  *p = dev;
  return FMCS_OK;
}

void Stager::set_out_bytes(const void* host, uint64_t bytes) {
  for (Out& o : outs_)
    if (o.host == host) o.bytes = std::min(o.bytes, bytes);
}

int Stager::finish() {
  if (outs_.empty()) return FMCS_OK;
  for (const Out& o : outs_)
    if (o.bytes) HIPCHK_C(ctx_, hipMemcpyAsync(o.host, o.dev, o.bytes, hipMemcpyDeviceToHost, ctx_->stream));
  HIPCHK_C(ctx_, hipStreamSynchronize(ctx_->stream));
  outs_.clear();
  return FMCS_OK;
}

}  // namespace internal
}  // namespace fmcs

// =========================================================================
// C ABI
// =========================================================================
// Message of the calling thread's last failed fmcs_ctx_create (no context
// exists to hold it): fmcs_last_error(NULL) returns it.
using fmcs::internal::Stager;

thread_local std::string g_create_err;
int create_fail(int code, const std::string& msg) {
  g_create_err = msg;
  return code;
}

namespace {
// fmcs_multi_round_sort after its buffer arguments are checked (and staged)
int mrs_body(fmcs_ctx* ctx, uint32_t nrounds, const void* const* cols_d, const uint32_t* banks,
             const uint32_t* bitwidths, uint64_t n, uint32_t* order_d, uint32_t* groups_d, int mode,
             uint32_t* ngroups);
}  // namespace

extern "C" {

int fmcs_ctx_create(int device, fmcs_ctx** out) {
  if (!out) return create_fail(FMCS_E_ARG, "null output pointer");
  *out = nullptr;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess) return create_fail(FMCS_E_HIP, std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
  if (ndev == 0) return create_fail(FMCS_E_HIP, "no GPU visible");
  if (device < 0 || device >= ndev) return create_fail(FMCS_E_ARG, "device index out of range");
  e = hipSetDevice(device);
  if (e != hipSuccess) return create_fail(FMCS_E_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
  fmcs_ctx* c = new fmcs_ctx();
  c->device = device;
  if (fmcs::internal::noalloc_setup(device, &c->scratch_ptr, &c->scratch_err) != FMCS_OK) c->scratch_ptr = nullptr;
  e = hipMalloc(&c->small_d, 64 * sizeof(u32));
  if (e != hipSuccess) {
    delete c;
    return create_fail(FMCS_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  e = hipHostMalloc((void**)&c->host_pin, PIN_WORDS * sizeof(u32), hipHostMallocDefault);
  if (e != hipSuccess) {
    (void)hipFree(c->small_d);
    delete c;
    return create_fail(FMCS_E_NOMEM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
  }
  for (auto& ev : c->ev) {
    e = hipEventCreate(&ev);
    if (e != hipSuccess) {
      fmcs_ctx_destroy(c);
      return create_fail(FMCS_E_HIP, std::string("hipEventCreate: ") + hipGetErrorString(e));
    }
  }
  *out = c;
  return FMCS_OK;
}

int fmcs_ctx_destroy(fmcs_ctx* ctx) {
  if (!ctx) return FMCS_E_ARG;
  (void)hipStreamSynchronize(ctx->stream);
  graph_drop(ctx);
  if (ctx->cap_stream) (void)hipStreamDestroy(ctx->cap_stream);
  for (auto& b : ctx->ws)
    if (b.p) (void)hipFree(b.p);
  (void)hipFree(ctx->small_d);
  (void)hipHostFree(ctx->host_pin);
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  prof_flush(ctx);
  for (auto& e : ctx->pool) (void)hipEventDestroy(e);
  delete ctx;
  return FMCS_OK;
}

int fmcs_ctx_set_stream(fmcs_ctx* ctx, void* stream) {
  if (!ctx) return FMCS_E_ARG;
  ctx->stream = (hipStream_t)stream;
  return FMCS_OK;
}

int fmcs_ctx_sync(fmcs_ctx* ctx) {
  if (!ctx) return FMCS_E_ARG;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (ctx->dd_pending) TRY(dd_finish(ctx, 1, true, nullptr));
  return FMCS_OK;
}

int fmcs_ctx_set_tuning(fmcs_ctx* ctx, const fmcs_tuning* t) {
  if (!ctx || !t) return FMCS_E_ARG;
  // small_max >= 64 bounds the device-driven path's segment lists by 256 / 65 entries per row
  if (t->small_max < 64 || t->small_max > (u32)TILESORT_TMAX / 2 || t->window < 1 || t->window > t->small_max + 1 ||
      t->msd_bits < 1 || t->msd_bits > (u32)RADIX_BITS || t->gather_window_kb > (1u << 24) || t->carry_next > 2 ||
      (t->med_max != 0 && (t->med_max < BIN_T || t->med_max > BIN_MED_MAX)))
    return fail(ctx, FMCS_E_ARG, "tuning out of range");
  ctx->tune = *t;
  return FMCS_OK;
}

int fmcs_ctx_get_tuning(fmcs_ctx* ctx, fmcs_tuning* t) {
  if (!ctx || !t) return FMCS_E_ARG;
  *t = ctx->tune;
  return FMCS_OK;
}

int fmcs_ctx_trim(fmcs_ctx* ctx) {
  if (!ctx) return FMCS_E_ARG;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  // an unreported device-driven key error goes with the word area: report it now
  const int pending = ctx->dd_pending ? dd_finish(ctx, 1, true, nullptr) : FMCS_OK;
  graph_drop(ctx);
  ctx->ws_epoch++;
  for (auto& b : ctx->ws) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
  }
  return pending;
}

const char* fmcs_last_error(fmcs_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_err.c_str(); }

uint32_t fmcs_classify_memory(int query_status, int memory_type, int owner_device, int ctx_device) {
  if (query_status != 0) return FMCS_MEM_HOST;  // unknown to the runtime: pageable host memory
  switch (memory_type) {
    case hipMemoryTypeUnregistered: return FMCS_MEM_HOST;
    case hipMemoryTypeHost: return FMCS_MEM_PINNED;
    case hipMemoryTypeDevice: return owner_device == ctx_device ? FMCS_MEM_DEVICE : FMCS_MEM_OTHER;
    case hipMemoryTypeManaged:
    case hipMemoryTypeUnified: return FMCS_MEM_MANAGED;
    default: return FMCS_MEM_OTHER;
  }
}

int fmcs_pointer_kind(fmcs_ctx* ctx, const void* p, uint32_t* kind) {
  if (!ctx || !kind) return FMCS_E_ARG;
  hipPointerAttribute_t a;
  std::memset(&a, 0, sizeof(a));
  const hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) (void)hipGetLastError();  // not an error of the call: the answer is "host"
  *kind = fmcs_classify_memory((int)e, (int)a.type, a.device, ctx->device);
  return FMCS_OK;
}

int fmcs_ctx_set_host_staging(fmcs_ctx* ctx, int on) {
  if (!ctx) return FMCS_E_ARG;
  ctx->host_staging = on != 0;
  return FMCS_OK;
}

uint32_t fmcs_bank_bytes(uint32_t width, int allow_u8) {
  if (allow_u8 && width <= 8) return 1;
  if (width <= 16) return 2;
  if (width <= 32) return 4;
  return 8;
}

int fmcs_construct_entries(uint32_t ncols, const uint32_t* bw, const uint32_t* col_order, uint32_t nsplit,
                           const uint32_t* split, fmcs_entry* entries, uint32_t max_entries, uint32_t* n_out) {
  if (!bw || !col_order || !split || !entries || !n_out || ncols == 0 || nsplit == 0) return FMCS_E_ARG;
  uint64_t so = 0, sn = 0;
  std::vector<char> seen(ncols, 0);
  for (u32 i = 0; i < ncols; ++i) {
    if (col_order[i] >= ncols || seen[col_order[i]]) return FMCS_E_ARG;
    seen[col_order[i]] = 1;
    if (bw[col_order[i]] == 0 || bw[col_order[i]] > 64) return FMCS_E_ARG;
    so += bw[col_order[i]];
  }
  for (u32 i = 0; i < nsplit; ++i) {
    if (split[i] == 0 || split[i] > 64) return FMCS_E_ARG;
    sn += split[i];
  }
  if (so != sn) return FMCS_E_ARG;
  u32 oi = 0, ni = 0, start = 1, n = 0;
  u32 ow = bw[col_order[0]], nw = split[0];
  while (ni < nsplit) {
    u32 m = std::min(ow, nw);
    ow -= m;
    nw -= m;
    if (n >= max_entries) return FMCS_E_ARG;
    entries[n++] = {ni, col_order[oi], start, start + m - 1};
    start += m;
    if (ow == 0) {
      ++oi;
      if (oi < ncols) ow = bw[col_order[oi]];
      start = 1;
    }
    if (nw == 0) {
      ++ni;
      if (ni < nsplit) nw = split[ni];
    }
  }
  *n_out = n;
  return FMCS_OK;
}

int fmcs_stitch(fmcs_ctx* ctx, uint64_t n, uint32_t nsrc, const void* const* src_d, const uint32_t* src_bank,
                const uint32_t* src_bw, uint32_t nent, const fmcs_entry* entries, uint32_t ndst, void* const* dst_d,
                const uint32_t* dst_bank) {
  return fmcs_stitch_dir(ctx, n, nsrc, src_d, src_bank, src_bw, nent, entries, ndst, dst_d, dst_bank, nullptr);
}

int fmcs_stitch_dir(fmcs_ctx* ctx, uint64_t n, uint32_t nsrc, const void* const* src_d, const uint32_t* src_bank,
                    const uint32_t* src_bw, uint32_t nent, const fmcs_entry* entries, uint32_t ndst,
                    void* const* dst_d, const uint32_t* dst_bank, const uint8_t* src_desc) {
  if (!ctx) return FMCS_E_ARG;
  if (nsrc == 0 || nsrc > STITCH_MAX_SRC || ndst == 0 || ndst > STITCH_MAX_DST || nent == 0 || nent > STITCH_MAX_ENT)
    return fail(ctx, FMCS_E_ARG, "stitch: too many columns or entries");
  if (!src_d || !src_bank || !src_bw || !entries || !dst_d || !dst_bank) return FMCS_E_ARG;
  for (u32 s = 0; s < nsrc; ++s)
    if (src_bank[s] > 8) return fail(ctx, FMCS_E_ARG, "stitch: bad source bank/width");
  for (u32 c = 0; c < ndst; ++c)
    if (dst_bank[c] > 8) return fail(ctx, FMCS_E_ARG, "stitch: destination too narrow");
  Stager st(ctx);
  const void* srcs[STITCH_MAX_SRC];
  void* dsts[STITCH_MAX_DST];
  for (u32 s = 0; s < nsrc; ++s) {
    srcs[s] = src_d[s];
    TRY(st.arg(srcs[s], n * src_bank[s], Stager::IN, "src_d"));
  }
  for (u32 c = 0; c < ndst; ++c) {
    dsts[c] = dst_d[c];
    TRY(st.arg(dsts[c], n * dst_bank[c], Stager::OUT, "dst_d"));
  }
  src_d = srcs;
  dst_d = dsts;
  StitchArgs a;
  std::memset(&a, 0, sizeof(a));
  a.n = n;
  a.ndst = ndst;
  for (u32 s = 0; s < nsrc; ++s) {
    if (!(src_bank[s] == 1 || bank_ok(src_bank[s])) || src_bw[s] == 0 || src_bw[s] > 8 * src_bank[s])
      return fail(ctx, FMCS_E_ARG, "stitch: bad source bank/width");
    if (((uintptr_t)src_d[s] & 15) != 0) return fail(ctx, FMCS_E_ARG, "stitch: source not 16-byte aligned");
    a.src[s] = src_d[s];
    a.src_bank[s] = src_bank[s];
    if (src_desc && src_desc[s]) a.invert |= 1u << s;
  }
  std::vector<u32> width(ndst, 0);
  u32 prev = 0;
  for (u32 e = 0; e < nent; ++e) {
    const fmcs_entry& x = entries[e];
    if (x.newcol >= ndst || x.src >= nsrc || x.start < 1 || x.end < x.start || x.end > src_bw[x.src] ||
        (e && x.newcol < prev))
      return fail(ctx, FMCS_E_ARG, "stitch: bad entry");
    prev = x.newcol;
    a.e_src[e] = (uint8_t)x.src;
    a.e_width[e] = (uint8_t)(x.end - x.start + 1);
    a.e_shift[e] = (uint8_t)(src_bw[x.src] - x.end);
    width[x.newcol] += x.end - x.start + 1;
  }
  u32 e = 0;
  for (u32 c = 0; c < ndst; ++c) {
    if (!(dst_bank[c] == 1 || bank_ok(dst_bank[c])) || width[c] == 0 || width[c] > 8 * dst_bank[c])
      return fail(ctx, FMCS_E_ARG, "stitch: destination too narrow");
    if (((uintptr_t)dst_d[c] & 15) != 0) return fail(ctx, FMCS_E_ARG, "stitch: destination not 16-byte aligned");
    a.dst[c] = dst_d[c];
    a.dst_bank[c] = dst_bank[c];
    a.ent_begin[c] = e;
    while (e < nent && entries[e].newcol == c) ++e;
  }
  a.ent_begin[ndst] = e;
  if (n == 0) return FMCS_OK;
  u64 threads = (n + 3) / 4;
  u32 blocks = (u32)((threads + 255) / 256);
  {
    // every source column that an entry names read once, every stitched column written once
    u64 used = 0;
    for (u32 i = 0; i < nent; ++i) used |= 1ull << entries[i].src;
    double row_bytes = 0;
    for (u32 s = 0; s < nsrc; ++s) row_bytes += ((used >> s) & 1) ? src_bank[s] : 0;
    for (u32 c = 0; c < ndst; ++c) row_bytes += dst_bank[c];
    Prof pf(ctx, "k_stitch", row_bytes * n);
    hipLaunchKernelGGL(k_stitch, dim3(blocks), dim3(256), 0, ctx->stream, a);
    TRY(launch_check(ctx, "k_stitch"));
  }
  return st.finish();
}

int fmcs_gather(fmcs_ctx* ctx, uint32_t bank, uint64_t n, const void* values_d, const uint32_t* order_d,
                void* out_d) {
  if (!ctx) return FMCS_E_ARG;
  if (!bank_ok(bank)) return fail(ctx, FMCS_E_ARG, "gather: bank must be 2, 4 or 8");
  if (n == 0) return FMCS_OK;
  Stager st(ctx);
  TRY(st.arg(values_d, n * bank, Stager::IN, "values_d"));
  TRY(st.arg(order_d, n * 4, Stager::IN, "order_d"));
  TRY(st.arg(out_d, n * bank, Stager::OUT, "out_d"));
  if (((uintptr_t)order_d & 15) != 0) return fail(ctx, FMCS_E_ARG, "gather: order must be 16-byte aligned");
  switch (bank) {
    case 2: TRY(gather_launch<u16>(ctx, (const u16*)values_d, order_d, (u16*)out_d, n)); break;
    case 4: TRY(gather_launch<u32>(ctx, (const u32*)values_d, order_d, (u32*)out_d, n)); break;
    default: TRY(gather_launch<u64>(ctx, (const u64*)values_d, order_d, (u64*)out_d, n)); break;
  }
  return st.finish();
}

int fmcs_gather_columns(fmcs_ctx* ctx, uint64_t n, uint32_t ncols, const void* const* cols_d, const uint32_t* banks,
                        const uint32_t* bits, const uint32_t* order_d, void* const* outs_d) {
  if (!ctx) return FMCS_E_ARG;
  if (ncols == 0 || ncols > 32 || !cols_d || !banks || !outs_d)
    return fail(ctx, FMCS_E_ARG, "gather_columns: 1 to 32 columns with their banks and outputs");
  for (u32 k = 0; k < ncols; ++k) {
    if (!bank_ok(banks[k])) return fail(ctx, FMCS_E_ARG, "gather_columns: banks must be 2, 4 or 8");
    if (bits && (bits[k] == 0 || bits[k] > 8 * banks[k]))
      return fail(ctx, FMCS_E_ARG, "gather_columns: bits must be 1 .. 8 x bank");
  }
  if (n == 0) return FMCS_OK;
  if (n > (1ull << 32)) return fail(ctx, FMCS_E_ARG, "gather_columns: more than 2^32 rows");
  Stager st(ctx);
  std::vector<const void*> src(cols_d, cols_d + ncols);
  std::vector<void*> dst(outs_d, outs_d + ncols);
  for (u32 k = 0; k < ncols; ++k) {
    TRY(st.arg(src[k], n * banks[k], Stager::IN, "cols_d"));
    TRY(st.arg(dst[k], n * banks[k], Stager::OUT, "outs_d"));
  }
  TRY(st.arg(order_d, n * 4, Stager::IN, "order_d"));
  if (((uintptr_t)order_d & 15) != 0) return fail(ctx, FMCS_E_ARG, "gather_columns: order must be 16-byte aligned");
  // Columns widest first, first fit into records: of at most 16 bytes
  // (each column at an offset aligned to its width), or, with declared
  // widths, of at most 64 bits.  Bit records are taken when they need no
  // more records a row than byte records (they are half the size or less).
  auto pack = [&](const uint32_t* w, u32 cap) {
    std::vector<u32> idx(ncols);
    for (u32 k = 0; k < ncols; ++k) idx[k] = k;
    std::stable_sort(idx.begin(), idx.end(), [&](u32 a, u32 b) { return w[a] > w[b]; });
    std::vector<std::vector<u32>> gs;
    std::vector<u32> fill;
    for (u32 k : idx) {
      size_t g = 0;
      while (g < gs.size() && (fill[g] + w[k] > cap || gs[g].size() == 8)) ++g;
      if (g == gs.size()) {
        gs.emplace_back();
        fill.push_back(0);
      }
      gs[g].push_back(k);
      fill[g] += w[k];
    }
    return gs;
  };
  std::vector<std::vector<u32>> groups = pack(banks, 16);
  bool use_bits = false;
  if (bits) {
    std::vector<std::vector<u32>> bg = pack(bits, 64);
    bool multi = false;
    for (auto& g : bg) multi = multi || g.size() > 1;
    if (multi && bg.size() <= groups.size()) {
      groups.swap(bg);
      use_bits = true;
    }
  }
  const u64 grid = (n + 1023) / 1024, ggrid = (n + 256 * REC_ROWS - 1) / (256 * REC_ROWS);
  auto column = [&](u32 k) -> int {
    switch (banks[k]) {
      case 2: return gather_launch<u16>(ctx, (const u16*)src[k], order_d, (u16*)dst[k], n);
      case 4: return gather_launch<u32>(ctx, (const u32*)src[k], order_d, (u32*)dst[k], n);
      default: return gather_launch<u64>(ctx, (const u64*)src[k], order_d, (u64*)dst[k], n);
    }
  };
  // a record group: two or more columns, all 16-byte aligned (else the
  // column gather takes them one by one)
  auto recordable = [&](const std::vector<u32>& gr) {
    bool a = gr.size() > 1;
    for (u32 k : gr) a = a && ((uintptr_t)src[k] & 15) == 0 && ((uintptr_t)dst[k] & 15) == 0;
    return a;
  };
  auto describe = [&](const std::vector<u32>& gr, RecCols& c) -> u32 {
    std::memset(&c, 0, sizeof(c));
    u32 off = 0;
    for (u32 k : gr) {
      c.src[c.ncol] = src[k];
      c.dst[c.ncol] = dst[k];
      c.width[c.ncol] = banks[k];
      c.bits[c.ncol] = use_bits ? bits[k] : 8 * banks[k];
      c.off[c.ncol] = off;
      off += use_bits ? bits[k] : banks[k];
      ++c.ncol;
    }
    return off;
  };
  if (use_bits) {
    // every bit group is packed, and the declared widths checked, before
    // any output is written
    std::vector<size_t> bg;
    for (size_t g = 0; g < groups.size(); ++g)
      if (recordable(groups[g])) bg.push_back(g);
    if (!bg.empty()) {
      u64* rec;
      u32* err;
      TRY(getbuf_t(ctx, WS_REC, n * bg.size(), &rec));
      TRY(getbuf_t(ctx, WS_RECERR, 1, &err));
      TRY(fill(ctx, err, 1, 0u));
      std::vector<RecCols> cs(bg.size());
      for (size_t i = 0; i < bg.size(); ++i) {
        describe(groups[bg[i]], cs[i]);
        double w = 0;
        for (u32 k : groups[bg[i]]) w += banks[k];
        TRY(klaunch(ctx, "k_bitrec_pack", k_bitrec_pack, (n + 2047) / 2048, (double)n * (w + 8), nullptr, 0.0, cs[i], rec + i * n,
                    n, err));
      }
      u32 h = 0;
      HIPCHK(hipMemcpyAsync(&h, err, sizeof(u32), hipMemcpyDeviceToHost, ctx->stream));
      HIPCHK(hipStreamSynchronize(ctx->stream));
      if (h) return fail(ctx, FMCS_E_ARG, "gather_columns: a value is wider than its declared bits");
      for (size_t i = 0; i < bg.size(); ++i) {
        double w = 0;
        for (u32 k : groups[bg[i]]) w += banks[k];
        TRY(klaunch(ctx, "k_bitrec_gather", k_bitrec_gather, grid, (double)n * (4 + 8 + w), nullptr, 0.0, cs[i],
                    (const u64*)(rec + i * n), order_d, n));
      }
    }
    for (size_t g = 0; g < groups.size(); ++g)
      if (!recordable(groups[g]))
        for (u32 k : groups[g]) TRY(column(k));
    return st.finish();
  }
  for (size_t g = 0; g < groups.size(); ++g) {
    if (!recordable(groups[g])) {
      for (u32 k : groups[g]) TRY(column(k));
      continue;
    }
    RecCols c;
    const u32 off = describe(groups[g], c);
    const u32 R = off <= 8 ? 8 : 16;
    void* rec;
    TRY(getbuf(ctx, WS_REC, n * R, &rec));
    if (R == 8) {
      TRY(klaunch(ctx, "k_rec_pack", k_rec_pack<8>, grid, (double)n * (off + 8), nullptr, 0.0, c, rec, n));
      TRY(klaunch(ctx, "k_rec_gather", k_rec_gather<8>, ggrid, (double)n * (4 + 8 + off), nullptr, 0.0, c,
                  (const void*)rec, order_d, n));
    } else {
      TRY(klaunch(ctx, "k_rec_pack", k_rec_pack<16>, grid, (double)n * (off + 16), nullptr, 0.0, c, rec, n));
      TRY(klaunch(ctx, "k_rec_gather", k_rec_gather<16>, ggrid, (double)n * (4 + 16 + off), nullptr, 0.0, c,
                  (const void*)rec, order_d, n));
    }
  }
  return st.finish();
}

int fmcs_extract_group_info(fmcs_ctx* ctx, uint32_t bank, uint64_t n, const void* sorted_d, const uint32_t* in_group_d,
                            uint32_t* out_group_d, uint32_t* seg_offsets_d, uint32_t* ngroups) {
  if (!ctx || !ngroups) return FMCS_E_ARG;
  if (!bank_ok(bank)) return fail(ctx, FMCS_E_ARG, "bank must be 2, 4 or 8");
  Stager st(ctx);
  uint32_t* const off_caller = seg_offsets_d;
  TRY(st.arg(sorted_d, n * bank, Stager::IN, "sorted_d"));
  TRY(st.arg(in_group_d, n * 4, Stager::IN, "in_group_d"));
  TRY(st.arg(out_group_d, n * 4, Stager::OUT, "out_group_d"));
  TRY(st.arg(seg_offsets_d, (n + 1) * 4, Stager::OUT, "seg_offsets_d"));
  switch (bank) {
    case 2: TRY(boundary<u16>(ctx, (const u16*)sorted_d, in_group_d, n, out_group_d, seg_offsets_d, ngroups)); break;
    case 4: TRY(boundary<u32>(ctx, (const u32*)sorted_d, in_group_d, n, out_group_d, seg_offsets_d, ngroups)); break;
    default: TRY(boundary<u64>(ctx, (const u64*)sorted_d, in_group_d, n, out_group_d, seg_offsets_d, ngroups)); break;
  }
  st.set_out_bytes(off_caller, ((u64)*ngroups + 1) * 4);  // the caller's array holds ngroups + 1 words
  return st.finish();
}

int fmcs_sort_pairs(fmcs_ctx* ctx, uint32_t bank, uint32_t keybits, uint64_t n, void* keys_d, uint32_t* oids_d) {
  if (!ctx) return FMCS_E_ARG;
  if (!bank_ok(bank)) return fail(ctx, FMCS_E_ARG, "bank must be 2, 4 or 8");
  if (keybits == 0 || keybits > 8 * bank) return fail(ctx, FMCS_E_ARG, "bad keybits");
  Stager st(ctx);
  TRY(st.arg(keys_d, n * bank, Stager::INOUT, "keys_d"));
  TRY(st.arg(oids_d, n * 4, Stager::INOUT, "oids_d"));
  switch (bank) {
    case 2: TRY(radix_sort<u16>(ctx, (u16*)keys_d, oids_d, (u16*)keys_d, oids_d, n, keybits)); break;
    case 4: TRY(radix_sort<u32>(ctx, (u32*)keys_d, oids_d, (u32*)keys_d, oids_d, n, keybits)); break;
    default: TRY(radix_sort<u64>(ctx, (u64*)keys_d, oids_d, (u64*)keys_d, oids_d, n, keybits)); break;
  }
  return st.finish();
}

int fmcs_fill_column(fmcs_ctx* ctx, uint32_t bank, uint64_t n, uint64_t seed, uint64_t card, int zipf,
                     uint32_t width, void* out_d) {
  if (!ctx || !out_d) return FMCS_E_ARG;
  if (!bank_ok(bank)) return fail(ctx, FMCS_E_ARG, "bank must be 2, 4 or 8");
  if (width == 0 || width > 8 * bank || card == 0) return fail(ctx, FMCS_E_ARG, "fill_column: bad width or card");
  if (n == 0) return FMCS_OK;
  Stager st(ctx);
  TRY(st.arg(out_d, n * bank, Stager::OUT, "out_d"));
  const u64 mask = width >= 64 ? ~0ull : ((1ull << width) - 1);
  const u64 span = width >= 64 ? ~0ull : (1ull << width);
  const u64 step = std::max<u64>(1, span / card);
  const double lc = std::log2((double)card);
  const u64 grid = cap_grid((n + 255) / 256, 8192);
  switch (bank) {
    case 2: TRY(klaunch(ctx, "k_fill_column", k_fill_column<u16>, grid, 0.0, nullptr, 0.0, (u16*)out_d, n, seed, card, (u32)(zipf != 0), lc, step, mask)); break;
    case 4: TRY(klaunch(ctx, "k_fill_column", k_fill_column<u32>, grid, 0.0, nullptr, 0.0, (u32*)out_d, n, seed, card, (u32)(zipf != 0), lc, step, mask)); break;
    default: TRY(klaunch(ctx, "k_fill_column", k_fill_column<u64>, grid, 0.0, nullptr, 0.0, (u64*)out_d, n, seed, card, (u32)(zipf != 0), lc, step, mask)); break;
  }
  return st.finish();
}

int fmcs_sort_pairs_merge(fmcs_ctx* ctx, uint32_t bank, uint64_t n, void* keys_d, uint32_t* oids_d) {
  if (!ctx) return FMCS_E_ARG;
  if (!bank_ok(bank)) return fail(ctx, FMCS_E_ARG, "bank must be 2, 4 or 8");
  Stager st(ctx);
  TRY(st.arg(keys_d, n * bank, Stager::INOUT, "keys_d"));
  TRY(st.arg(oids_d, n * 4, Stager::INOUT, "oids_d"));
  switch (bank) {
    case 2: TRY(merge_sort<u16>(ctx, (u16*)keys_d, oids_d, (u16*)keys_d, oids_d, n)); break;
    case 4: TRY(merge_sort<u32>(ctx, (u32*)keys_d, oids_d, (u32*)keys_d, oids_d, n)); break;
    default: TRY(merge_sort<u64>(ctx, (u64*)keys_d, oids_d, (u64*)keys_d, oids_d, n)); break;
  }
  return st.finish();
}

int fmcs_segmented_sort(fmcs_ctx* ctx, uint32_t bank, uint32_t keybits, uint64_t n, void* keys_d, uint32_t* oids_d,
                        const uint32_t* seg_offsets_d, uint32_t nseg, int mode, fmcs_group_info* info) {
  if (!ctx) return FMCS_E_ARG;
  if (!bank_ok(bank)) return fail(ctx, FMCS_E_ARG, "bank must be 2, 4 or 8");
  if (keybits == 0 || keybits > 8 * bank) return fail(ctx, FMCS_E_ARG, "bad keybits");
  if (n >= (1ull << 30)) return fail(ctx, FMCS_E_UNSUPPORTED, "n < 2^30 required");
  Stager st(ctx);
  TRY(st.arg(keys_d, n * bank, Stager::INOUT, "keys_d"));
  TRY(st.arg(oids_d, n * 4, Stager::INOUT, "oids_d"));
  TRY(st.arg(seg_offsets_d, ((u64)nseg + 1) * 4, Stager::IN, "seg_offsets_d"));
  if (info) TRY(group_info_from_offsets(ctx, bank, seg_offsets_d, nseg, info));
  const u32* grp = nullptr;
  if (mode == FMCS_MODE_UNBINNED) {
    // group ids from the offsets: ids of a run of equal ids are its heads
    u32* g;
    TRY(getbuf_t(ctx, WS_GROUPTMP, n, &g));
    std::vector<u32> off((size_t)nseg + 1);
    HIPCHK(hipMemcpyAsync(off.data(), seg_offsets_d, ((size_t)nseg + 1) * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    std::vector<u32> hg(n);
    for (u32 s = 0; s < nseg; ++s)
      for (u32 i = off[s]; i < off[s + 1] && i < n; ++i) hg[i] = s;
    HIPCHK(hipMemcpyAsync(g, hg.data(), n * 4, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    grp = g;
  }
  switch (bank) {
    case 2: TRY(seg_sort<u16>(ctx, (u16*)keys_d, oids_d, n, keybits, seg_offsets_d, nseg, mode, grp)); break;
    case 4: TRY(seg_sort<u32>(ctx, (u32*)keys_d, oids_d, n, keybits, seg_offsets_d, nseg, mode, grp)); break;
    default: TRY(seg_sort<u64>(ctx, (u64*)keys_d, oids_d, n, keybits, seg_offsets_d, nseg, mode, grp)); break;
  }
  // the device-driven forms report keys above keybits through the context's
  // error word, read by the next call that returns a count or by fmcs_ctx_sync
  if ((mode == FMCS_MODE_DEVICE || mode == FMCS_MODE_GRAPH) && n > 1 && nseg) ctx->dd_pending = true;
  return st.finish();
}

int fmcs_sort_by_group(fmcs_ctx* ctx, uint32_t bank, uint32_t keybits, uint64_t n, const void* column_d,
                       uint32_t* order_d, const uint32_t* in_group_d, uint32_t* out_group_d, int last_round, int mode,
                       void* sorted_d, fmcs_group_info* info, uint32_t* ngroups) {
  if (!ctx) return FMCS_E_ARG;
  if (keybits == 0 || keybits > 8 * bank) return fail(ctx, FMCS_E_ARG, "bad keybits");
  if (!last_round && !out_group_d) return fail(ctx, FMCS_E_ARG, "out_group required unless last round");
  if (n >= (1ull << 30)) return fail(ctx, FMCS_E_UNSUPPORTED, "n < 2^30 required");
  if (!bank_ok(bank)) return fail(ctx, FMCS_E_ARG, "bank must be 2, 4 or 8");
  // the reference's buffers (src/composer.cpp:1222-1261): the column and
  // the previous groups are read, the order is read and rewritten, the new
  // groups (and the sorted values) are written
  Stager st(ctx);
  TRY(st.arg(column_d, n * bank, Stager::IN, "column_d"));
  TRY(st.arg(order_d, n * 4, Stager::INOUT, "order_d"));
  TRY(st.arg(in_group_d, n * 4, Stager::IN, "in_group_d"));
  TRY(st.arg(out_group_d, n * 4, Stager::OUT, "out_group_d"));
  TRY(st.arg(sorted_d, n * bank, Stager::OUT, "sorted_d"));
  switch (bank) {
    case 2: TRY(sort_by_group_t<u16>(ctx, keybits, n, (const u16*)column_d, order_d, in_group_d, out_group_d, last_round, mode, (u16*)sorted_d, info, ngroups, nullptr, 0, nullptr, nullptr, nullptr, nullptr)); break;
    case 4: TRY(sort_by_group_t<u32>(ctx, keybits, n, (const u32*)column_d, order_d, in_group_d, out_group_d, last_round, mode, (u32*)sorted_d, info, ngroups, nullptr, 0, nullptr, nullptr, nullptr, nullptr)); break;
    default: TRY(sort_by_group_t<u64>(ctx, keybits, n, (const u64*)column_d, order_d, in_group_d, out_group_d, last_round, mode, (u64*)sorted_d, info, ngroups, nullptr, 0, nullptr, nullptr, nullptr, nullptr)); break;
  }
  // a later round's device-driven segmented sort reports keys above keybits
  // through the error word (the next count readback or fmcs_ctx_sync)
  if (in_group_d && n > 1 && (mode == FMCS_MODE_DEVICE || mode == FMCS_MODE_GRAPH)) ctx->dd_pending = true;
  return st.finish();
}

int fmcs_multi_round_sort(fmcs_ctx* ctx, uint32_t nrounds, const void* const* cols_d, const uint32_t* banks,
                          const uint32_t* bitwidths, uint64_t n, uint32_t* order_d, uint32_t* groups_d, int mode,
                          uint32_t* ngroups) {
  if (!ctx || !cols_d || !banks || !bitwidths || !order_d || nrounds == 0) return FMCS_E_ARG;
  if (n >= (1ull << 30)) return fail(ctx, FMCS_E_UNSUPPORTED, "n < 2^30 required");
  for (u32 r = 0; r < nrounds; ++r)
    if (!bank_ok(banks[r]) || bitwidths[r] == 0 || bitwidths[r] > 8 * banks[r])
      return fail(ctx, FMCS_E_ARG, "bad bank/bitwidth");
  if (n == 0) return FMCS_OK;
  // the reference's column banks (src/composer.cpp:48-135) are read; the
  // final order and groups are written
  Stager st(ctx);
  std::vector<const void*> cols(cols_d, cols_d + nrounds);
  for (u32 r = 0; r < nrounds; ++r) TRY(st.arg(cols[r], n * banks[r], Stager::IN, "cols_d"));
  TRY(st.arg(order_d, n * 4, Stager::OUT, "order_d"));
  TRY(st.arg(groups_d, n * 4, Stager::OUT, "groups_d"));
  TRY(mrs_body(ctx, nrounds, cols.data(), banks, bitwidths, n, order_d, groups_d, mode, ngroups));
  return st.finish();
}

}  // extern "C"

namespace {

int mrs_body(fmcs_ctx* ctx, uint32_t nrounds, const void* const* cols_d, const uint32_t* banks,
             const uint32_t* bitwidths, uint64_t n, uint32_t* order_d, uint32_t* groups_d, int mode,
             uint32_t* ngroups) {
  if (mode == FMCS_MODE_DEVICE || mode == FMCS_MODE_GRAPH) {
    if (nrounds > DD_ROUNDS_MAX) return fail(ctx, FMCS_E_UNSUPPORTED, "device-driven path supports at most 64 rounds");
    if (((uintptr_t)order_d & 15) != 0) return fail(ctx, FMCS_E_ARG, "order must be 16-byte aligned");
    if (mode == FMCS_MODE_DEVICE) {
      TRY(mrs_dd(ctx, nrounds, cols_d, banks, bitwidths, n, order_d, groups_d));
      if (ctx->prof) TRY(prof_flush(ctx));
    } else {
      TRY(mrs_graph(ctx, nrounds, cols_d, banks, bitwidths, n, order_d, groups_d));
    }
    ctx->dd_pending = true;
    if (ngroups) TRY(dd_finish(ctx, nrounds, groups_d != nullptr, ngroups));
    return FMCS_OK;
  }
  // order = identity (src/composer.cpp:60-63)
  hipLaunchKernelGGL(k_iota_u32, dim3((u32)std::min<u64>((n + 255) / 256, 8192)), dim3(256), 0, ctx->stream, order_d, n);
  TRY(launch_check(ctx, "k_iota_u32"));
  u32 *ga, *gb, *offa, *offb;
  TRY(getbuf_t(ctx, WS_HASH, n, &ga));
  TRY(getbuf_t(ctx, WS_HASH2, n, &gb));
  TRY(getbuf_t(ctx, WS_SEGOFF, n + 1, &offa));
  TRY(getbuf_t(ctx, WS_SEGOFF2, n + 1, &offb));
  const u32* gin = nullptr;
  const u32* soff = nullptr;
  u32 nseg = 0, ng = 0;
  double tg = 0, ts1 = 0, ts = 0, tx = 0;
  for (u32 r = 0; r < nrounds; ++r) {
    bool last = (r + 1 == nrounds);
    bool want_groups = !last || groups_d;
    u32* gout = last && groups_d ? groups_d : ((gin == ga) ? gb : ga);
    u32* oout = (soff == offa) ? offb : offa;
    double* tsort = (r == 0) ? &ts1 : &ts;
    int rc;
    switch (banks[r]) {
      case 2: rc = sort_by_group_t<u16>(ctx, bitwidths[r], n, (const u16*)cols_d[r], order_d, gin, gout, !want_groups, mode, nullptr, nullptr, &ng, soff, nseg, oout, tsort, &tg, &tx); break;
      case 4: rc = sort_by_group_t<u32>(ctx, bitwidths[r], n, (const u32*)cols_d[r], order_d, gin, gout, !want_groups, mode, nullptr, nullptr, &ng, soff, nseg, oout, tsort, &tg, &tx); break;
      default: rc = sort_by_group_t<u64>(ctx, bitwidths[r], n, (const u64*)cols_d[r], order_d, gin, gout, !want_groups, mode, nullptr, nullptr, &ng, soff, nseg, oout, tsort, &tg, &tx); break;
    }
    if (rc != FMCS_OK) return rc;
    if (!last) {
      gin = gout;
      soff = oout;
      nseg = ng;
    }
  }
  ctx->breakdown[0] = tg;
  ctx->breakdown[1] = ts1;
  ctx->breakdown[2] = ts;
  ctx->breakdown[3] = tx;
  if (ngroups) *ngroups = groups_d ? ng : nseg;
  return FMCS_OK;
}

}  // namespace

extern "C" {

int fmcs_ctx_set_profiling(fmcs_ctx* ctx, int on) {
  if (!ctx) return FMCS_E_ARG;
  TRY(prof_flush(ctx));
  ctx->prof = on != 0;
  return FMCS_OK;
}

int fmcs_kernel_stats(fmcs_ctx* ctx, char* buf, uint64_t cap) {
  if (!ctx || !buf || cap == 0) return FMCS_E_ARG;
  TRY(prof_flush(ctx));
  std::string js = "{";
  bool first = true;
  for (auto& kv : ctx->agg) {
    char line[256];
    std::snprintf(line, sizeof(line),
                  "%s\"%s\": {\"launches\": %llu, \"ms\": %.6f, \"bytes\": %.0f, \"extra_bytes\": %.0f}",
                  first ? "" : ", ", kv.first.c_str(), (unsigned long long)kv.second.launches, kv.second.ms,
                  kv.second.bytes, kv.second.extra);
    js += line;
    first = false;
  }
  js += "}";
  if (js.size() + 1 > cap) return fail(ctx, FMCS_E_ARG, "stats buffer too small");
  std::memcpy(buf, js.c_str(), js.size() + 1);
  return FMCS_OK;
}

int fmcs_kernel_stats_reset(fmcs_ctx* ctx) {
  if (!ctx) return FMCS_E_ARG;
  TRY(prof_flush(ctx));
  ctx->agg.clear();
  return FMCS_OK;
}

int fmcs_last_breakdown(fmcs_ctx* ctx, double* ms4) {
  if (!ctx || !ms4) return FMCS_E_ARG;
  for (int i = 0; i < 4; ++i) ms4[i] = ctx->breakdown[i];
  return FMCS_OK;
}

int fmcs_hash_round(fmcs_ctx* ctx, uint32_t bank, uint64_t n, const void* values_d, const uint32_t* order_d,
                    const uint32_t* in_group_d, uint32_t* gid_d, uint32_t* out_order_d, uint32_t* out_group_d,
                    int last_round, uint32_t* ngroups) {
  if (!ctx || !gid_d) return FMCS_E_ARG;
  if (!(bank == 1 || bank_ok(bank))) return fail(ctx, FMCS_E_ARG, "bank must be 1, 2, 4 or 8");
  if (!last_round && (!out_order_d || !out_group_d || !order_d))
    return fail(ctx, FMCS_E_ARG, "order, out_order and out_group required unless last round");
  Stager st(ctx);
  TRY(st.arg(values_d, n * bank, Stager::IN, "values_d"));
  TRY(st.arg(order_d, n * 4, Stager::IN, "order_d"));
  TRY(st.arg(in_group_d, n * 4, Stager::IN, "in_group_d"));
  TRY(st.arg(gid_d, n * 4, Stager::OUT, "gid_d"));
  TRY(st.arg(out_order_d, n * 4, Stager::OUT, "out_order_d"));
  TRY(st.arg(out_group_d, n * 4, Stager::OUT, "out_group_d"));
  const void* keys = values_d;
  if (in_group_d && n) {
    if (!order_d) return fail(ctx, FMCS_E_ARG, "order required after round 1");
    if (((uintptr_t)order_d & 15) != 0) return fail(ctx, FMCS_E_ARG, "order must be 16-byte aligned");
    void* ka;
    TRY(getbuf(ctx, WS_KA, n * bank, &ka));
    u32 blocks = (u32)((n + 1023) / 1024);
    Prof pf(ctx, "k_gather", (double)n * (4 + 2 * bank));
    switch (bank) {
      case 1: hipLaunchKernelGGL((k_gather<uint8_t>), dim3(blocks), dim3(256), 0, ctx->stream, (const uint8_t*)values_d, order_d, (uint8_t*)ka, n); break;
      case 2: hipLaunchKernelGGL((k_gather<u16>), dim3(blocks), dim3(256), 0, ctx->stream, (const u16*)values_d, order_d, (u16*)ka, n); break;
      case 4: hipLaunchKernelGGL((k_gather<u32>), dim3(blocks), dim3(256), 0, ctx->stream, (const u32*)values_d, order_d, (u32*)ka, n); break;
      default: hipLaunchKernelGGL((k_gather<u64>), dim3(blocks), dim3(256), 0, ctx->stream, (const u64*)values_d, order_d, (u64*)ka, n); break;
    }
    TRY(launch_check(ctx, "k_gather"));
    keys = ka;
  }
  u32 ng = 0;
  TRY(hash_ids(ctx, bank, keys, in_group_d, n, gid_d, &ng));
  if (!last_round && n) {
    u32 bits = 1;
    while (bits < 32 && (1ull << bits) < ng) ++bits;
    TRY(radix_sort<u32>(ctx, gid_d, order_d, out_group_d, out_order_d, n, bits));
  }
  if (ngroups) *ngroups = ng;
  return st.finish();
}

int fmcs_probe_handoff(fmcs_ctx* ctx, uint32_t tasks, uint32_t grid, double* ms4) {
  if (!ctx || !ms4 || tasks == 0 || tasks > (1u << 17) || grid == 0 || grid > (1u << 16))
    return fail(ctx, FMCS_E_ARG, "probe: 1..2^17 tasks, 1..2^16 blocks");
  const u64 rows = (u64)tasks * PROBE_ROWS;
  void *in, *out, *fl;
  TRY(internal::scratch(ctx, 13, rows * 8, &in));
  TRY(internal::scratch(ctx, 14, rows * 8, &out));
  TRY(internal::scratch(ctx, 15, 4096, &fl));
  TRY(fill(ctx, (u32*)in, rows * 2, 0));
  TRY(fill(ctx, (u32*)fl, 1024, 0));
  void (*fns[4])(const u64*, u64*, u32, u32*) = {k_handoff_probe<0>, k_handoff_probe<1>, k_handoff_probe<2>,
                                                  k_handoff_probe<3>};
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  int rc = FMCS_OK;
  for (int m = 0; m < 4 && rc == FMCS_OK; ++m) {
    float best = 1e30f;
    for (int r = 0; r < 5 && rc == FMCS_OK; ++r) {
      hipError_t e = hipEventRecord(e0, ctx->stream);
      if (e == hipSuccess) {
        hipLaunchKernelGGL(fns[m], dim3(grid), dim3(256), 0, ctx->stream, (const u64*)in, (u64*)out, tasks, (u32*)fl);
        e = hipGetLastError();
      }
      if (e == hipSuccess) e = hipEventRecord(e1, ctx->stream);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0.0f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      if (e != hipSuccess) rc = fail(ctx, FMCS_E_HIP, hipGetErrorString(e));
      best = std::min(best, ms);
    }
    ms4[m] = best;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

int fmcs_distinct_count(fmcs_ctx* ctx, uint32_t bank, uint64_t n, const void* col_d, uint64_t* count) {
  if (!ctx || !count || (n && !col_d) || !(bank == 1 || bank == 2 || bank == 4 || bank == 8))
    return FMCS_E_ARG;
  if (n >= (1ull << 30)) return fail(ctx, FMCS_E_UNSUPPORTED, "at most 2^30 - 1 rows");
  Stager st(ctx);
  TRY(st.arg(col_d, n * bank, Stager::IN, "col_d"));
  u32* gid;
  TRY(getbuf_t(ctx, WS_HASH3, std::max<u64>(n, 1), &gid));
  u32 ng = 0;
  TRY(hash_ids(ctx, bank, col_d, nullptr, n, gid, &ng));
  *count = ng;
  return FMCS_OK;
}

int fmcs_multi_round_hash(fmcs_ctx* ctx, uint32_t nrounds, const void* const* cols_d, const uint32_t* banks,
                          uint64_t n, int hash_type, uint32_t* order_d, uint32_t* gid_d, uint32_t* ngroups) {
  if (!ctx || !cols_d || !banks || !order_d || !gid_d || nrounds == 0) return FMCS_E_ARG;
  if (n >= (1ull << 30)) return fail(ctx, FMCS_E_UNSUPPORTED, "n < 2^30 required");
  for (u32 r = 0; r < nrounds; ++r)
    if (!(banks[r] == 1 || bank_ok(banks[r]))) return fail(ctx, FMCS_E_ARG, "bad bank");
  if (n == 0) {
    if (ngroups) *ngroups = 0;
    return FMCS_OK;
  }
  Stager st(ctx);
  std::vector<const void*> colv(cols_d, cols_d + nrounds);
  for (u32 r = 0; r < nrounds; ++r) TRY(st.arg(colv[r], n * banks[r], Stager::IN, "cols_d"));
  cols_d = colv.data();
  TRY(st.arg(order_d, n * 4, Stager::OUT, "order_d"));
  TRY(st.arg(gid_d, n * 4, Stager::OUT, "gid_d"));
  hipLaunchKernelGGL(k_iota_u32, dim3((u32)std::min<u64>((n + 255) / 256, 8192)), dim3(256), 0, ctx->stream, order_d, n);
  TRY(launch_check(ctx, "k_iota_u32"));
  HIPCHK(hipEventRecord(ctx->ev[0], ctx->stream));
  u32 ng = 0;
  if (hash_type == 0) {
    // hashByGroup_non_reordered (src/composer.cpp:1153-1219): every round
    // hashes (previous id, value) over the rows in their original order.
    u32* prev;
    TRY(getbuf_t(ctx, WS_GROUPTMP, n, &prev));
    for (u32 r = 0; r < nrounds; ++r) {
      if (r) HIPCHK(hipMemcpyAsync(prev, gid_d, n * 4, hipMemcpyDeviceToDevice, ctx->stream));
      TRY(hash_ids(ctx, banks[r], cols_d[r], r ? prev : nullptr, n, gid_d, &ng));
    }
  } else {
    // hashByGroup_reordered (src/composer.cpp:814-976)
    u32 *ob, *ga, *gb;
    TRY(getbuf_t(ctx, WS_HASH, n, &ga));
    TRY(getbuf_t(ctx, WS_HASH2, n, &gb));
    TRY(getbuf_t(ctx, WS_HASH3, n, &ob));
    u32* cur_order = order_d;
    u32* other_order = ob;
    const u32* gin = nullptr;
    for (u32 r = 0; r < nrounds; ++r) {
      bool last = (r + 1 == nrounds);
      u32* gout = (gin == ga) ? gb : ga;
      TRY(fmcs_hash_round(ctx, banks[r], n, cols_d[r], cur_order, gin, gid_d, last ? nullptr : other_order,
                          last ? nullptr : gout, last ? 1 : 0, &ng));
      if (!last) {
        std::swap(cur_order, other_order);
        gin = gout;
      }
    }
    if (cur_order != order_d) HIPCHK(hipMemcpyAsync(order_d, cur_order, n * 4, hipMemcpyDeviceToDevice, ctx->stream));
  }
  HIPCHK(hipEventRecord(ctx->ev[1], ctx->stream));
  HIPCHK(hipEventSynchronize(ctx->ev[1]));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
  ctx->breakdown[0] = 0;
  ctx->breakdown[1] = ms;
  ctx->breakdown[2] = 0;
  ctx->breakdown[3] = 0;
  if (ngroups) *ngroups = ng;
  return st.finish();
}

}  // extern "C"
