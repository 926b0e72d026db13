// fmcs_bins.hip -- the persistent size-binned launch of the batched
// segmented sort.
//
// The device-driven radix levels split every segment above med_max rows
// (device-wide MSD passes, one level at a time).  Everything else that the
// levels leave behind is sorted by one persistent launch with two size bins:
//
// * single segments of small_max + 1 .. med_max rows.  Up to one window tile
//   (BIN_T rows) they are a window of their own; larger ones ("medium") are
//   sorted end to end by the block that takes them: an MSD level of its own
//   (a digit histogram of the whole segment, then a stable LDS-staged scatter
//   tile by tile into the other buffer), then the child segments packed into
//   windows and sorted like any other window, children still above BIN_T rows
//   pushed on a per-block stack for another level.  No row of a medium
//   segment is touched by another block, so nothing crosses between blocks
//   (or XCDs) inside the launch;
// * windows of packed small segments.
//
// Every window goes through the bucket sort.  A window it refuses (a bin of
// more than BUCKET_LIMIT rows: heavy duplicates) is listed for the radix tile
// sort that runs after the launch (k_tile_sort, then k_tile_rank for windows
// too wide for a segment index), with its segment starts copied into a
// pool indexed by row (pool[lo + p] = lo + p where a segment starts, else
// hi), so the list stays valid whatever the block does next.  Keeping those
// rarely used sorts out of this kernel keeps its registers at the level of
// the bucket sort and the medium scatter.
//
// Work is dealt XCD-aware: die x takes the x-th contiguous eighth of each
// bin (singles first: they hold the largest items), its blocks walking it
// together.
#include "fmcs_kernels.h"
#include "fmcs_tile.h"

namespace fmcs {

namespace {

constexpr int MED_ITEMS = 4;                        // rows a thread in a medium scatter tile
constexpr u32 MED_TILE = 256 * MED_ITEMS;           // rows of a medium scatter tile
constexpr u32 SC_CS = 0;                            // child starts (RADIX + 1 words)
constexpr u32 SC_STACK = 264;                       // BIN_STACK segments (lo, hi, key bits, buffer)
constexpr u32 SC_WIN = SC_STACK + 4 * BIN_STACK;    // up to RADIX windows (lo, hi, s0, s1)
static_assert(SC_WIN + 4 * RADIX <= BIN_SCRATCH, "block scratch too small");
static_assert(BIN_MED_MAX / (BIN_T + 1) * (64 / RADIX_BITS) + 1 <= BIN_STACK, "medium stack bound");

template <class K>
struct MedLds {
  u32 whist[4][RADIX];  // per-wave digit counts, then per-wave LDS offsets
  u32 goff[RADIX];      // next global row of every child
  u32 gt[RADIX];        // a digit's global row minus its first staging slot
  u32 cnt[RADIX];       // child sizes
  u32 cst[RADIX];       // child starts
  u32 wid[RADIX];       // child windows (~0: not a small child)
  u32 tmp[8];
  __align__(16) K keys[MED_TILE];
  u32 vals[MED_TILE];
};

template <class K>
union BinLds {
  BucketLds<K, BIN_E> b;
  MedLds<K> m;
  uint8_t f[BIN_T];  // a refused window's segment-start flags
};

// v[L] for a level L known only at run time, without indexing the kernel
// argument block (which would copy it to private memory).
// A value every lane holds alike (read from LDS or memory at one address),
// kept in a scalar register.
__device__ __forceinline__ u32 uni(u32 v) { return (u32)__builtin_amdgcn_readfirstlane((int)v); }

__device__ __forceinline__ u32 pick(const u32 (&v)[BIN_LVL_MAX], u32 L) {
  u32 r = v[0];
#pragma unroll
  for (u32 q = 1; q < BIN_LVL_MAX; ++q)
    if (L == q) r = v[q];
  return r;
}

template <class K>
__device__ __forceinline__ const K* kin_of(const BinArgs<K>& a, u32 s) {
  return s ? a.kin[1] : a.kin[0];
}
template <class K>
__device__ __forceinline__ const u32* vin_of(const BinArgs<K>& a, u32 s) {
  return s ? a.vin[1] : a.vin[0];
}

// Inclusive max-scan over the block's 256 threads.  s_tmp: 4 words.
// Contains __syncthreads.
__device__ __forceinline__ u32 block_max_incl(u32 v, u32* s_tmp) {
  const u32 lane = __lane_id(), wv = threadIdx.x / WAVE;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    const u32 t = __shfl_up(v, (unsigned)o);
    if (lane >= (u32)o) v = max(v, t);
  }
  if (lane == WAVE - 1) s_tmp[wv] = v;
  __syncthreads();
  u32 pre = 0;
  for (u32 w = 0; w < wv; ++w) pre = max(pre, s_tmp[w]);
  __syncthreads();
  return max(v, pre);
}

// One window of rows of buffer src, sorted into the final rows by the bucket
// sort, or (refused) listed for the radix tile sort after the launch.
template <class K>
__device__ __forceinline__ void bin_window(const BinArgs<K>& a, u32 src, const u32* starts, u32 kb, TileWin win,
                                           bool check, BinLds<K>& U) {
  WinRows<K, BIN_E> r;
  r.w = 0;
  load_rows<K, BIN_E>(kin_of(a, src), vin_of(a, src), starts, win.lo, win.hi, win.s0, win.s1, r);
  if (!tile_bucket_body<K, BIN_E>(r, starts, kb, a.kout, a.vout, a.bad, check ? a.err : nullptr, U.b)) return;
  __syncthreads();
  const u32 tid = threadIdx.x, cnt = win.hi - win.lo;
  for (u32 i = tid; i < BIN_T; i += 256) U.f[i] = 0;
  __syncthreads();
  for (u32 s = win.s0 + tid; s <= win.s1; s += 256) {
    const u32 st = starts[s];
    if (st >= win.lo && st < win.hi) U.f[st - win.lo] = 1;
  }
  if (tid == 0) U.f[0] = 1;
  __syncthreads();
  for (u32 p = tid; p < cnt; p += 256) a.fbpool[win.lo + p] = U.f[p] ? win.lo + p : win.hi;
  if (tid == 0) {
    const u32 k = atomicAdd(a.fbcount + src, 1u);
    const u32 e = src * a.fbcap + k;
    a.fblo[e] = win.lo;
    a.fbhi[e] = win.hi;
    a.fbs0[e] = win.lo;
    a.fbs1[e] = win.hi - 1;
  }
}

// One MSD level of segment [lo, hi) (more than BIN_T rows, all in buffer
// src, kb key bits still to sort) by this block alone.  Its children are
// listed in the block's scratch: windows (sorted next by the caller, from
// buffer src ^ 1 with kb - D key bits) and, on the stack, the children still
// above BIN_T rows.  Returns the children's key bits.
template <class K>
__device__ u32 med_node(const BinArgs<K>& a, u32 lo, u32 hi, u32 kb, u32 src, bool check, BinLds<K>& U, u32* sc,
                        u32& s_nstack, u32& s_nwin) {
  constexpr int NT = 256, NW = NT / WAVE;
  MedLds<K>& L = U.m;
  const u32 tid = threadIdx.x, wv = tid / WAVE, lane = tid & (WAVE - 1);
  const K* kin = kin_of(a, src);
  const u32* vin = vin_of(a, src);
  if (kb == 0) {
    // every key bit consumed: the rows are in order (each level's scatter
    // is stable); they only move to the final rows
    if (src != 0)
      for (u32 i = lo + tid; i < hi; i += NT) {
        if (a.kout) a.kout[i] = kin[i];
        a.vout[i] = vin[i];
      }
    return 0;
  }
  const u32 D = min((u32)RADIX_BITS, kb), shift = kb - D, mask = (1u << D) - 1;
  // 1. the digit counts of the whole segment -> child starts
  for (u32 i = tid; i < (u32)(NW * RADIX); i += NT) (&L.whist[0][0])[i] = 0;
  __syncthreads();
  u64 bad = 0;
  for (u32 base = lo; base < hi; base += MED_TILE) {
    K k[MED_ITEMS];
#pragma unroll
    for (int j = 0; j < MED_ITEMS; ++j) k[j] = kin[min(base + j * NT + tid, hi - 1)];
#pragma unroll
    for (int j = 0; j < MED_ITEMS; ++j)
      if (base + j * NT + tid < hi) {
        bad |= (u64)k[j] & a.bad;
        atomicAdd(&L.whist[wv][digit_of(k[j], shift, mask)], 1u);
      }
  }
  if (check && a.err && __ballot(bad != 0) && lane == 0) atomicOr(a.err, 1u);
  __syncthreads();
  {
    u32 c = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) c += L.whist[w][tid];
    u32 tot;
    const u32 ex = block_excl_scan<NT>(c, L.tmp, &tot);
    L.cnt[tid] = c;
    L.cst[tid] = lo + ex;
    L.goff[tid] = lo + ex;
    sc[SC_CS + tid] = lo + ex;
    if (tid == 0) sc[SC_CS + RADIX] = hi;
  }
  // 2. stable scatter into the other buffer, tile by tile: item j of lane l
  // of wave w is row w * 256 + j * 64 + l of the tile, ranked by ballot
  // matching into per-wave LDS counters and staged in LDS, so every digit's
  // rows leave as one contiguous run
  K* kdst = src ? a.kbuf[0] : a.kbuf[1];
  u32* vdst = src ? a.vbuf[0] : a.vbuf[1];
  const u64 lt = lanemask_lt();
  for (u32 t0 = lo; t0 < hi; t0 += MED_TILE) {
    const u32 tend = min(hi, t0 + MED_TILE);
    const u32 wbase = t0 + wv * (WAVE * MED_ITEMS);
    K keys[MED_ITEMS];
    u32 vals[MED_ITEMS], rank[MED_ITEMS], old[MED_ITEMS];
#pragma unroll
    for (int j = 0; j < MED_ITEMS; ++j) keys[j] = kin[min(wbase + j * WAVE + lane, tend - 1)];
#pragma unroll
    for (int j = 0; j < MED_ITEMS; ++j) vals[j] = vin[min(wbase + j * WAVE + lane, tend - 1)];
    __syncthreads();
    for (u32 i = tid; i < (u32)(NW * RADIX); i += NT) (&L.whist[0][0])[i] = 0;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < MED_ITEMS; ++j) {
      const bool ok = wbase + j * WAVE + lane < tend;
      const u32 dg = digit_of(keys[j], shift, mask);
      // lanes whose digit differs from this lane's in any bit
      u32 dlo = 0, dhi = 0;
#pragma unroll
      for (int b = 0; b < RADIX_BITS; ++b) {
        const u32 own = (u32)__builtin_amdgcn_sbfe((int)dg, b, 1);
        const u64 bb = __ballot(own != 0);
        dlo = __builtin_amdgcn_bitop3_b32(dlo, (u32)bb, own, BITOP3_OR_XOR);
        dhi = __builtin_amdgcn_bitop3_b32(dhi, (u32)(bb >> 32), own, BITOP3_OR_XOR);
      }
      const u64 peers = ok ? __ballot(ok) & ~(((u64)dhi << 32) | dlo) : 0ull;
      const u32 lead = ok ? (u32)(__ffsll((long long)peers) - 1) : lane;
      rank[j] = (lead << 8) | (u32)__popcll(peers & lt);
      old[j] = 0;
      if (ok && lane == lead) old[j] = atomicAdd(&L.whist[wv][dg], (u32)__popcll(peers));
    }
#pragma unroll
    for (int j = 0; j < MED_ITEMS; ++j) rank[j] = __shfl(old[j], (int)(rank[j] >> 8)) + (rank[j] & 0xffu);
    __syncthreads();
    {
      // digit tid: the waves' counts -> per-wave staging offsets; the tile's
      // count advances the child's next global row
      u32 wc[NW], run = 0;
#pragma unroll
      for (int w = 0; w < NW; ++w) {
        wc[w] = L.whist[w][tid];
        run += wc[w];
      }
      u32 ttot;
      const u32 tdig = block_excl_scan<NT>(run, L.tmp, &ttot);
      u32 o = tdig;
#pragma unroll
      for (int w = 0; w < NW; ++w) {
        L.whist[w][tid] = o;
        o += wc[w];
      }
      const u32 g = L.goff[tid];
      L.gt[tid] = g - tdig;
      L.goff[tid] = g + run;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < MED_ITEMS; ++j)
      if (wbase + j * WAVE + lane < tend) {
        const u32 pos = L.whist[wv][digit_of(keys[j], shift, mask)] + rank[j];
        L.keys[pos] = keys[j];
        L.vals[pos] = vals[j];
      }
    __syncthreads();
    const u32 cnt = tend - t0;
    for (u32 i = tid; i < cnt; i += NT) {
      const K k = L.keys[i];
      const u32 g = L.gt[digit_of(k, shift, mask)] + i;
      kdst[g] = k;
      vdst[g] = L.vals[i];
    }
  }
  // 3. the children: runs of small ones packed into windows by position
  // (an empty child counts at start - 1 within the segment), single ones a
  // window each, larger ones pushed for another level
  const u32 nkb = kb - D, nsrc = src ^ 1u;
  {
    const u32 d = tid;
    const u32 z = L.cnt[d], st = L.cst[d];
    const bool small = z <= a.small_max;
    const u32 pos = (z == 0 && st > lo) ? st - 1 : st;
    L.wid[d] = small ? (pos - lo) / a.W : ~0u;
    __syncthreads();
    const u32 wd = L.wid[d];
    const bool first = small && (d == 0 || L.wid[d - 1] != wd);
    const bool last = small && (d == RADIX - 1 || L.wid[d + 1] != wd);
    const u32 fidx = block_max_incl(first ? d : 0u, L.tmp);
    if (last) {
      const u32 wlo = L.cst[fidx], whi = st + z;
      if (whi > wlo) {
        const u32 k = atomicAdd(&s_nwin, 1u);
        sc[SC_WIN + 4 * k] = wlo;
        sc[SC_WIN + 4 * k + 1] = whi;
        sc[SC_WIN + 4 * k + 2] = fidx;
        sc[SC_WIN + 4 * k + 3] = d;
      }
    } else if (!small && z <= BIN_T) {
      const u32 k = atomicAdd(&s_nwin, 1u);
      sc[SC_WIN + 4 * k] = st;
      sc[SC_WIN + 4 * k + 1] = st + z;
      sc[SC_WIN + 4 * k + 2] = d;
      sc[SC_WIN + 4 * k + 3] = d;
    } else if (z > BIN_T) {
      const u32 k = atomicAdd(&s_nstack, 1u);
      if (k < BIN_STACK) {
        sc[SC_STACK + 4 * k] = st;
        sc[SC_STACK + 4 * k + 1] = st + z;
        sc[SC_STACK + 4 * k + 2] = nkb;
        sc[SC_STACK + 4 * k + 3] = nsrc;
      } else if (a.err) {
        atomicOr(a.err, 2u);
      }
    }
  }
  __syncthreads();
  if (tid == 0 && s_nstack > BIN_STACK) s_nstack = BIN_STACK;
  return nkb;
}

}  // namespace

// Block loop over the two bins.  An item is a window (sorted at once) or a
// medium segment, whose levels run from the block's stack, each leaving its
// children's windows in the block's window list; the windows of a list share
// a source buffer, key bits and segment starts.  One call site for each body
// keeps the kernel's registers at those of its largest body.
// The register budget: waves a SIMD the kernel is compiled for (the bucket
// sort's windows are bound by the latency of their barrier phases, so the
// windows in flight a CU are what pays).
#ifndef FMCS_BINS_WAVES
#define FMCS_BINS_WAVES 6
#endif
template <class K>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FMCS_BINS_WAVES, 8))) void k_seg_bins(
    BinArgs<K> a) {
  __shared__ BinLds<K> U;
  __shared__ u32 s_cnt[2][BIN_LVL_MAX];
  __shared__ u32 s_nstack, s_nwin;
  const u32 tid = threadIdx.x;
  u32* sc = a.scratch + (u64)blockIdx.x * BIN_SCRATCH;
  if (tid < BIN_LVL_MAX) {
    const bool in = tid < a.nlev;
    s_cnt[0][tid] = in ? a.lv[(u64)tid * DD_LVL_WORDS + LV_NSINGLE] : 0u;
    s_cnt[1][tid] = in ? a.lv[(u64)tid * DD_LVL_WORDS + LV_WCOUNT] : 0u;
  }
  if (tid == 0) s_nstack = s_nwin = 0;
  __syncthreads();
  const u32 xcd = blockIdx.x % XCDS, per_x = gridDim.x / XCDS, slot = blockIdx.x / XCDS;
  for (u32 bin = 0; bin < 2; ++bin) {
    u32 total = 0;
    for (u32 q = 0; q < BIN_LVL_MAX; ++q) total += uni(s_cnt[bin][q]);
    const u32 end = (u32)((u64)total * (xcd + 1) / XCDS);
    for (u32 i = (u32)((u64)total * xcd / XCDS) + slot; i < end; i += per_x) {
      u32 L = 0, j = i;
      while (L + 1 < BIN_LVL_MAX && j >= uni(s_cnt[bin][L])) {
        j -= uni(s_cnt[bin][L]);
        ++L;
      }
      // the item: one window, or a medium segment pushed on the stack
      u32 src = L & 1u, kb = pick(a.kbs, L);
      const u32* starts = a.starts + pick(a.loff, L);
      bool check = L == 0;
      TileWin w0;
      bool one = true;
      if (bin == 0) {
        const u32 o = pick(a.soff, L) + j;
        w0 = TileWin{uni(a.slo[o]), uni(a.shi[o]), uni(a.sseg[o]), uni(a.sseg[o])};
        if (w0.hi - w0.lo > BIN_T) {
          one = false;
          if (tid == 0) {
            sc[SC_STACK] = w0.lo;
            sc[SC_STACK + 1] = w0.hi;
            sc[SC_STACK + 2] = kb;
            sc[SC_STACK + 3] = src;
            s_nstack = 1;
          }
        }
      } else {
        const u32 w = pick(a.woff, L) + uni(a.wlist[pick(a.woff, L) + j]);
        w0 = TileWin{uni(a.tlo[w]), uni(a.thi[w]), uni(a.ts0[w]), uni(a.ts1[w])};
      }
      __syncthreads();
      for (;;) {
        if (!one) {
          // next level of the medium segment on top of the stack
          const u32 ns = uni(s_nstack);
          if (ns == 0) break;
          const u32 e = SC_STACK + 4 * (ns - 1);
          const u32 nlo = uni(sc[e]), nhi = uni(sc[e + 1]), nkb = uni(sc[e + 2]), nsrc = uni(sc[e + 3]);
          __syncthreads();
          if (tid == 0) {
            s_nstack = ns - 1;
            s_nwin = 0;
          }
          __syncthreads();
          kb = med_node(a, nlo, nhi, nkb, nsrc, check, U, sc, s_nstack, s_nwin);
          src = nsrc ^ 1u;
          starts = sc + SC_CS;
          check = false;
        }
        const u32 nw = one ? 1u : uni(s_nwin);
        for (u32 k = 0; k < nw; ++k) {
          const TileWin win = one ? w0
                                  : TileWin{uni(sc[SC_WIN + 4 * k]), uni(sc[SC_WIN + 4 * k + 1]),
                                            uni(sc[SC_WIN + 4 * k + 2]), uni(sc[SC_WIN + 4 * k + 3])};
          __syncthreads();
          bin_window(a, src, starts, kb, win, check, U);
        }
        __syncthreads();
        if (one) break;
      }
    }
  }
}

// Probe of what a hand-off between workgroups costs inside one launch: every
// block copies tasks of PROBE_ROWS u64 rows; after each task mode 1 has
// every wave release its stores at agent scope (a fence, as when each
// wave publishes its own output), mode 2 one release after a barrier (the
// least a task whose output another workgroup reads must do), mode 3 the
// same after every 8th task of a block; mode 0 none.  On MI355X an
// agent-scope release writes back the XCD's L2 (buffer_wbl2), which
// serialises per XCD.
template <int MODE>
__global__ __launch_bounds__(256) void k_handoff_probe(const u64* __restrict__ in, u64* __restrict__ out, u32 ntasks,
                                                       u32* __restrict__ flags) {
  for (u32 t = blockIdx.x; t < ntasks; t += gridDim.x) {
    const u64* src = in + (u64)t * PROBE_ROWS;
    u64* dst = out + (u64)t * PROBE_ROWS;
#pragma unroll
    for (u32 j = 0; j < PROBE_ROWS / 256; ++j) dst[j * 256 + threadIdx.x] = src[j * 256 + threadIdx.x] + 1;
    if (MODE == 1) {
      __threadfence();
      if (threadIdx.x == 0)
        __hip_atomic_fetch_add(flags + (blockIdx.x % XCDS) * 16, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else if (MODE == 2 || (MODE == 3 && (t / gridDim.x) % 8 == 7)) {
      __syncthreads();
      if (threadIdx.x == 0)
        __hip_atomic_fetch_add(flags + (blockIdx.x % XCDS) * 16, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}
template __global__ void k_handoff_probe<0>(const u64*, u64*, u32, u32*);
template __global__ void k_handoff_probe<1>(const u64*, u64*, u32, u32*);
template __global__ void k_handoff_probe<2>(const u64*, u64*, u32, u32*);
template __global__ void k_handoff_probe<3>(const u64*, u64*, u32, u32*);

template __global__ void k_seg_bins<uint16_t>(BinArgs<uint16_t>);
template __global__ void k_seg_bins<uint32_t>(BinArgs<uint32_t>);
template __global__ void k_seg_bins<uint64_t>(BinArgs<uint64_t>);

}  // namespace fmcs
