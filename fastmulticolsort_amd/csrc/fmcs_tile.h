// fmcs_tile.h -- the LDS window sorts of the batched segmented sort.
//
// A window is a range of rows that holds whole segments, at most 256 E rows
// in all.  Three sorts order a window's rows by (segment, key, oid): the
// bucket sort (the fast path), the radix tile sort (windows whose bins
// overflow: heavy duplicates) and the comparison rank (windows whose varying
// key bits leave no room for a segment index).  Each body takes its LDS as a
// struct, so a kernel can hold it in a __shared__ variable of its own or
// overlay several bodies in one union.
#pragma once
#include "fmcs_kernels.h"

namespace fmcs {

// Rows [lo, hi) holding whole segments whose starts are seg_start[s0 .. s1].
struct TileWin {
  u32 lo, hi, s0, s1;
};

// ---------------------------------------------------------------------------
// LDS tile sort for packed small segments.  A tile is a window of rows that
// holds whole segments; the segment index is recomputed from the segment
// starts (no per-row segment ids are read).  The tile is sorted by the
// composite (local segment << kb) | (key & mask(kb)) with an LSD radix sort
// in LDS: each pass ranks keys stably with 64-lane ballot matching into
// per-wave histograms and scatters them through LDS.  Stability keeps equal
// keys in their incoming (ascending-oid) order, so no oid compare is needed.
// Passes cover only the bits that vary inside the tile.  The top (already
// partitioned) key bits of every segment are kept in LDS and restored on the
// way out.
// ---------------------------------------------------------------------------

// Stable ranking of E items per lane (item j of lane l is the j*64+l-th
// item of the wave) into the wave's LDS digit counters.
template <int E>
__device__ __forceinline__ void wave_rank(const u32 (&dig)[E], const bool (&ok)[E], u32 dbits, u32* whist,
                                          u32 (&rank)[E], u32 jn = E) {
  const u32 lane = __lane_id();
  const u64 lt = lanemask_lt();
#pragma unroll
  for (int j = 0; j < E; ++j) {
    if ((u32)j >= jn) break;  // wave-uniform: no item row beyond jn holds a row
    u32 d = dig[j];
    u64 peers = __ballot(ok[j]);
    for (u32 b = 0; b < dbits; ++b) {
      u64 bb = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? bb : ~bb;
    }
    if (!ok[j]) peers = 0;
    u32 leader = ok[j] ? (u32)(__ffsll((long long)peers) - 1) : lane;
    u32 old = 0;
    if (ok[j] && lane == leader) {
      old = whist[d];
      whist[d] = old + (u32)__popcll(peers);
    }
    old = __shfl(old, (int)leader);
    rank[j] = old + (u32)__popcll(peers & lt);
  }
}

template <int NT>
__device__ __forceinline__ void block_or_and(u64 v_or, u64 v_and, u64* s_red, u64* out_or, u64* out_and) {
  constexpr int NW = NT / WAVE;
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) {
    v_or |= __shfl_xor(v_or, o);
    v_and &= __shfl_xor(v_and, o);
  }
  u32 wv = threadIdx.x / WAVE;
  if ((threadIdx.x & (WAVE - 1)) == 0) {
    s_red[wv] = v_or;
    s_red[NW + wv] = v_and;
  }
  __syncthreads();
  u64 ro = 0, ra = ~0ull;
  for (int w = 0; w < NW; ++w) {
    ro |= s_red[w];
    ra &= s_red[NW + w];
  }
  *out_or = ro;
  *out_and = ra;
  __syncthreads();
}

// LDS of the radix tile sort (T = 256 E rows).
template <class K, int E>
struct TileSortLds {
  static constexpr int NT = 256, NW = NT / WAVE, T = NT * E;
  u64 c[T];
  u32 o[T];
  u64 top[T];
  uint8_t f[T];
  u32 whist[NW][RADIX];
  u32 dig[RADIX];
  u32 tmp[NW + 1];
  u64 red[2 * NW];
};

// Sorts window `win` (rows [lo, hi) of kin/vin holding whole segments whose
// starts are seg_start[s0 .. s1]) into kout/vout at the same rows.  Returns
// true, without writing anything, when the key bits that vary inside the
// window leave no room for the segment index: the comparison rank
// (tile_rank_win) sorts such a window (w is then appended to rklist, if set).
template <class K, int E>
__device__ __forceinline__ bool tile_sort_win(const K* __restrict__ kin, const u32* __restrict__ vin, K* kout,
                                              u32* vout, const u32* __restrict__ seg_start, u32 kb, TileWin win,
                                              TileSortLds<K, E>& L, u32* rklist = nullptr, u32* rkcount = nullptr,
                                              u32 w = 0) {
  constexpr int NT = 256, NW = NT / WAVE, T = NT * E;
  u64* s_c = L.c;
  u32* s_o = L.o;
  u64* s_top = L.top;
  uint8_t* s_f = L.f;
  auto& s_whist = L.whist;
  u32* s_dig = L.dig;
  u32* s_tmp = L.tmp;
  u64* s_red = L.red;
  const u32 tid = threadIdx.x, wv = tid / WAVE, lane = tid & (WAVE - 1);
  const u32 lo = win.lo, hi = win.hi, s0 = win.s0, s1 = win.s1;
  if (lo >= hi) return false;
  const u32 cnt = hi - lo;
  const u64 kmask = kb >= 64 ? ~0ull : ((1ull << kb) - 1);
  // stage rows (coalesced) and mark segment starts
  for (u32 i = tid; i < (u32)T; i += NT) s_f[i] = 0;
  __syncthreads();
  // both global read phases (rows, segment starts) in flight together
  for (u32 i = tid; i < cnt; i += NT) {
    s_c[i] = (u64)kin[lo + i];
    s_o[i] = vin[lo + i];
  }
  for (u32 s = s0 + tid; s <= s1; s += NT) {
    u32 st = seg_start[s];
    if (st >= lo && st < hi) s_f[st - lo] = 1;
  }
  if (tid == 0) s_f[0] = 1;
  __syncthreads();
  // blocked scan of the flags -> local segment index; build composites
  u32 f[E];
  u32 sum = 0;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    f[e] = s_f[tid * E + e];
    sum += f[e];
  }
  u32 nsegs;
  u32 run = block_excl_scan<NT>(sum, s_tmp, &nsegs);
  // Key bits above the highest one that varies inside the window are the
  // same in every row: the composite needs only the kbe bits up to it (the
  // rest is restored from s_top with the segment's partitioned top bits).
  u64 kor = 0, kand = ~0ull;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const u32 p = tid * E + e;
    if (p < cnt) {
      const u64 k = s_c[p] & kmask;
      kor |= k;
      kand &= k;
    }
  }
  u64 kcor, kcand;
  block_or_and<NT>(kor, kand, s_red, &kcor, &kcand);  // contains __syncthreads
  const u64 kvary = kcor ^ kcand;
  const u32 kbe = kvary ? 64u - (u32)__clzll((long long)kvary) : 0u;
  u32 segbits = 0;
  while (segbits < 32 && (1u << segbits) < nsegs) ++segbits;
  if (kbe + segbits > 64) {
    // no room beside the key for the segment index: the comparison rank
    // (k_tile_rank over a.rklist) sorts this window
    if (tid == 0 && rklist) rklist[atomicAdd(rkcount, 1u)] = w;
    return true;
  }
  const u64 emask = kbe >= 64 ? ~0ull : ((1ull << kbe) - 1);
  u64 vor = 0, vand = ~0ull;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    u32 p = tid * E + e;
    run += f[e];
    if (p < cnt) {
      u64 k = s_c[p];
      u32 sg = run - 1;
      if (f[e]) s_top[sg] = k & ~emask;
      u64 c = (kbe >= 64 ? 0ull : ((u64)sg << kbe)) | (k & emask);
      s_c[p] = c;
      vor |= c;
      vand &= c;
    }
  }
  u64 cor, cand;
  block_or_and<NT>(vor, vand, s_red, &cor, &cand);  // contains __syncthreads
  const u64 vary = cor ^ cand;
  const u32 B = kbe + segbits;
  // Rows are spread over all waves: Ee items per lane, so the ranking work
  // follows the rows actually in the tile, not its capacity.
  const u32 Ee = (cnt + NT - 1) / NT;
  const u32 WEe = WAVE * Ee;
  const u32 jn = (cnt > wv * WEe) ? min(Ee, (cnt - wv * WEe + WAVE - 1) / WAVE) : 0u;
  for (u32 shift = 0; shift < B; shift += RADIX_BITS) {
    const u32 dbits = min((u32)RADIX_BITS, B - shift);
    const u32 dmask = (1u << dbits) - 1;
    if (((vary >> shift) & dmask) == 0) continue;  // every row has the same digit
    u64 c[E];
    u32 o[E], dig[E], rank[E];
    bool ok[E];
#pragma unroll
    for (int j = 0; j < E; ++j) {
      u32 idx = wv * WEe + j * WAVE + lane;
      ok[j] = (u32)j < Ee && idx < cnt;
      c[j] = ok[j] ? s_c[idx] : 0ull;
      o[j] = ok[j] ? s_o[idx] : 0u;
      dig[j] = (u32)(c[j] >> shift) & dmask;
    }
    for (u32 i = tid; i < (u32)(NW * RADIX); i += NT) (&s_whist[0][0])[i] = 0;
    __syncthreads();
    wave_rank<E>(dig, ok, dbits, s_whist[wv], rank, jn);
    __syncthreads();
    {
      u32 d = tid, r = 0;
      for (int w = 0; w < NW; ++w) {
        u32 x = s_whist[w][d];
        s_whist[w][d] = r;
        r += x;
      }
      u32 tot;
      s_dig[d] = block_excl_scan<NT>(r, s_tmp, &tot);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < E; ++j) {
      if (ok[j]) {
        u32 pos = s_dig[dig[j]] + s_whist[wv][dig[j]] + rank[j];
        s_c[pos] = c[j];
        s_o[pos] = o[j];
      }
    }
    __syncthreads();
  }
  for (u32 i = tid; i < cnt; i += NT) {
    u64 c = s_c[i];
    u32 sg = (u32)(c >> kbe);
    if (kbe >= 64) sg = 0;
    u64 k = s_top[sg] | (c & emask);
    if (kout) kout[lo + i] = (K)k;
    vout[lo + i] = s_o[i];
  }
  return false;
}

// ---------------------------------------------------------------------------
// Fast path of the tile sort: a per-segment bucket sort.  Segment s of m_s
// rows gets 2^ceil(log2 m_s) bins (about one row per bin) cut from the top of
// its unsorted key bits; bins are laid out in segment order.  Rows are
// counted into bins (LDS atomics), the counts scanned, rows scattered into
// their bin's range, and every row then ranks itself inside its (short,
// contiguous) bin by (key, oid): its final place is bin start + rank.  The
// work no longer grows with the key width and has no dependent LDS chains.
// A tile whose largest bin exceeds BUCKET_LIMIT rows (heavy duplicates or
// skew) is flagged in `fallback` and left to the radix tile sort.
// ---------------------------------------------------------------------------

// LDS vectors of the bucket sort's fixed-layout phases: N (2, 4 or 8)
// consecutive u8 / u16 / u32 at an address aligned to their total size, in
// one LDS instruction (or two for 8 u32) instead of N.  The sort is bound
// by its LDS instructions, so the per-thread runs of flags, segment ids,
// segment starts and bin counts move as vectors.
template <int N>
__device__ __forceinline__ void lds_ld_u8(const uint8_t* p, u32 (&v)[N]) {
  if constexpr (N == 8) {
    const uint2 x = *(const uint2*)p;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      v[i] = (x.x >> (8 * i)) & 0xffu;
      v[4 + i] = (x.y >> (8 * i)) & 0xffu;
    }
  } else if constexpr (N == 4) {
    const u32 x = *(const u32*)p;
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = (x >> (8 * i)) & 0xffu;
  } else {
    const u32 x = *(const u16*)p;
    v[0] = x & 0xffu;
    v[1] = x >> 8;
  }
}
template <int N>
__device__ __forceinline__ void lds_st_u8(uint8_t* p, const u32 (&v)[N]) {
  auto pack4 = [&](int i) {
    return (v[i] & 0xffu) | ((v[i + 1] & 0xffu) << 8) | ((v[i + 2] & 0xffu) << 16) | (v[i + 3] << 24);
  };
  if constexpr (N == 8)
    *(uint2*)p = make_uint2(pack4(0), pack4(4));
  else if constexpr (N == 4)
    *(u32*)p = pack4(0);
  else
    *(u16*)p = (u16)((v[0] & 0xffu) | (v[1] << 8));
}
template <int N>
__device__ __forceinline__ void lds_ld_u16(const u16* p, u32 (&v)[N]) {
  u32 w[N / 2];
  if constexpr (N == 8) {
    const uint4 x = *(const uint4*)p;
    w[0] = x.x; w[1] = x.y; w[2] = x.z; w[3] = x.w;
  } else if constexpr (N == 4) {
    const uint2 x = *(const uint2*)p;
    w[0] = x.x; w[1] = x.y;
  } else {
    w[0] = *(const u32*)p;
  }
#pragma unroll
  for (int i = 0; i < N / 2; ++i) {
    v[2 * i] = w[i] & 0xffffu;
    v[2 * i + 1] = w[i] >> 16;
  }
}
template <int N>
__device__ __forceinline__ void lds_st_u16(u16* p, const u32 (&v)[N]) {
  u32 w[N / 2];
#pragma unroll
  for (int i = 0; i < N / 2; ++i) w[i] = (v[2 * i] & 0xffffu) | (v[2 * i + 1] << 16);
  if constexpr (N == 8)
    *(uint4*)p = make_uint4(w[0], w[1], w[2], w[3]);
  else if constexpr (N == 4)
    *(uint2*)p = make_uint2(w[0], w[1]);
  else
    *(u32*)p = w[0];
}
template <int N>
__device__ __forceinline__ void lds_ld_u32(const u32* p, u32 (&v)[N]) {
#pragma unroll
  for (int i = 0; i < N; i += 4) {
    if constexpr (N >= 4) {
      const uint4 x = *(const uint4*)(p + i);
      v[i] = x.x; v[i + 1] = x.y; v[i + 2] = x.z; v[i + 3] = x.w;
    } else {
      const uint2 x = *(const uint2*)p;
      v[0] = x.x; v[1] = x.y;
    }
  }
}
template <int N>
__device__ __forceinline__ void lds_st_u32(u32* p, const u32 (&v)[N]) {
#pragma unroll
  for (int i = 0; i < N; i += 4) {
    if constexpr (N >= 4)
      *(uint4*)(p + i) = make_uint4(v[i], v[i + 1], v[i + 2], v[i + 3]);
    else
      *(uint2*)p = make_uint2(v[0], v[1]);
  }
}

// LDS of the bucket sort.  The segment arrays are dead once the rows are
// scattered into their bins, so they share LDS with the row buffers (fewer
// bytes a block, more windows in flight a CU): region A holds the rows' keys
// from the scatter on and, before it, the local segment ids [0, 2T), the
// segments' first bins [2T, 4T), their log2 bin counts [4T, 5T) and the
// segment-start flags [5T, 6T); region B holds the oids from the scatter on
// and, before it, the segment starts.
template <class K, int E>
struct BucketLds {
  static constexpr int NT = 256, T = NT * E, NB = T;
  static constexpr int AB = (int)sizeof(K) * T > 6 * T ? (int)sizeof(K) * T : 6 * T;
  static constexpr int BB = 4 * T;
  __align__(16) unsigned char raw[AB + BB];
  u32 cnt[NB];
  u32 tmp[2 * (NT / WAVE)];  // two buffers of the one-barrier block scans
  u32 over;
};

// A window's range, its segment range and the starts of its first 256 E
// segments (sst[e]: segment s0 + e * 256 + tid), and its rows in registers
// (row p = e * 256 + tid).
template <class K, int E>
struct WinRows {
  u32 w, lo, hi, s0, s1;
  u32 sst[E];
  K kk[E];
  u32 oo[E];
};

// The segment starts are loaded before the rows: the flag phase needs only
// them, and memory operations complete in order, so it need not wait for
// the rows.  Both without per-row branches: indices past the end repeat the
// last segment / row (a flag set twice; rows past cnt are never used).
// The rows and segment starts of window (lo, hi, s0, s1) (r.w is the
// caller's), loaded as described above.
template <class K, int E>
__device__ __forceinline__ void load_rows(const K* __restrict__ kin, const u32* __restrict__ vin,
                                          const u32* __restrict__ seg_start, u32 lo, u32 hi, u32 s0, u32 s1,
                                          WinRows<K, E>& r) {
  r.lo = lo;
  r.hi = hi;
  r.s0 = s0;
  r.s1 = s1;
  const u32 cnt = r.hi > r.lo ? r.hi - r.lo : 0u;
  if (cnt == 0 || r.s1 < r.s0) {
    r.hi = r.lo;
    return;
  }
  const u32 nsg = r.s1 - r.s0 + 1;
#pragma unroll
  for (int e = 0; e < E; ++e) r.sst[e] = seg_start[r.s0 + min(e * 256 + threadIdx.x, nsg - 1)];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const u32 p = min(e * 256 + threadIdx.x, cnt - 1);
    r.kk[e] = kin[r.lo + p];
    r.oo[e] = vin[r.lo + p];
  }
}

// The window's rows r of kin (in registers) sorted into kout/vout at the
// same rows.  Returns true, without writing anything, when a bin holds more
// than BUCKET_LIMIT rows: the radix tile sort (tile_sort_win) sorts such a
// window.  A key with a bit in bad sets *err (when err is set).
template <class K, int E>
__device__ __forceinline__ bool tile_bucket_body(const WinRows<K, E>& r, const u32* __restrict__ seg_start, u32 kb,
                                                 K* kout, u32* vout, u64 bad, u32* err, BucketLds<K, E>& L) {
  constexpr int NT = 256, T = NT * E, NB = T, BPT = NB / NT, SE = T / NT;
  constexpr int AB = BucketLds<K, E>::AB;
  unsigned char* s_raw = L.raw;
  u32* s_cnt = L.cnt;  // bin counts -> bin cursors -> bin ends
  u32* s_tmp = L.tmp;
  u32& s_over = L.over;
  K* s_k = (K*)s_raw;                           // rows in bin order, then in final order
  u32* s_o = (u32*)(s_raw + AB);
  u16* s_sid = (u16*)s_raw;                     // local segment of every row position
  u16* s_segbase = (u16*)(s_raw + 2 * T);       // first bin of every segment
  uint8_t* s_lb = s_raw + 4 * T;                // log2 bins of every segment
  uint8_t* s_f = s_raw + 5 * T;                 // segment-start flags
  u16* s_segstart = (u16*)(s_raw + AB);         // T + 1 entries
  const u32 tid = threadIdx.x;
  const u32 lo = r.lo, hi = r.hi;
  if (lo >= hi) return false;
  if (err) {  // rows past the window's end repeat its last row
    u64 b = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) b |= (u64)r.kk[e] & bad;
    if (__ballot(b != 0) && __lane_id() == 0) atomicOr(err, 1u);
  }
  const u32 s0 = r.s0, s1 = r.s1;
  const u32 cnt = hi - lo;
  const u64 kmask = kb >= 64 ? ~0ull : ((1ull << kb) - 1);
  const K (&kk)[E] = r.kk;
  const u32 (&oo)[E] = r.oo;
  static_assert(E == 2 || E == 4 || E == 8, "LDS vectors of 2, 4 or 8");
  // flags and bin counts cleared with 16-byte stores (s_f at 5T and the
  // counts are 16-byte aligned)
  for (u32 i = tid; i < (u32)T / 16; i += NT) ((uint4*)s_f)[i] = make_uint4(0u, 0u, 0u, 0u);
  for (u32 i = tid; i < (u32)NB / 4; i += NT) ((uint4*)s_cnt)[i] = make_uint4(0u, 0u, 0u, 0u);
  if (tid == 0) s_over = 0;
  __syncthreads();
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const u32 st = r.sst[e];
    if (st >= lo && st < hi) s_f[st - lo] = 1;
  }
  for (u32 s = s0 + T + tid; s <= s1; s += NT) {  // segments beyond the loaded ones
    const u32 st = seg_start[s];
    if (st >= lo && st < hi) s_f[st - lo] = 1;
  }
  if (tid == 0) s_f[0] = 1;
  __syncthreads();
  // local segment of every position (blocked scan of the flags)
  {
    u32 f[E];
    lds_ld_u8<E>(s_f + tid * E, f);
    u32 sum = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) sum += f[e];
    u32 nsegs;
    u32 run = block_excl_scan1<NT>(sum, s_tmp, &nsegs);
    u32 sid[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      run += f[e];
      const u32 p = tid * E + e;
      sid[e] = run - 1;
      if (f[e] && p < cnt) s_segstart[run - 1] = (u16)p;
    }
    lds_st_u16<E>(s_sid + tid * E, sid);
    if (tid == 0) s_segstart[nsegs] = (u16)cnt;
    __syncthreads();
    // bins per segment: 2^lb >= m, laid out in segment order (entries past
    // nsegs are computed from stale starts and never read)
    u32 ss[SE];
    lds_ld_u16<SE>(s_segstart + tid * SE, ss);
    const u32 snext = s_segstart[tid * SE + SE];  // T + 1 entries
    u32 bsum = 0;
    u32 lbs[SE];
#pragma unroll
    for (int e = 0; e < SE; ++e) {
      const u32 sg = tid * SE + e;
      u32 lb = 0;
      if (sg < nsegs) {
        const u32 m = (e + 1 < SE ? ss[e + 1] : snext) - ss[e];
        lb = m <= 2 ? 0u : 31u - __clz(m - 1);  // 2^lb >= m / 2: about two rows a bin
        bsum += 1u << lb;
      }
      lbs[e] = lb;
    }
    u32 btot;
    u32 brun = block_excl_scan1<NT>(bsum, s_tmp + NT / WAVE, &btot);
    u32 segb[SE];
#pragma unroll
    for (int e = 0; e < SE; ++e) {
      segb[e] = brun;
      if (tid * SE + e < nsegs) brun += 1u << lbs[e];
    }
    lds_st_u8<SE>(s_lb + tid * SE, lbs);
    lds_st_u16<SE>(s_segbase + tid * SE, segb);
  }
  __syncthreads();
  // bin of every row, and the bin counts
  u32 bin[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const u32 p = e * NT + tid;
    bin[e] = 0;
    if (p < cnt) {
      const u32 sg = s_sid[p];
      const u64 km = (u64)kk[e] & kmask;
      const u32 lb = s_lb[sg];
      u32 off = 0;
      if (lb) off = kb >= lb ? (u32)(km >> (kb - lb)) : (u32)(km << (lb - kb));
      bin[e] = (u32)s_segbase[sg] + off;
      atomicAdd(&s_cnt[bin[e]], 1u);
    }
  }
  __syncthreads();
  // exclusive scan of the counts (BPT consecutive bins per thread) -> cursors
  {
    u32 c[BPT];
    lds_ld_u32<BPT>(s_cnt + tid * BPT, c);
    u32 csum = 0, cmax = 0;
#pragma unroll
    for (int j = 0; j < BPT; ++j) {
      csum += c[j];
      cmax = max(cmax, c[j]);
    }
    if (cmax > BUCKET_LIMIT) s_over = 1;
    u32 tot;
    u32 crun = block_excl_scan1<NT>(csum, s_tmp, &tot);
    if (s_over) return true;
    u32 cur[BPT];
#pragma unroll
    for (int j = 0; j < BPT; ++j) {
      cur[j] = crun;
      crun += c[j];
    }
    lds_st_u32<BPT>(s_cnt + tid * BPT, cur);
  }
  __syncthreads();
  // scatter rows into their bins (order inside a bin is arbitrary); the
  // cursors end as the bin ends, and bin b starts where bin b - 1 ends
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const u32 p = e * NT + tid;
    if (p < cnt) {
      const u32 slot = atomicAdd(&s_cnt[bin[e]], 1u);
      s_k[slot] = kk[e];
      s_o[slot] = oo[e];
    }
  }
  __syncthreads();
  // rank inside the bin by (key, oid)
  u32 rk[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const u32 p = e * NT + tid;
    rk[e] = 0;
    if (p < cnt) {
      const u32 end = s_cnt[bin[e]], start = bin[e] ? s_cnt[bin[e] - 1] : 0u;
      u32 r = start;
      for (u32 q = start; q < end; ++q) {
        const K kq = s_k[q];
        r += (kq < kk[e]) || (kq == kk[e] && s_o[q] < oo[e]);
      }
      rk[e] = r;
    }
  }
  __syncthreads();
#pragma unroll
  for (int e = 0; e < E; ++e) {
    if (e * NT + tid < cnt) {
      s_k[rk[e]] = kk[e];
      s_o[rk[e]] = oo[e];
    }
  }
  __syncthreads();
  for (u32 i = tid; i < cnt; i += NT) {
    if (kout) kout[lo + i] = s_k[i];
    vout[lo + i] = s_o[i];
  }
  return false;
}

// Fallback of the bucket tile sort for keys too wide to pack a segment index
// beside them (kb + 11 > 64 bits): every row's place inside its segment is
// the number of rows of the segment ordered before it by (key, oid), found
// by comparing against the whole segment (segments hold at most small_max
// rows; only tiles the bucket sort refused come here).
template <class K, int E>
struct RankLds {
  static constexpr int NT = 256, T = NT * E;
  K k[T];
  u32 o[T];
  u16 sid[T];
  u16 start[T + 1];
  uint8_t f[T];
  u32 tmp[NT / WAVE + 1];
};

template <class K, int E>
__device__ __forceinline__ void tile_rank_win(const K* __restrict__ kin, const u32* __restrict__ vin, K* kout,
                                              u32* vout, const u32* __restrict__ seg_start, TileWin win,
                                              RankLds<K, E>& L) {
  constexpr int NT = 256, T = NT * E;
  K* s_k = L.k;
  u32* s_o = L.o;
  u16* s_sid = L.sid;
  u16* s_start = L.start;
  uint8_t* s_f = L.f;
  u32* s_tmp = L.tmp;
  const u32 tid = threadIdx.x;
  const u32 lo = win.lo, hi = win.hi;
  if (lo >= hi) return;
  const u32 s0 = win.s0, s1 = win.s1;
  const u32 cnt = hi - lo;
  for (u32 i = tid; i < (u32)T; i += NT) s_f[i] = 0;
  __syncthreads();
  for (u32 i = tid; i < cnt; i += NT) {
    s_k[i] = kin[lo + i];
    s_o[i] = vin[lo + i];
  }
  for (u32 s = s0 + tid; s <= s1; s += NT) {
    const u32 st = seg_start[s];
    if (st >= lo && st < hi) s_f[st - lo] = 1;
  }
  if (tid == 0) s_f[0] = 1;
  __syncthreads();
  u32 f[E];
  u32 sum = 0;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    f[e] = s_f[tid * E + e];
    sum += f[e];
  }
  u32 nsegs;
  u32 run = block_excl_scan<NT>(sum, s_tmp, &nsegs);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    run += f[e];
    const u32 p = tid * E + e;
    s_sid[p] = (u16)(run - 1);
    if (f[e] && p < cnt) s_start[run - 1] = (u16)p;
  }
  if (tid == 0) s_start[nsegs] = (u16)cnt;
  __syncthreads();
  for (u32 p = tid; p < cnt; p += NT) {
    const u32 sg = s_sid[p];
    const u32 b0 = s_start[sg], b1 = s_start[sg + 1];
    const K kp = s_k[p];
    const u32 op = s_o[p];
    u32 r = b0;
    for (u32 q = b0; q < b1; ++q) {
      const K kq = s_k[q];
      r += (kq < kp) || (kq == kp && s_o[q] < op);
    }
    if (kout) kout[lo + r] = kp;
    vout[lo + r] = op;
  }
}


}  // namespace fmcs
