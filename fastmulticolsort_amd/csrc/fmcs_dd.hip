// fmcs_dd.hip -- device-driven list kernels of the batched segmented sort.
//
// Every size these kernels work on (segment-list lengths, large-segment and
// tile counts, window counts) is read from device memory, and the launches
// are grid-stride or persistent, so a whole round runs without one host
// synchronisation and can be captured in a hipGraph.
#include "fmcs_kernels.h"

namespace fmcs {

__global__ void kd_set4(u32* __restrict__ p, u32 a, u32 b, u32 c, u32 d) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    p[0] = a;
    p[1] = b;
    p[2] = c;
    p[3] = d;
  }
}

__global__ void kd_offsets_to_list(const u32* __restrict__ off, const u32* __restrict__ nseg_d, u32* __restrict__ st,
                                   u32* __restrict__ sz) {
  const u32 nseg = *nseg_d;
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < nseg; i += gridDim.x * blockDim.x) {
    u32 a = off[i], b = off[i + 1];
    st[i] = a;
    sz[i] = b - a;
  }
}

// Window packing (see k_windows) over device-sized lists; an empty segment
// is placed at start - 1 within its parent.  A block takes WIN_CHUNK
// consecutive segments at a time, collects the windows that start a run in
// LDS (one LDS atomic per wave) and appends them to the compact list with
// one global atomic per chunk, so the list counter sees few contended
// updates.
constexpr u32 WIN_CHUNK = 2048;
// Segments of small_max + 1 .. single_max rows (single_max > small_max: a
// whole tile) are listed apart, each one its own window: (slo, shi, sseg)
// appended at *scount, collected in LDS like the windows (one global atomic
// per chunk).
__global__ __launch_bounds__(256) void kd_windows(const u32* __restrict__ st, const u32* __restrict__ sz,
                                                  const u32* __restrict__ nseg_d, u32 W, u32 small_max,
                                                  const u32* __restrict__ plo, const u32* __restrict__ pwb, u32 wpt,
                                                  u32* __restrict__ tlo, u32* __restrict__ thi, u32* __restrict__ ts0,
                                                  u32* __restrict__ ts1, u32* __restrict__ wlist,
                                                  u32* __restrict__ wcount, u32 single_max, u32* __restrict__ slo,
                                                  u32* __restrict__ shi, u32* __restrict__ sseg,
                                                  u32* __restrict__ scount, u32* __restrict__ scan_status,
                                                  u32* __restrict__ scan_status2) {
  constexpr u32 PER = WIN_CHUNK / 256;
  __shared__ u32 s_w[WIN_CHUNK];
  // the chunk's segments' windows and sizes at 1 .. WIN_CHUNK, the segment
  // before the chunk at 0 and the one after it at WIN_CHUNK + 1
  __shared__ u32 s_wid[WIN_CHUNK + 2];
  __shared__ u32 s_sz[WIN_CHUNK + 2];
  __shared__ u32 s_si[WIN_CHUNK];  // the chunk's single segments
  __shared__ u32 s_n, s_base, s_ns, s_sbase;
  const u32 nseg = *nseg_d;
  const u32 tid = threadIdx.x, lane = __lane_id();
  // the look-back status words of the level's two scans: over the segment
  // sizes, and over the large segments' tile counts (no more of those)
  const u32 nst = (nseg + SCAN_TILE - 1) / SCAN_TILE;
  for (u32 i = blockIdx.x * blockDim.x + tid; i < nst; i += gridDim.x * blockDim.x) {
    scan_status[i] = 0;
    scan_status2[i] = 0;
  }
  // window of a segment from its start s, size z and (below the first level)
  // its parent's first row lo and first window pb
  auto window = [&](u32 s, u32 z, u32 lo, u32 pb) -> u32 {
    if (!plo) return ((z == 0 && s > 0) ? s - 1 : s) / W;
    return pb * wpt + (((z == 0 && s > lo) ? s - 1 : s) - lo) / W;
  };
  for (u64 c0 = (u64)blockIdx.x * WIN_CHUNK; c0 < nseg; c0 += (u64)gridDim.x * WIN_CHUNK) {
    if (tid == 0) s_n = s_ns = 0;
    // every load of the chunk first, at clamped indices and without
    // branches, so they are in flight together
    u32 sv[PER], zv[PER], lv[PER], bv[PER];
#pragma unroll
    for (u32 j = 0; j < PER; ++j) {
      const u32 ic = (u32)min(c0 + j * 256 + tid, (u64)nseg - 1);
      sv[j] = st[ic];
      zv[j] = sz[ic];
      lv[j] = plo ? plo[ic >> RADIX_BITS] : 0u;
      bv[j] = plo ? pwb[ic >> RADIX_BITS] : 0u;
    }
    // threads 0 and 1: the segments either side of the chunk
    const u32 ie = tid == 0 ? (u32)(c0 > 0 ? c0 - 1 : 0) : (u32)min(c0 + WIN_CHUNK, (u64)nseg - 1);
    const u32 es = st[ie], ez = sz[ie];
    const u32 el = plo ? plo[ie >> RADIX_BITS] : 0u, eb = plo ? pwb[ie >> RADIX_BITS] : 0u;
#pragma unroll
    for (u32 j = 0; j < PER; ++j) {
      s_wid[1 + j * 256 + tid] = window(sv[j], zv[j], lv[j], bv[j]);
      s_sz[1 + j * 256 + tid] = zv[j];
    }
    if (tid < 2) {
      s_wid[tid == 0 ? 0 : WIN_CHUNK + 1] = window(es, ez, el, eb);
      s_sz[tid == 0 ? 0 : WIN_CHUNK + 1] = ez;
    }
    __syncthreads();
#pragma unroll
    for (u32 j = 0; j < PER; ++j) {
      const u64 i64 = c0 + j * 256 + tid;
      const u32 k = 1 + j * 256 + tid;
      bool first = false, single = false;
      u32 w = 0, size = 0;
      const u32 i = (u32)i64;
      if (i64 < nseg) {
        size = s_sz[k];
        if (size <= small_max) {
          w = s_wid[k];
          first = (i == 0) || (s_wid[k - 1] != w);
          const bool last = (i + 1 == nseg) || (s_wid[k + 1] != w) || (s_sz[k + 1] > small_max);
          if (first) { tlo[w] = sv[j]; ts0[w] = i; }
          if (last) { thi[w] = sv[j] + size; ts1[w] = i; }
        } else {
          single = size <= single_max;
        }
      }
      const u64 ms = __ballot(single);
      if (ms) {
        const u32 leader = (u32)(__ffsll((long long)ms) - 1);
        u32 b = 0;
        if (lane == leader) b = atomicAdd(&s_ns, (u32)__popcll(ms));
        b = __shfl(b, (int)leader);
        if (single) s_si[b + (u32)__popcll(ms & lanemask_lt())] = i;
      }
      const u64 m = __ballot(first);
      if (m) {
        const u32 leader = (u32)(__ffsll((long long)m) - 1);
        u32 b = 0;
        if (lane == leader) b = atomicAdd(&s_n, (u32)__popcll(m));
        b = __shfl(b, (int)leader);
        if (first) s_w[b + (u32)__popcll(m & lanemask_lt())] = w;
      }
    }
    __syncthreads();
    if (tid == 0) {
      s_base = s_n ? atomicAdd(wcount, s_n) : 0u;
      s_sbase = s_ns ? atomicAdd(scount, s_ns) : 0u;
    }
    __syncthreads();
    for (u32 k = tid; k < s_n; k += 256) wlist[s_base + k] = s_w[k];
    for (u32 k = tid; k < s_ns; k += 256) {
      const u32 i = s_si[k];
      const u32 s0 = st[i];
      slo[s_sbase + k] = s0;
      shi[s_sbase + k] = s0 + s_sz[1 + (i - (u32)c0)];
      sseg[s_sbase + k] = i;
    }
    __syncthreads();
  }
}

// Exclusive scan with decoupled look-back, persistent: blocks take tiles from
// the counter until ceil(n / TILE) are done.  *total = the sum (0 when n = 0).
// FLAGS: in holds the sizes of segments starting at seg_start and the
// scanned value is (size > small_max): *total is the large segments' count,
// and large segment i, the p-th, gets llo[p] = its first row, lhi[p] = its
// end, ltiles[p] = its tiles of tile_rows rows, and (hist set) zeroed digit
// counts hist[p * 256 .. + 256); out is not written.  rows (if set) [0] and
// [1] get the small and the large segments' rows.
template <int ITEMS, bool FLAGS>
__global__ __launch_bounds__(256) void kd_scan_excl(const u32* __restrict__ in, u32* __restrict__ out,
                                                    const u32* __restrict__ n_d, u32* __restrict__ status,
                                                    u32* __restrict__ counter, u32* __restrict__ total,
                                                    u32 small_max, u32* __restrict__ rows,
                                                    const u32* __restrict__ seg_start, u32 tile_rows,
                                                    u32* __restrict__ llo, u32* __restrict__ lhi,
                                                    u32* __restrict__ ltiles, u32* __restrict__ hist) {
  constexpr int NT = 256, TILE = NT * ITEMS;
  __shared__ u32 s_tile, s_excl;
  __shared__ u32 s_tmp[NT / WAVE + 1];
  __shared__ u32 s_lpos[FLAGS ? TILE : 1];  // FLAGS: the tile's large segments' places
  const u32 n = *n_d;
  const u32 ntiles = (n + TILE - 1) / TILE;
  const u32 tid = threadIdx.x;
  if (n == 0) {
    if (blockIdx.x == 0 && tid == 0) *total = 0;
    return;
  }
  u32 rs = 0, rl = 0;  // FLAGS: this thread's small and large rows
  __shared__ u32 s_rows[2];
  if (FLAGS && tid == 0) s_rows[0] = s_rows[1] = 0;
  for (;;) {
    if (tid == 0) s_tile = atomicAdd(counter, 1u);
    __syncthreads();
    const u32 tile = s_tile;
    if (tile >= ntiles) break;
    const u64 base = (u64)tile * TILE + (u64)tid * ITEMS;
    u32 v[ITEMS];
    u32 zs[FLAGS ? ITEMS : 1], ss[FLAGS ? ITEMS : 1];  // FLAGS: sizes and starts
    u32 sum = 0;
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) v[j] = (base + j < n) ? in[base + j] : 0u;
    if constexpr (FLAGS) {
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) ss[j] = seg_start[min(base + j, (u64)n - 1)];
    }
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      if constexpr (FLAGS) {
        zs[j] = v[j];
        const bool large = v[j] > small_max;
        rs += large ? 0u : v[j];
        rl += large ? v[j] : 0u;
        v[j] = large ? 1u : 0u;
      }
      sum += v[j];
    }
    u32 tot;
    u32 ex = block_excl_scan<NT>(sum, s_tmp, &tot);
    if (tid < WAVE) {
      const u32 lane = tid;
      u32 e = 0;
      if (tile == 0) {
        if (lane == 0) status_store(status, ST_INC | tot);
      } else {
        if (lane == 0) status_store(status + tile, ST_AGG | tot);
        e = lookback_wave(status, tile);
        if (lane == 0) status_store(status + tile, ST_INC | (e + tot));
      }
      if (lane == 0) {
        s_excl = e;
        if (tile + 1 == ntiles) *total = e + tot;
      }
    }
    __syncthreads();
    u32 run = s_excl + ex;
    if constexpr (FLAGS) {
      // rows past n were read as size 0: never large
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) {
        if (v[j]) {
          llo[run] = ss[j];
          lhi[run] = ss[j] + zs[j];
          ltiles[run] = (zs[j] + tile_rows - 1) / tile_rows;
          s_lpos[run - s_excl] = run;
        }
        run += v[j];
      }
      if (hist) {
        __syncthreads();
        // one 256-count row per large segment, a row per block step
        for (u32 k = 0; k < tot; ++k) hist[(u64)s_lpos[k] * RADIX + tid] = 0;
      }
    } else {
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) {
        if (base + j < n) out[base + j] = run;
        run += v[j];
      }
    }
    __syncthreads();
  }
  if constexpr (FLAGS) {
    // every thread left the loop together (s_tile is block-uniform); one
    // global atomic per block and counter
    if (rows) {
      rs = __shfl(wave_incl_scan(rs), WAVE - 1);
      rl = __shfl(wave_incl_scan(rl), WAVE - 1);
      if ((tid & (WAVE - 1)) == 0) {
        atomicAdd(&s_rows[0], rs);
        atomicAdd(&s_rows[1], rl);
      }
      __syncthreads();
      if (tid == 0) {
        if (s_rows[0]) atomicAdd(rows, s_rows[0]);
        if (s_rows[1]) atomicAdd(rows + 1, s_rows[1]);
      }
    }
  }
}

// Per-segment digit histogram, tiles never crossing a segment, grid-stride
// over the *ntiles_d tiles of the *nsegs_d segments.  tile_counts (if set)
// gets every tile's 256 counts, the input of the tile-offset scan of the
// scatter pass (same tiling); hist (if set) gets the per-segment totals.
// A block keeps its digit's sum in a register while its tiles stay in one
// segment and adds it to hist when the segment changes, so a huge segment
// (the whole array in an LSD pass) sees one global atomic per block and
// digit.  Full tiles that start on a 16-byte boundary are read with 16-byte
// loads (the rate of 4-byte loads is the limit for 16- and 32-bit keys).
// A key with a bit in bad_mask (above the declared width) sets *err.
// tile_seg (if set) gets every tile's segment, for the passes after it.
template <class K, int ITEMS, int RB>
__global__ __launch_bounds__(256) void kd_seg_hist(const K* __restrict__ keys, const u32* __restrict__ seg_lo,
                                                   const u32* __restrict__ seg_hi, const u32* __restrict__ tile_base,
                                                   const u32* __restrict__ nsegs_d, const u32* __restrict__ ntiles_d,
                                                   u32 shift, u32 dbits, u32* __restrict__ hist,
                                                   u32* __restrict__ tile_counts, u64 bad_mask,
                                                   u32* __restrict__ err, u32* __restrict__ tile_seg) {
  constexpr int TILE = 256 * ITEMS;
  constexpr int PV = 16 / sizeof(K);         // keys per 16-byte vector
  constexpr int NVT = TILE / PV;             // vectors in a full tile
  constexpr int NV = (NVT + 255) / 256;      // per thread (the last one ragged unless NVT % 256 == 0)
  constexpr bool VEC = TILE % PV == 0;       // a full tile is whole vectors
  // HCOPY sub-histograms, lane l counting into copy l % HCOPY: lanes of one
  // instruction that share a digit mostly hit different copies, so a skewed
  // key does not serialize the LDS atomics.  Copies are HSTRIDE words apart
  // so the copies of one digit sit in different banks.
  constexpr u32 R = 1u << RB, DPT = R / 256;  // RB-bit digits; thread t owns digits t + 256 q
  constexpr u32 HCOPY = RB > 8 ? 4 : 8, HSTRIDE = R + 1;
  __shared__ u32 sh[HCOPY * HSTRIDE];
  __shared__ u32 s_seg;
  const u32 nsegs = *nsegs_d, ntiles = *ntiles_d;
  const u32 mask = (1u << dbits) - 1;
  u32* myh = sh + (threadIdx.x % HCOPY) * HSTRIDE;
  u32 acc[DPT];
#pragma unroll
  for (u32 q = 0; q < DPT; ++q) acc[q] = 0;
  u32 cur = 0xffffffffu;
  u64 bad = 0;
  for (u32 tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    for (u32 i = threadIdx.x; i < HCOPY * HSTRIDE; i += 256) sh[i] = 0;
    if (threadIdx.x == 0) {
      s_seg = upper_bound_m1(tile_base, nsegs, tile);
      if (tile_seg) tile_seg[tile] = s_seg;
    }
    __syncthreads();
    const u32 seg = s_seg;
    if (seg != cur) {
#pragma unroll
      for (u32 q = 0; q < DPT; ++q) {
        if (hist && acc[q]) atomicAdd(&hist[(u64)cur * R + threadIdx.x + 256 * q], acc[q]);
        acc[q] = 0;
      }
      cur = seg;
    }
    const u64 row0 = (u64)seg_lo[seg] + (u64)(tile - tile_base[seg]) * TILE;
    const u64 rowend = min((u64)seg_hi[seg], row0 + TILE);
    if (VEC && rowend - row0 == (u64)TILE && ((uintptr_t)(keys + row0) & 15) == 0) {
      const uint4* vp = (const uint4*)(keys + row0);
      uint4 x[VEC ? NV : 1];
      // every load unconditional (a ragged last vector re-reads the tile's
      // last one and is not counted)
#pragma unroll
      for (int v = 0; v < NV; ++v) x[v] = vp[min(v * 256 + (int)threadIdx.x, NVT - 1)];
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        if (NVT % 256 == 0 || v * 256 + (int)threadIdx.x < NVT) {
          K k[PV];
          unpack16<K>(x[v], k);
#pragma unroll
          for (int i = 0; i < PV; ++i) {
            atomicAdd(&myh[digit_of(k[i], shift, mask)], 1u);
            bad |= (u64)k[i] & bad_mask;
          }
        }
      }
    } else {
      // rows past the tile's end load its last row (no branch around the
      // loads, so their waits stay per load) and are not counted
      K k[ITEMS];
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) k[j] = keys[min(row0 + (u64)j * 256 + threadIdx.x, rowend - 1)];
#pragma unroll
      for (int j = 0; j < ITEMS; ++j)
        if (row0 + (u64)j * 256 + threadIdx.x < rowend) {
          atomicAdd(&myh[digit_of(k[j], shift, mask)], 1u);
          bad |= (u64)k[j] & bad_mask;
        }
    }
    __syncthreads();
#pragma unroll
    for (u32 q = 0; q < DPT; ++q) {
      const u32 d = threadIdx.x + 256 * q;
      u32 c = 0;
#pragma unroll
      for (u32 h = 0; h < HCOPY; ++h) c += sh[h * HSTRIDE + d];
      acc[q] += c;
      if (tile_counts) tile_counts[(u64)tile * R + d] = c;
    }
    __syncthreads();
  }
#pragma unroll
  for (u32 q = 0; q < DPT; ++q)
    if (hist && acc[q]) atomicAdd(&hist[(u64)cur * R + threadIdx.x + 256 * q], acc[q]);
  if (err && __ballot(bad != 0) && __lane_id() == 0) atomicOr(err, 1u);
}

// Per-segment exclusive scan of the tile counts along the tiles, in two
// steps.  kd_tscan_local: for every chunk of TSCAN_CT tiles, thread d
// replaces counts[t][d] by the sum of counts[t'][d] over earlier tiles t' of
// the chunk in the same segment; carry[c][d] = that running sum at the chunk
// end, cflag[c] = 1 when a segment starts inside chunk c.  tile_seg (if set)
// holds every tile's segment (else found by search over tile_base).
template <int RB>
__global__ __launch_bounds__(256) void kd_tscan_local(u32* __restrict__ counts, const u32* __restrict__ tile_base,
                                                      const u32* __restrict__ nsegs_d,
                                                      const u32* __restrict__ ntiles_d, u32* __restrict__ carry,
                                                      u32* __restrict__ cflag, const u32* __restrict__ tile_seg) {
  constexpr u32 R = 1u << RB;
  __shared__ u32 s_first[TSCAN_CT];
  const u32 nsegs = *nsegs_d, ntiles = *ntiles_d;
  const u32 nch = (ntiles + TSCAN_CT - 1) / TSCAN_CT;
  for (u32 c = blockIdx.x; c < nch; c += gridDim.x) {
    const u32 t0 = c * TSCAN_CT;
    const u32 m = min(ntiles - t0, TSCAN_CT);
    if (threadIdx.x < m) {
      const u32 t = t0 + threadIdx.x;
      s_first[threadIdx.x] = tile_seg ? (t == 0 || tile_seg[t] != tile_seg[t - 1])
                                      : tile_base[upper_bound_m1(tile_base, nsegs, t)] == t;
    }
    __syncthreads();
    u32 any = 0;
    for (u32 d = threadIdx.x; d < R; d += 256) {
      u32 v[TSCAN_CT];
#pragma unroll
      for (u32 j = 0; j < TSCAN_CT; ++j) v[j] = j < m ? counts[(u64)(t0 + j) * R + d] : 0u;
      u32 run = 0;
#pragma unroll
      for (u32 j = 0; j < TSCAN_CT; ++j) {
        if (j < m) {
          if (s_first[j]) {
            run = 0;
            any = 1;
          }
          counts[(u64)(t0 + j) * R + d] = run;
          run += v[j];
        }
      }
      carry[(u64)c * R + d] = run;
    }
    if (threadIdx.x == 0) cflag[c] = any;
    __syncthreads();
  }
}

// kd_tscan_carry: block d (grid = 2^RB) turns carry[.][d] into carry_in:
// the sum over earlier chunks back to (and including the tail of) the chunk
// where the segment of the chunk's first tile starts; a segmented exclusive
// scan over the chunks with resets at chunks that hold a segment start.
template <int RB>
__global__ __launch_bounds__(256) void kd_tscan_carry(u32* __restrict__ carry, const u32* __restrict__ cflag,
                                                      const u32* __restrict__ ntiles_d) {
  constexpr u32 R = 1u << RB;
  __shared__ u32 s_f[256], s_v[256];
  const u32 d = blockIdx.x, tid = threadIdx.x;
  const u32 nch = (*ntiles_d + TSCAN_CT - 1) / TSCAN_CT;
  const u32 per = (nch + 255) / 256;
  const u32 c0 = min(nch, tid * per), c1 = min(nch, c0 + per);
  // this thread's chunks reduced to (has reset, sum since the last reset)
  u32 f = 0, v = 0;
  for (u32 c = c0; c < c1; ++c) {
    const u32 x = carry[(u64)c * R + d];
    if (cflag[c]) {
      f = 1;
      v = x;
    } else {
      v += x;
    }
  }
  // inclusive segmented scan over the threads: (a) then (b) is
  // (a.f | b.f, b.f ? b.v : a.v + b.v)
  s_f[tid] = f;
  s_v[tid] = v;
  __syncthreads();
  for (u32 o = 1; o < 256; o <<= 1) {
    const u32 pf = tid >= o ? s_f[tid - o] : 0u, pv = tid >= o ? s_v[tid - o] : 0u;
    __syncthreads();
    if (tid >= o && !f) {
      v += pv;
      f |= pf;
    }
    s_f[tid] = f;
    s_v[tid] = v;
    __syncthreads();
  }
  u32 acc = tid ? s_v[tid - 1] : 0u;
  for (u32 c = c0; c < c1; ++c) {
    const u32 x = carry[(u64)c * R + d];
    carry[(u64)c * R + d] = acc;
    acc = cflag[c] ? x : acc + x;
  }
}

// digit_base[s][d] = lo[s] + prefix; bucket list (start, size) in (s, d)
// order; loops over *nsets_d sets.
template <int RB>
__global__ __launch_bounds__(256) void kd_digit_base(const u32* __restrict__ hist, const u32* __restrict__ lo,
                                                     u32* __restrict__ digit_base, u32* __restrict__ bstart,
                                                     u32* __restrict__ bsize, const u32* __restrict__ nsets_d,
                                                     u32* __restrict__ nbuckets) {
  constexpr u32 R = 1u << RB, DPT = R / 256;  // thread t owns digits t * DPT + q
  __shared__ u32 tmp[8];
  const u32 nsets = *nsets_d;
  // (if set) the bucket count, the next level's segment count
  if (nbuckets && blockIdx.x == 0 && threadIdx.x == 0) *nbuckets = nsets * R;
  for (u32 s = blockIdx.x; s < nsets; s += gridDim.x) {
    u32 c[DPT];
    u32 sum = 0;
#pragma unroll
    for (u32 q = 0; q < DPT; ++q) {
      c[q] = hist[(u64)s * R + threadIdx.x * DPT + q];
      sum += c[q];
    }
    u32 tot;
    u32 b = lo[s] + block_excl_scan<256>(sum, tmp, &tot);
#pragma unroll
    for (u32 q = 0; q < DPT; ++q) {
      const u32 d = threadIdx.x * DPT + q;
      digit_base[(u64)s * R + d] = b;
      if (bstart) {
        bstart[(u64)s * R + d] = b;
        bsize[(u64)s * R + d] = c[q];
      }
      b += c[q];
    }
  }
}

template <class K>
__global__ void kd_copy_segs(const K* __restrict__ kin, const u32* __restrict__ vin, K* __restrict__ kout,
                             u32* __restrict__ vout, const u32* __restrict__ llo, const u32* __restrict__ lhi,
                             const u32* __restrict__ tile_base, const u32* __restrict__ nsegs_d,
                             const u32* __restrict__ ntiles_d, u32 tile_rows) {
  __shared__ u32 s_seg;
  const u32 nsegs = *nsegs_d, ntiles = *ntiles_d;
  for (u32 tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    if (threadIdx.x == 0) s_seg = upper_bound_m1(tile_base, nsegs, tile);
    __syncthreads();
    const u32 seg = s_seg;
    const u64 r0 = (u64)llo[seg] + (u64)(tile - tile_base[seg]) * tile_rows;
    const u64 r1 = min((u64)lhi[seg], r0 + tile_rows);
    for (u64 r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
      if (kout) kout[r] = kin[r];
      vout[r] = vin[r];
    }
    __syncthreads();
  }
}

template <class K>
__global__ void kd_copy_pairs(const K* __restrict__ kin, const u32* __restrict__ vin, K* __restrict__ kout,
                              u32* __restrict__ vout, u64 n) {
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) {
    kout[i] = kin[i];
    vout[i] = vin ? vin[i] : (u32)i;  // no vin: the identity order
  }
}

template <class T>
__global__ void kd_copy_col(const T* __restrict__ in, T* __restrict__ out, u64 n) {
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) out[i] = in[i];
}

// ---- profiling helper: snapshot a device counter into a record slot
__global__ void kd_copy1(const u32* __restrict__ src, u32* __restrict__ dst) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *dst = *src;
}

#define FMCS_DD_INST(K)                                                                                            \
  template __global__ void kd_seg_hist<K, ONESWEEP_ITEMS, 8>(const K*, const u32*, const u32*, const u32*,           \
                                              const u32*, const u32*, u32, u32, u32*, u32*, u64, u32*, u32*);      \
  template __global__ void kd_seg_hist<K, WIDE_ITEMS, 8>(const K*, const u32*, const u32*, const u32*,               \
                                              const u32*, const u32*, u32, u32, u32*, u32*, u64, u32*, u32*);      \
  template __global__ void kd_copy_segs<K>(const K*, const u32*, K*, u32*, const u32*, const u32*, const u32*,   \
                                           const u32*, const u32*, u32);                            \
  template __global__ void kd_copy_pairs<K>(const K*, const u32*, K*, u32*, u64);                                  \
  template __global__ void kd_copy_col<K>(const K*, K*, u64);
// the next-digit bytes of the packed round-1 passes
template __global__ void kd_seg_hist<uint8_t, WIDE_ITEMS, 8>(const uint8_t*, const u32*, const u32*, const u32*,
                                                             const u32*, const u32*, u32, u32, u32*, u32*, u64, u32*,
                                                             u32*);
FMCS_DD_INST(u16)
FMCS_DD_INST(u32)
FMCS_DD_INST(u64)
#define FMCS_SCAN_INST(F)                                                                                         \
  template __global__ void kd_scan_excl<8, F>(const u32*, u32*, const u32*, u32*, u32*, u32*, u32, u32*, const u32*, \
                                              u32, u32*, u32*, u32*, u32*);
FMCS_SCAN_INST(false)
FMCS_SCAN_INST(true)
template __global__ void kd_tscan_local<8>(u32*, const u32*, const u32*, const u32*, u32*, u32*, const u32*);
template __global__ void kd_tscan_carry<8>(u32*, const u32*, const u32*);
template __global__ void kd_digit_base<8>(const u32*, const u32*, u32*, u32*, u32*, const u32*, u32*);

}  // namespace fmcs
