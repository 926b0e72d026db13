// fmcs_merge.hip -- comparison sort of (key, oid) pairs: 2048-row tiles
// sorted by bitonic networks in registers (lane exchanges by DPP, swizzle
// and bpermute inside a wave) and merge path across the tile's waves, then
// merge-path merge passes over global memory with their splits found by a
// kernel of their own.
//
// The structure of the reference's AVX2 merge sort
// (Mergesort::avxmergesort_anytype, src/mergesort.h:2878-3096: in-register
// sorting networks, src/mergesort.h:560-903 and src/avxcommon.h:118-502,
// then a 2-way merge tree).  Pairs are ordered by (key, oid); oids are
// distinct, so the order is total and the result equals the stable sort by
// key.
#include "fmcs_kernels.h"

namespace fmcs {

template <class K>
__device__ __forceinline__ bool pair_less(K ka, u32 oa, K kb, u32 ob) {
  return ka < kb || (ka == kb && oa < ob);
}

// ---- block sort: bitonic networks in registers, merge path across waves ----
//
// A thread holds BE consecutive rows of the tile (row x = thread * BE + e),
// so a wave holds 64 * BE consecutive rows.  Every wave sorts its rows with
// a bitonic network: partners less than BE apart are in the thread's own
// registers; partners 8 .. 256 apart are in another lane of the wave and are
// exchanged by DPP lane moves (lane xor 1, 2, 4, 8), ds_swizzle (xor 16) or
// ds_bpermute (xor 32), with no LDS traffic and no barrier.  The waves'
// sorted runs are then merged pairwise in LDS by merge path.
constexpr int BE = 8;                // rows a thread holds
constexpr int WROWS = WAVE * BE;     // rows a wave sorts on its own

// The value of lane (this lane xor M) of the wave.
template <int M>
__device__ __forceinline__ u32 lane_xor(u32 x) {
  if constexpr (M == 1) {
    return dpp_move<0xB1>(x);  // quad_perm [1, 0, 3, 2]
  } else if constexpr (M == 2) {
    return dpp_move<0x4E>(x);  // quad_perm [2, 3, 0, 1]
  } else if constexpr (M == 4) {
    // row_ror:4 gives lane i - 4, row_ror:12 lane i + 4 (within 16 lanes)
    const u32 down = dpp_move<0x124>(x), up = dpp_move<0x12C>(x);
    return (__lane_id() & 4u) ? down : up;
  } else if constexpr (M == 8) {
    return dpp_move<0x128>(x);  // row_ror:8 within 16 lanes is xor 8
  } else if constexpr (M == 16) {
    return (u32)__builtin_amdgcn_ds_swizzle((int)x, 0x401F);  // and 0x1F, xor 0x10 within 32 lanes
  } else {
    return (u32)__shfl_xor((int)x, M);
  }
}

template <int M, class K>
__device__ __forceinline__ K lane_xor_key(K k) {
  if constexpr (sizeof(K) == 8) {
    const u32 lo = lane_xor<M>((u32)k), hi = lane_xor<M>((u32)((u64)k >> 32));
    return (K)(((u64)hi << 32) | lo);
  } else {
    return (K)lane_xor<M>((u32)k);
  }
}

// Keeps the smaller (take_min) or the larger (key, oid) of a and b in a.
template <class K>
__device__ __forceinline__ void keep(K& ka, u32& va, K kb, u32 vb, bool take_min) {
  const bool b_less = pair_less(kb, vb, ka, va);
  if (b_less == take_min) {
    ka = kb;
    va = vb;
  }
}

// One step of the bitonic merge of size S with partner distance J; blocks
// of S rows alternate ascending / descending, except at S == WROWS, where
// every wave sorts ascending.
template <int S, int J, class K>
__device__ __forceinline__ void bitonic_step(K (&k)[BE], u32 (&v)[BE], u32 t) {
  if constexpr (J < BE) {
#pragma unroll
    for (int e = 0; e < BE; ++e) {
      if (e & J) continue;
      const u32 x = t * BE + e;
      const bool asc = S == WROWS || (x & S) == 0;
      K kl = k[e], kh = k[e | J];
      u32 vl = v[e], vh = v[e | J];
      const bool swap = pair_less(kh, vh, kl, vl) == asc;
      k[e] = swap ? kh : kl;
      v[e] = swap ? vh : vl;
      k[e | J] = swap ? kl : kh;
      v[e | J] = swap ? vl : vh;
    }
  } else {
    constexpr int M = J / BE;
#pragma unroll
    for (int e = 0; e < BE; ++e) {
      const u32 x = t * BE + e;
      const bool asc = S == WROWS || (x & S) == 0;
      const bool lower = (x & J) == 0;
      const K pk = lane_xor_key<M>(k[e]);
      const u32 pv = lane_xor<M>(v[e]);
      keep(k[e], v[e], pk, pv, lower == asc);
    }
  }
  if constexpr (J > 1) bitonic_step<S, J / 2>(k, v, t);
}

template <int S, class K>
__device__ __forceinline__ void bitonic_wave_sort(K (&k)[BE], u32 (&v)[BE], u32 t) {
  bitonic_step<S, S / 2>(k, v, t);
  if constexpr (S < WROWS) bitonic_wave_sort<2 * S>(k, v, t);
}

// Number of elements taken from A among the first d outputs of merge(A, B)
// (the merge-path split of diagonal d): the largest i with
// A[i-1] < B[d-i], over i in [max(0, d-nb), min(d, na)].
template <class K>
__device__ __forceinline__ u32 merge_split(const K* ak, const u32* av, u32 na, const K* bk, const u32* bv, u32 nb,
                                           u32 d);

// Sorts every MERGE_TILE-row tile (rows past n padded with the largest
// (key, oid), which no row has: oids are below n).
template <class K>
__global__ __launch_bounds__(256) void k_bitonic_tile(const K* __restrict__ kin, const u32* __restrict__ vin,
                                                      K* __restrict__ kout, u32* __restrict__ vout, u64 n) {
  constexpr u32 T = MERGE_TILE, NT = 256;
  static_assert(T == NT * BE, "a tile is BE rows a thread");
  __shared__ K s_k[T];
  __shared__ u32 s_v[T];
  const u32 t = threadIdx.x;
  const u64 base = (u64)blockIdx.x * T;
  for (u32 i = t; i < T; i += NT) {
    const u64 r = base + i;
    s_k[i] = r < n ? kin[r] : (K)~(K)0;
    s_v[i] = r < n ? vin[r] : 0xffffffffu;
  }
  __syncthreads();
  K k[BE];
  u32 v[BE];
#pragma unroll
  for (int e = 0; e < BE; ++e) {
    k[e] = s_k[t * BE + e];
    v[e] = s_v[t * BE + e];
  }
  bitonic_wave_sort<2>(k, v, t);
  // runs of WROWS rows, sorted; merge them pairwise in LDS
  for (u32 run = WROWS; run < T; run *= 2) {
    __syncthreads();  // every read of the LDS tile above is done
#pragma unroll
    for (int e = 0; e < BE; ++e) {
      s_k[t * BE + e] = k[e];
      s_v[t * BE + e] = v[e];
    }
    __syncthreads();
    const u32 x0 = t * BE, pair0 = x0 / (2 * run) * (2 * run), d = x0 - pair0;
    const K* ak = s_k + pair0;
    const u32* av = s_v + pair0;
    u32 ia = merge_split(ak, av, run, ak + run, av + run, run, d);
    u32 ib = d - ia;
#pragma unroll
    for (int e = 0; e < BE; ++e) {
      const bool has_a = ia < run, has_b = ib < run;
      bool take_a = has_a;
      if (has_a && has_b) take_a = pair_less(ak[ia], av[ia], ak[run + ib], av[run + ib]);
      if (take_a) {
        k[e] = ak[ia];
        v[e] = av[ia];
        ++ia;
      } else {
        k[e] = ak[run + ib];
        v[e] = av[run + ib];
        ++ib;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int e = 0; e < BE; ++e) {
    s_k[t * BE + e] = k[e];
    s_v[t * BE + e] = v[e];
  }
  __syncthreads();
  for (u32 i = t; i < T; i += NT) {
    const u64 r = base + i;
    if (r < n) {
      kout[r] = s_k[i];
      vout[r] = s_v[i];
    }
  }
}

// Number of elements taken from A among the first d outputs of merge(A, B)
// (the merge-path split of diagonal d): the largest i with
// A[i-1] < B[d-i], over i in [max(0, d-nb), min(d, na)].
template <class K>
__device__ __forceinline__ u32 merge_split(const K* ak, const u32* av, u32 na, const K* bk, const u32* bv, u32 nb,
                                           u32 d) {
  u32 lo = d > nb ? d - nb : 0u, hi = min(d, na);
  while (lo < hi) {
    const u32 i = (lo + hi + 1) >> 1;  // candidate: take i from A
    // valid if A[i-1] < B[d-i] (B[d-i] exists since i > d - nb)
    if (pair_less(ak[i - 1], av[i - 1], bk[d - i], bv[d - i]))
      lo = i;
    else
      hi = i - 1;
  }
  return lo;
}

// Merge-path splits of one merge pass, all at once: split[c] = the rows
// taken from run A among the outputs of the pair before block c's first
// output.  One thread a split, so the binary searches over global memory
// overlap instead of each merge block waiting for its own.  (A 64-ary
// search by a wave per split took a quarter of the dependent loads but four
// times the time: its candidates are 64 different lines at every step, where
// the binary searches of neighbouring splits share their first steps' lines.)
template <class K>
__global__ __launch_bounds__(256) void k_merge_splits(const K* __restrict__ kin, const u32* __restrict__ vin, u64 n,
                                                      u64 run, u32 nblocks, u32* __restrict__ split) {
  const u32 c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nblocks) return;
  const u64 o0 = (u64)c * MERGE_TILE;
  const u64 pair0 = o0 / (2 * run) * (2 * run);
  const u64 a0 = pair0, a1 = min(n, pair0 + run);
  const u64 b0 = a1, b1 = min(n, pair0 + 2 * run);
  split[c] = merge_split(kin + a0, vin + a0, (u32)(a1 - a0), kin + b0, vin + b0, (u32)(b1 - b0), (u32)(o0 - pair0));
}

// One merge pass: sorted runs of `run` rows are merged pairwise.  Block c
// writes outputs [c*MERGE_TILE, +MERGE_TILE): its two merge-path splits
// come from k_merge_splits (the end split is the next block's, or all of A
// at the pair's end), it stages both input windows in LDS, and every thread
// merges MERGE_ITEMS outputs from its own split of the LDS windows.
template <class K>
__global__ __launch_bounds__(256) void k_merge_pass(const K* __restrict__ kin, const u32* __restrict__ vin,
                                                    K* __restrict__ kout, u32* __restrict__ vout, u64 n, u64 run,
                                                    const u32* __restrict__ split) {
  constexpr u32 T = MERGE_TILE, NT = 256, E = T / NT;
  __shared__ K s_k[T];
  __shared__ u32 s_v[T];
  const u64 o0 = (u64)blockIdx.x * T;
  const u64 pair0 = o0 / (2 * run) * (2 * run);
  const u64 a0 = pair0, a1 = min(n, pair0 + run);
  const u64 b0 = a1, b1 = min(n, pair0 + 2 * run);
  const u32 na = (u32)(a1 - a0);
  const u32 d0 = (u32)(o0 - pair0);
  const u32 cnt = (u32)min((u64)T, n - o0);
  // the pair's last block ends where the pair ends: all of A taken
  const bool last_of_pair = o0 + T >= b1;
  const u32 ia0 = split[blockIdx.x], ia1 = last_of_pair ? na : split[blockIdx.x + 1];
  const u32 ib0 = d0 - ia0, ib1 = d0 + cnt - ia1;
  const u32 la = ia1 - ia0, lb = ib1 - ib0;  // la + lb == cnt
  for (u32 i = threadIdx.x; i < cnt; i += NT) {
    if (i < la) {
      s_k[i] = kin[a0 + ia0 + i];
      s_v[i] = vin[a0 + ia0 + i];
    } else {
      s_k[i] = kin[b0 + ib0 + (i - la)];
      s_v[i] = vin[b0 + ib0 + (i - la)];
    }
  }
  __syncthreads();
  // this thread's E outputs, merged from the LDS windows [0, la) and [la, cnt)
  const u32 d = min(threadIdx.x * E, cnt);
  u32 ia = merge_split(s_k, s_v, la, s_k + la, s_v + la, lb, d);
  u32 ib = d - ia;
  K rk[E];
  u32 rv[E];
#pragma unroll
  for (u32 e = 0; e < E; ++e) {
    const bool has_a = ia < la, has_b = ib < lb;
    bool take_a = has_a;
    if (has_a && has_b) take_a = pair_less(s_k[ia], s_v[ia], s_k[la + ib], s_v[la + ib]);
    if (take_a) {
      rk[e] = s_k[ia];
      rv[e] = s_v[ia];
      ++ia;
    } else if (has_b) {
      rk[e] = s_k[la + ib];
      rv[e] = s_v[la + ib];
      ++ib;
    }
  }
  __syncthreads();
#pragma unroll
  for (u32 e = 0; e < E; ++e) {
    const u32 o = threadIdx.x * E + e;
    if (o < cnt) {
      s_k[o] = rk[e];
      s_v[o] = rv[e];
    }
  }
  __syncthreads();
  for (u32 i = threadIdx.x; i < cnt; i += NT) {
    kout[o0 + i] = s_k[i];
    vout[o0 + i] = s_v[i];
  }
}

#define FMCS_MERGE_INST(K)                                                                  \
  template __global__ void k_bitonic_tile<K>(const K*, const u32*, K*, u32*, u64);          \
  template __global__ void k_merge_pass<K>(const K*, const u32*, K*, u32*, u64, u64, const u32*);  \
  template __global__ void k_merge_splits<K>(const K*, const u32*, u64, u64, u32, u32*);
FMCS_MERGE_INST(u16)
FMCS_MERGE_INST(u32)
FMCS_MERGE_INST(u64)

}  // namespace fmcs
