// fmcs_merge.hip -- comparison sort of (key, oid) pairs: LDS bitonic tiles
// followed by merge-path merge passes.
//
// The comparator of the reference's AVX2 merge sort
// (Mergesort::avxmergesort_anytype, src/mergesort.h:2878-3096: in-register
// sorting networks over L2-sized blocks, then a 2-way merge tree).  Pairs
// are ordered by (key, oid); oids are distinct, so the order is total and
// the result equals the stable sort by key.
#include "fmcs_kernels.h"

namespace fmcs {

template <class K>
__device__ __forceinline__ bool pair_less(K ka, u32 oa, K kb, u32 ob) {
  return ka < kb || (ka == kb && oa < ob);
}

// Sorts every MERGE_TILE-row tile in LDS with a bitonic network.  Rows past
// n are padded with the largest (key, oid) and not written back.
template <class K>
__global__ __launch_bounds__(256) void k_bitonic_tile(const K* __restrict__ kin, const u32* __restrict__ vin,
                                                      K* __restrict__ kout, u32* __restrict__ vout, u64 n) {
  constexpr u32 T = MERGE_TILE, NT = 256;
  __shared__ K s_k[T];
  __shared__ u32 s_v[T];
  const u64 base = (u64)blockIdx.x * T;
  for (u32 i = threadIdx.x; i < T; i += NT) {
    const u64 r = base + i;
    s_k[i] = r < n ? kin[r] : (K)~(K)0;
    s_v[i] = r < n ? vin[r] : 0xffffffffu;
  }
  __syncthreads();
  for (u32 k = 2; k <= T; k <<= 1) {
    for (u32 j = k >> 1; j > 0; j >>= 1) {
      for (u32 p = threadIdx.x; p < T / 2; p += NT) {
        const u32 lo = 2 * j * (p / j) + (p % j), hi = lo + j;
        const bool up = (lo & k) == 0;
        const K a = s_k[lo], b = s_k[hi];
        const u32 oa = s_v[lo], ob = s_v[hi];
        if (pair_less(b, ob, a, oa) == up) {
          s_k[lo] = b;
          s_k[hi] = a;
          s_v[lo] = ob;
          s_v[hi] = oa;
        }
      }
      __syncthreads();
    }
  }
  for (u32 i = threadIdx.x; i < T; i += NT) {
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

// One merge pass: sorted runs of `run` rows are merged pairwise.  Block c
// writes outputs [c*MERGE_TILE, +MERGE_TILE): it finds its two merge-path
// splits, stages both input windows in LDS, and every thread merges
// MERGE_ITEMS outputs from its own split of the LDS windows.
template <class K>
__global__ __launch_bounds__(256) void k_merge_pass(const K* __restrict__ kin, const u32* __restrict__ vin,
                                                    K* __restrict__ kout, u32* __restrict__ vout, u64 n, u64 run) {
  constexpr u32 T = MERGE_TILE, NT = 256, E = T / NT;
  __shared__ K s_k[T];
  __shared__ u32 s_v[T];
  __shared__ u32 s_split[2];
  const u64 o0 = (u64)blockIdx.x * T;
  const u64 pair0 = o0 / (2 * run) * (2 * run);
  const u64 a0 = pair0, a1 = min(n, pair0 + run);
  const u64 b0 = a1, b1 = min(n, pair0 + 2 * run);
  const u32 na = (u32)(a1 - a0), nb = (u32)(b1 - b0);
  const u32 d0 = (u32)(o0 - pair0);
  const u32 cnt = (u32)min((u64)T, n - o0);
  if (threadIdx.x < 2) {
    const u32 d = threadIdx.x ? d0 + cnt : d0;
    s_split[threadIdx.x] = merge_split(kin + a0, vin + a0, na, kin + b0, vin + b0, nb, d);
  }
  __syncthreads();
  const u32 ia0 = s_split[0], ia1 = s_split[1];
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
  template __global__ void k_merge_pass<K>(const K*, const u32*, K*, u32*, u64, u64);
FMCS_MERGE_INST(u16)
FMCS_MERGE_INST(u32)
FMCS_MERGE_INST(u64)

}  // namespace fmcs
