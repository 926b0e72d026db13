// fmcs_hash.hip -- hash grouping over rows whose (segment, key) composite
// fits in 64 bits (Hashing::subgroup_range, src/hashing.h:289-385: ids in
// order of first appearance within each segment, numbered across segments).
//
// The composite c = (segment << key bits) | key goes through a bijective
// 64-bit mixer, so two rows have the same hash exactly when they have the
// same composite: no key is ever compared through a row index.  Every
// grouping step below produces, for every position p, the first position
// F[p] of p's group; then p is a group head when F[p] == p, ids are the
// exclusive scan of the heads (k_scan_excl's head mode reads F directly),
// and gid[p] = prefix[F[p]].
//
// F comes from LDS hash tables over chunks of 2048 rows, one of two ways:
//  * local: chunks of consecutive positions.  Each chunk's distinct hashes
//    (with their first position in the chunk) are merged into one global
//    table, then each chunk looks its distinct hashes up there.  Global
//    work is one operation per distinct hash per chunk, so values repeated
//    many times cost little, and F is written in position order.
//  * partitioned, when the chunks hold many distinct hashes: the (hash,
//    position) pairs are radix-partitioned on the hash's low PB bits (about
//    512 rows a partition, stable), and a window of PW partitions, which
//    holds every row of its hashes, is grouped in LDS on its own.  Windows
//    too large for LDS go through the local form's global table.
#include "fmcs_kernels.h"

namespace fmcs {

constexpr u64 HG_EMPTY = ~0ull;
constexpr int HG_E = 8;
constexpr u32 HG_CAP = 256 * HG_E;  // rows of a chunk or window
constexpr u32 HG_TS = 2 * HG_CAP;   // LDS table slots
constexpr u32 HG_NONE = ~0u;

// SplitMix64's finalizer: xor-shifts and odd multiplications, each a
// bijection of the 64-bit words.
__device__ __forceinline__ u64 mix_bijective(u64 x) {
  x ^= x >> 30;
  x *= 0xBF58476D1CE4E5B9ull;
  x ^= x >> 27;
  x *= 0x94D049BB133111EBull;
  x ^= x >> 31;
  return x;
}

template <class K>
__device__ __forceinline__ u64 composite_hash(const K* __restrict__ keys, const u32* __restrict__ seg, u64 i,
                                              u32 kb) {
  u64 c = (u64)keys[i];
  if (seg) c |= (u64)seg[i] << kb;
  return mix_bijective(c);
}

// An LDS table of the distinct hashes of up to HG_CAP rows (E per thread)
// with the first position of each; the one composite whose hash is the
// empty marker keeps its first position in a word of its own.  Slots come
// from the hash's top bits (the low bits are the partition in a window).
struct LdsTable {
  u64* key;
  u32* first;
  u32* marker;
  u32 tmask;
};

// ts: a power of two >= 2 x the rows (so every probe ends).  Contains __syncthreads.
__device__ __forceinline__ void lds_clear(const LdsTable& t, u32 ts) {
  for (u32 i = threadIdx.x; i < ts; i += 256) {
    t.key[i] = HG_EMPTY;
    t.first[i] = HG_NONE;
  }
  if (threadIdx.x == 0) *t.marker = HG_NONE;
  __syncthreads();
}

// slot of a row: its table slot, HG_TS for the marker's rows, HG_NONE for no row
__device__ __forceinline__ u32 lds_insert(const LdsTable& t, u64 h, u32 pos, bool ok) {
  if (!ok) return HG_NONE;
  if (h == HG_EMPTY) {
    atomicMin(t.marker, pos);
    return HG_TS;
  }
  u32 s = (u32)(h >> 40) & t.tmask;
  for (;;) {
    u64 cur = t.key[s];
    if (cur == HG_EMPTY) {
      cur = atomicCAS((unsigned long long*)&t.key[s], (unsigned long long)HG_EMPTY, (unsigned long long)h);
      if (cur == HG_EMPTY) break;
    }
    if (cur == h) break;
    s = (s + 1) & t.tmask;
  }
  // a wave whose active lanes all found one slot (a value repeated many
  // times): its first active lane holds the wave's smallest position (rows
  // ascend with the lane), so it alone updates the slot
  const u64 active = __ballot(true);
  const u32 s0 = (u32)__builtin_amdgcn_readfirstlane((int)s);
  if (__ballot(s == s0) == active) {
    if ((active & lanemask_lt()) == 0) atomicMin(&t.first[s], pos);
  } else {
    atomicMin(&t.first[s], pos);
  }
  return s;
}

__device__ __forceinline__ u32 table_size(u32 rows) { return max(64u, 1u << (32 - __clz(2 * rows - 1))); }

// ---- global table of exact hashes ------------------------------------------

// At most tmask + 1 probes: a full table sets *overflow and drops the row
// (the caller then runs again with a larger table).
__device__ __forceinline__ void global_insert_min(u64* __restrict__ tkeys, u32* __restrict__ tfirst, u64 tmask,
                                                  u64 h, u32 pos, u32* __restrict__ overflow) {
  u64 s = (h >> 20) & tmask;
  for (u64 probes = 0;; ++probes) {
    if (probes > tmask) {
      atomicOr(overflow, 1u);
      return;
    }
    u64 cur = __hip_atomic_load(tkeys + s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (cur == HG_EMPTY) {
      cur = atomicCAS((unsigned long long*)(tkeys + s), (unsigned long long)HG_EMPTY, (unsigned long long)h);
      if (cur == HG_EMPTY) break;
    }
    if (cur == h) break;
    s = (s + 1) & tmask;
  }
  atomicMin(tfirst + s, pos);
}

// Claims h's slot in a table of hashes only (the sample's distinct count);
// 1 when this call claimed it.  At most tmask + 1 probes (the sample table
// has room for twice the sampled rows).
__device__ __forceinline__ u32 global_insert_new(u64* __restrict__ tkeys, u64 tmask, u64 h) {
  u64 s = (h >> 20) & tmask;
  for (u64 probes = 0; probes <= tmask; ++probes, s = (s + 1) & tmask) {
    u64 cur = __hip_atomic_load(tkeys + s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (cur == HG_EMPTY) {
      cur = atomicCAS((unsigned long long*)(tkeys + s), (unsigned long long)HG_EMPTY, (unsigned long long)h);
      if (cur == HG_EMPTY) return 1u;
    }
    if (cur == h) return 0u;
  }
  return 0u;
}

// After every insert has finished (an earlier kernel) with no overflow, the
// hash is there; the probes stay bounded all the same.
__device__ __forceinline__ u32 global_first(const u64* __restrict__ tkeys, const u32* __restrict__ tfirst, u64 tmask,
                                            u64 h) {
  u64 s = (h >> 20) & tmask;
  for (u64 probes = 0; probes <= tmask; ++probes, s = (s + 1) & tmask)
    if (tkeys[s] == h) return tfirst[s];
  return HG_NONE;
}

// ---- local form: chunks of consecutive positions -----------------------------

// Pass 1 (insert set): every chunk's distinct hashes into the global table,
// first positions by atomicMin.  Sampling (insert clear): every stride-th
// chunk; distinct[0] sums the chunks' own distinct hashes (the local form's
// global work, which chooses between the two forms) and distinct[1] counts
// the slots they claim in a table of hashes only (tkeys): the sample's
// distinct count, which sizes the local form's table.
template <class K>
__global__ __launch_bounds__(256) void k_hash_local(const K* __restrict__ keys, const u32* __restrict__ seg, u64 n,
                                                    u32 kb, u64* __restrict__ tkeys, u32* __restrict__ tfirst,
                                                    u64 tmask, u32* __restrict__ marker, u32* __restrict__ distinct,
                                                    u32 insert, u32 stride, u32* __restrict__ overflow) {
  __shared__ u64 s_key[HG_TS];
  __shared__ u32 s_first[HG_TS];
  __shared__ u32 s_marker, s_count;
  const LdsTable t{s_key, s_first, &s_marker, HG_TS - 1};
  const u64 nch = (n + HG_CAP - 1) / HG_CAP;
  for (u64 c = (u64)blockIdx.x * stride; c < nch; c += (u64)gridDim.x * stride) {
    const u64 lo = c * HG_CAP;
    const u32 rows = (u32)min<u64>(HG_CAP, n - lo);
    u64 h[HG_E];  // all of a thread's loads in flight before the first insert
#pragma unroll
    for (int e = 0; e < HG_E; ++e) {
      const u32 r = e * 256 + threadIdx.x;
      h[e] = r < rows ? composite_hash(keys, seg, lo + r, kb) : 0ull;
    }
    lds_clear(t, HG_TS);
    if (threadIdx.x == 0) s_count = 0;
#pragma unroll
    for (int e = 0; e < HG_E; ++e) {
      const u32 r = e * 256 + threadIdx.x;
      lds_insert(t, h[e], (u32)(lo + r), r < rows);
    }
    __syncthreads();
    u32 mine = 0, claimed = 0;
    for (u32 i = threadIdx.x; i < HG_TS; i += 256) {
      if (s_key[i] == HG_EMPTY) continue;
      ++mine;
      if (insert)
        global_insert_min(tkeys, tfirst, tmask, s_key[i], s_first[i], overflow);
      else
        claimed += global_insert_new(tkeys, tmask, s_key[i]);
    }
    if (threadIdx.x == 0 && s_marker != HG_NONE) {
      ++mine;
      ++claimed;  // the marker composite: once per sampled chunk holding it (an estimate)
      if (insert) atomicMin(marker, s_marker);
    }
    if (!insert) {
      // the counts only sample (a few hundred chunks): one counter word
      // taking an atomic per chunk of the whole array would serialise
      if (mine) atomicAdd(&s_count, mine);
      if (claimed) atomicAdd(distinct + 1, claimed);
    }
    __syncthreads();
    if (!insert && threadIdx.x == 0) atomicAdd(distinct, s_count);
    __syncthreads();
  }
}

// Pass 2: F in position order, one global lookup per distinct hash of a chunk.
template <class K>
__global__ __launch_bounds__(256) void k_hash_local_first(const K* __restrict__ keys, const u32* __restrict__ seg,
                                                          u64 n, u32 kb, const u64* __restrict__ tkeys,
                                                          const u32* __restrict__ tfirst, u64 tmask,
                                                          const u32* __restrict__ marker, u32* __restrict__ F) {
  __shared__ u64 s_key[HG_TS];
  __shared__ u32 s_first[HG_TS];
  __shared__ u32 s_marker;
  const LdsTable t{s_key, s_first, &s_marker, HG_TS - 1};
  const u64 nch = (n + HG_CAP - 1) / HG_CAP;
  for (u64 c = blockIdx.x; c < nch; c += gridDim.x) {
    const u64 lo = c * HG_CAP;
    const u32 rows = (u32)min<u64>(HG_CAP, n - lo);
    u64 h[HG_E];
    u32 slot[HG_E];
#pragma unroll
    for (int e = 0; e < HG_E; ++e) {
      const u32 r = e * 256 + threadIdx.x;
      h[e] = r < rows ? composite_hash(keys, seg, lo + r, kb) : 0ull;
    }
    lds_clear(t, HG_TS);
#pragma unroll
    for (int e = 0; e < HG_E; ++e) {
      const u32 r = e * 256 + threadIdx.x;
      slot[e] = lds_insert(t, h[e], (u32)(lo + r), r < rows);
    }
    __syncthreads();
    for (u32 i = threadIdx.x; i < HG_TS; i += 256)
      if (s_key[i] != HG_EMPTY) s_first[i] = global_first(tkeys, tfirst, tmask, s_key[i]);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < HG_E; ++e) {
      const u32 r = e * 256 + threadIdx.x;
      if (slot[e] != HG_NONE) F[lo + r] = slot[e] == HG_TS ? *marker : s_first[slot[e]];
    }
    __syncthreads();
  }
}

// ---- partitioned form -----------------------------------------------------------

template <class K>
__global__ __launch_bounds__(256) void k_hash_prep(const K* __restrict__ keys, const u32* __restrict__ seg, u64 n,
                                                   u32 kb, u64* __restrict__ hk, u32* __restrict__ pos) {
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) {
    hk[i] = composite_hash(keys, seg, i, kb);
    pos[i] = (u32)i;
  }
}

// First sorted index whose partition (the hash's low bits) is at least p.
__device__ __forceinline__ u32 part_start(const u64* __restrict__ hs, u32 n, u32 pmask, u32 p) {
  u32 lo = 0, hi = n;
  while (lo < hi) {
    const u32 mid = (lo + hi) >> 1;
    if ((u32)(hs[mid] & pmask) < p) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// pstart[p] = the first sorted index of partition p, for p in [0, nparts]
// (one binary search a thread, all at once).
__global__ __launch_bounds__(256) void k_hash_pstart(const u64* __restrict__ hs, u32 n, u32 pmask,
                                                     u32* __restrict__ pstart) {
  const u32 nparts = pmask + 1;
  for (u32 p = blockIdx.x * blockDim.x + threadIdx.x; p <= nparts; p += gridDim.x * blockDim.x)
    pstart[p] = p == nparts ? n : part_start(hs, n, pmask, p);
}

// One block per window of pw partitions.  A window of at most HG_CAP rows
// is grouped in LDS; a larger one is listed (lo, hi) for the global table.
__global__ __launch_bounds__(256) void k_hash_window(const u64* __restrict__ hs, const u32* __restrict__ ps,
                                                     const u32* __restrict__ pstart, u32 nparts, u32 pw,
                                                     u32* __restrict__ F, u32* __restrict__ large,
                                                     u32* __restrict__ nlarge) {
  __shared__ u64 s_key[HG_TS];
  __shared__ u32 s_first[HG_TS];
  __shared__ u32 s_marker;
  const u32 p0 = blockIdx.x * pw, p1 = min(p0 + pw, nparts);
  const u32 lo = pstart[p0], hi = pstart[p1];
  if (lo >= hi) return;
  if (hi - lo > HG_CAP) {
    if (threadIdx.x == 0) {
      const u32 k = atomicAdd(nlarge, 1u);
      large[2 * k] = lo;
      large[2 * k + 1] = hi;
    }
    return;
  }
  const u32 ts = table_size(hi - lo);
  const LdsTable t{s_key, s_first, &s_marker, ts - 1};
  u64 h[HG_E];
  u32 pos[HG_E], slot[HG_E];
#pragma unroll
  for (int e = 0; e < HG_E; ++e) {
    const u32 j = lo + e * 256 + threadIdx.x;
    h[e] = j < hi ? hs[j] : 0ull;
    pos[e] = j < hi ? ps[j] : 0u;
  }
  lds_clear(t, ts);
#pragma unroll
  for (int e = 0; e < HG_E; ++e) slot[e] = lds_insert(t, h[e], pos[e], lo + e * 256 + threadIdx.x < hi);
  __syncthreads();
#pragma unroll
  for (int e = 0; e < HG_E; ++e)
    if (slot[e] != HG_NONE) F[pos[e]] = slot[e] == HG_TS ? s_marker : s_first[slot[e]];
}

// The large windows' rows in chunks of HG_CAP sorted positions (clo, chi):
// their distinct hashes into the global table (first positions by atomicMin).
__global__ __launch_bounds__(256) void k_hash_large(const u64* __restrict__ hs, const u32* __restrict__ ps,
                                                    const u32* __restrict__ clo, const u32* __restrict__ chi,
                                                    u32 nchunks, u64* __restrict__ tkeys, u32* __restrict__ tfirst,
                                                    u64 tmask, u32* __restrict__ marker, u32* __restrict__ overflow) {
  __shared__ u64 s_key[HG_TS];
  __shared__ u32 s_first[HG_TS];
  __shared__ u32 s_marker;
  const LdsTable t{s_key, s_first, &s_marker, HG_TS - 1};
  for (u32 c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const u32 lo = clo[c], hi = chi[c];
    u64 h[HG_E];
    u32 pos[HG_E];
#pragma unroll
    for (int e = 0; e < HG_E; ++e) {
      const u32 j = lo + e * 256 + threadIdx.x;
      h[e] = j < hi ? hs[j] : 0ull;
      pos[e] = j < hi ? ps[j] : 0u;
    }
    lds_clear(t, HG_TS);
#pragma unroll
    for (int e = 0; e < HG_E; ++e) lds_insert(t, h[e], pos[e], lo + e * 256 + threadIdx.x < hi);
    __syncthreads();
    for (u32 i = threadIdx.x; i < HG_TS; i += 256)
      if (s_key[i] != HG_EMPTY) global_insert_min(tkeys, tfirst, tmask, s_key[i], s_first[i], overflow);
    if (threadIdx.x == 0 && s_marker != HG_NONE) atomicMin(marker, s_marker);
    __syncthreads();
  }
}

// F of the large windows' rows: one global lookup per distinct hash of a chunk.
__global__ __launch_bounds__(256) void k_hash_large_first(const u64* __restrict__ hs, const u32* __restrict__ ps,
                                                          const u32* __restrict__ clo, const u32* __restrict__ chi,
                                                          u32 nchunks, const u64* __restrict__ tkeys,
                                                          const u32* __restrict__ tfirst, u64 tmask,
                                                          const u32* __restrict__ marker, u32* __restrict__ F) {
  __shared__ u64 s_key[HG_TS];
  __shared__ u32 s_first[HG_TS];
  __shared__ u32 s_marker;
  const LdsTable t{s_key, s_first, &s_marker, HG_TS - 1};
  for (u32 c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const u32 lo = clo[c], hi = chi[c];
    u64 h[HG_E];
    u32 pos[HG_E], slot[HG_E];
#pragma unroll
    for (int e = 0; e < HG_E; ++e) {
      const u32 j = lo + e * 256 + threadIdx.x;
      h[e] = j < hi ? hs[j] : 0ull;
      pos[e] = j < hi ? ps[j] : 0u;
    }
    lds_clear(t, HG_TS);
#pragma unroll
    for (int e = 0; e < HG_E; ++e) slot[e] = lds_insert(t, h[e], pos[e], lo + e * 256 + threadIdx.x < hi);
    __syncthreads();
    for (u32 i = threadIdx.x; i < HG_TS; i += 256)
      if (s_key[i] != HG_EMPTY) s_first[i] = global_first(tkeys, tfirst, tmask, s_key[i]);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < HG_E; ++e)
      if (slot[e] != HG_NONE) F[pos[e]] = slot[e] == HG_TS ? *marker : s_first[slot[e]];
    __syncthreads();
  }
}

// gid[p] = the rank of p's group head among the heads.
__global__ __launch_bounds__(256) void k_hash_assign_f(const u32* __restrict__ F, const u32* __restrict__ prefix,
                                                       u64 n, u32* __restrict__ gid) {
  for (u64 p = (u64)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (u64)gridDim.x * blockDim.x)
    gid[p] = prefix[F[p]];
}

#define FMCS_INST_HASH(K)                                                                                        \
  template __global__ void k_hash_prep<K>(const K*, const u32*, u64, u32, u64*, u32*);                           \
  template __global__ void k_hash_local<K>(const K*, const u32*, u64, u32, u64*, u32*, u64, u32*, u32*, u32, u32,  \
                                           u32*);                                                                \
  template __global__ void k_hash_local_first<K>(const K*, const u32*, u64, u32, const u64*, const u32*, u64,  \
                                                 const u32*, u32*);
FMCS_INST_HASH(uint8_t)
FMCS_INST_HASH(u16)
FMCS_INST_HASH(u32)
FMCS_INST_HASH(u64)

}  // namespace fmcs
