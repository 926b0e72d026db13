// fmcs_common.h -- shared device helpers for the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fmcs {

typedef uint16_t u16;
typedef uint32_t u32;
typedef uint64_t u64;

constexpr int WAVE = 64;
constexpr int RADIX_BITS = 8;
constexpr int RADIX = 1 << RADIX_BITS;

// Decoupled look-back status word: 2 flag bits + 30 value bits.
constexpr u32 ST_AGG = 1u << 30;
constexpr u32 ST_INC = 2u << 30;
constexpr u32 ST_VAL = (1u << 30) - 1;

__device__ __forceinline__ u32 status_load(const u32* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void status_store(u32* p, u32 v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Truth table of a | (b ^ c) for v_bitop3_b32 (bit (a << 2 | b << 1 | c) of
// the table is the result for those input bits).
constexpr unsigned BITOP3_OR_XOR = 0xF6;

__device__ __forceinline__ u64 lanemask_lt() {
  u32 lane = __lane_id();
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Inclusive wave scan of u32 through DPP lane moves (no LDS crossbar
// instructions, which a shuffle costs): row_shr 1, 2, 4, 8 inside each
// 16-lane row, then row_bcast:15 (row r's last lane into row r + 1) and
// row_bcast:31 (lane 31 into rows 2 and 3).  A lane whose source lies
// outside its row or below it ignores the moved value.
template <int CTRL>
__device__ __forceinline__ u32 dpp_move(u32 v) {
  return (u32)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xf, 0xf, false);
}
__device__ __forceinline__ u32 wave_incl_scan(u32 v) {
  const u32 lane = __lane_id(), rl = lane & 15u;
  u32 t;
  t = dpp_move<0x111>(v);  // row_shr:1
  if (rl >= 1) v += t;
  t = dpp_move<0x112>(v);  // row_shr:2
  if (rl >= 2) v += t;
  t = dpp_move<0x114>(v);  // row_shr:4
  if (rl >= 4) v += t;
  t = dpp_move<0x118>(v);  // row_shr:8
  if (rl >= 8) v += t;
  t = dpp_move<0x142>(v);  // row_bcast:15
  if ((lane & 31u) >= 16) v += t;
  t = dpp_move<0x143>(v);  // row_bcast:31
  if (lane >= 32) v += t;
  return v;
}

// Decoupled look-back by a whole wave (64 lanes, all active, the same t >
// 0): walks back from tile t (exclusive) over status[0 .. t), each round
// loading the 64 nearest unvisited predecessors (lane l the (l+1)-th
// nearest).  It waits until every word below the nearest inclusive prefix
// holds a flag, then adds the values up to and including that prefix (all
// 64 when there is none, and goes on).  Every predecessor took its tile id
// from an atomic counter before this tile did, so it is resident and
// publishes its aggregate without waiting on anything: the spin always
// ends.  Returns the exclusive prefix in every lane.
__device__ __forceinline__ u32 lookback_wave(const u32* status, u32 t) {
  const u32 lane = __lane_id();
  u32 excl = 0;
  for (u32 p = t; p > 0; p = p > (u32)WAVE ? p - (u32)WAVE : 0u) {
    const bool in = lane < p;  // lanes past tile 0 read as an inclusive 0
    const u32* a = status + (in ? p - 1 - lane : 0u);
    u32 v = in ? status_load(a) : ST_INC;
    u64 inc;
    for (;;) {
      const u64 wait = __ballot((v & ~ST_VAL) == 0);
      inc = __ballot((v & ST_INC) != 0);
      const u64 below = inc ? (inc & (~inc + 1)) - 1 : ~0ull;
      if ((wait & below) == 0) break;
      if ((v & ~ST_VAL) == 0) v = status_load(a);
    }
    // lanes up to and including the nearest inclusive one
    const u64 upto = inc ? ((inc & (~inc + 1)) << 1) - 1 : ~0ull;
    const u32 c = ((upto >> lane) & 1ull) ? (v & ST_VAL) : 0u;
    excl += (u32)__builtin_amdgcn_readlane((int)wave_incl_scan(c), WAVE - 1);
    if (inc) break;
  }
  return excl;
}

// Exclusive scan over a block of NT threads (NT/64 waves), one value each.
// s_tmp needs NT/64 + 1 words.  Returns the exclusive prefix; *total gets
// the block sum.  Contains __syncthreads.
template <int NT>
__device__ __forceinline__ u32 block_excl_scan(u32 v, u32* s_tmp, u32* total) {
  constexpr int NW = NT / WAVE;
  u32 lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
  u32 inc = wave_incl_scan(v);
  if (lane == WAVE - 1) s_tmp[wv] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    u32 run = 0;
    for (int w = 0; w < NW; ++w) {
      u32 c = s_tmp[w];
      s_tmp[w] = run;
      run += c;
    }
    s_tmp[NW] = run;
  }
  __syncthreads();
  u32 r = s_tmp[wv] + inc - v;
  *total = s_tmp[NW];
  __syncthreads();
  return r;
}

// The same scan with one barrier: every thread adds up the wave totals
// before its own wave (NT / WAVE broadcast reads).  s_tmp (NT / WAVE words)
// must not be written again until every thread has passed a later barrier:
// callers with scans in a row alternate two buffers.
template <int NT>
__device__ __forceinline__ u32 block_excl_scan1(u32 v, u32* s_tmp, u32* total) {
  constexpr int NW = NT / WAVE;
  const u32 lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
  const u32 inc = wave_incl_scan(v);
  if (lane == WAVE - 1) s_tmp[wv] = inc;
  __syncthreads();
  u32 pre = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    const u32 c = s_tmp[w];
    pre += (u32)w < wv ? c : 0u;
    tot += c;
  }
  *total = tot;
  return pre + inc - v;
}

template <class K>
__device__ __forceinline__ u32 digit_of(K k, u32 shift, u32 mask) {
  return (u32)((u64)k >> shift) & mask;
}

// Keys of one 16-byte vector.
template <class K>
__device__ __forceinline__ void unpack16(const uint4& x, K (&k)[16 / sizeof(K)]) {
  const u32 w[4] = {x.x, x.y, x.z, x.w};
  if constexpr (sizeof(K) == 8) {
    k[0] = (K)w[0] | ((K)w[1] << 32);
    k[1] = (K)w[2] | ((K)w[3] << 32);
  } else if constexpr (sizeof(K) == 4) {
#pragma unroll
    for (int i = 0; i < 4; ++i) k[i] = (K)w[i];
  } else if constexpr (sizeof(K) == 1) {
#pragma unroll
    for (int i = 0; i < 16; ++i) k[i] = (K)(w[i / 4] >> (8 * (i % 4)));
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      k[2 * i] = (K)(w[i] & 0xffffu);
      k[2 * i + 1] = (K)(w[i] >> 16);
    }
  }
}

// Largest i in [0, n) with a[i] <= x (a sorted ascending, a[0] <= x).
__device__ __forceinline__ u32 upper_bound_m1(const u32* a, u32 n, u32 x) {
  u32 lo = 0, hi = n;
  while (hi - lo > 1) {
    u32 mid = (lo + hi) >> 1;
    if (a[mid] <= x) lo = mid; else hi = mid;
  }
  return lo;
}

}  // namespace fmcs
