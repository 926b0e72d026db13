// fmcs_kernels.hip -- gfx950 kernels of the multi-column sorting path.
//
// Every kernel here is integer, bandwidth-oriented work: no MFMA.  Keys live
// in 16/32/64-bit banks (structure-of-arrays), oids are u32.  The kernels are
// wave64-native: ranking uses 64-bit ballots, scans use wave shuffles.
#include "fmcs_kernels.h"
#include "fmcs_tile.h"

namespace fmcs {

// ---------------------------------------------------------------------------
// a3: fused stitch / bit-borrow.  One read of each source bank and one write
// of each destination bank per row (src/column.h:105-1059 does one
// read-modify-write pass per entry plus a shift pass per entry).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void load4(const void* p, u32 bank, u64 i, bool full, u64 n, u64 v[4]) {
  if (full) {
    switch (bank) {
      case 1: {
        u32 x = *(const u32*)((const uint8_t*)p + i);
        for (int j = 0; j < 4; ++j) v[j] = (x >> (8 * j)) & 0xff;
        break;
      }
      case 2: {
        uint2 x = *(const uint2*)((const u16*)p + i);
        v[0] = x.x & 0xffff; v[1] = x.x >> 16; v[2] = x.y & 0xffff; v[3] = x.y >> 16;
        break;
      }
      case 4: {
        uint4 x = *(const uint4*)((const u32*)p + i);
        v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
        break;
      }
      default: {
        const ulonglong2* q = (const ulonglong2*)((const u64*)p + i);
        ulonglong2 a = q[0], b = q[1];
        v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
        break;
      }
    }
  } else {
    for (int j = 0; j < 4; ++j) {
      u64 r = i + j;
      if (r >= n) { v[j] = 0; continue; }
      switch (bank) {
        case 1: v[j] = ((const uint8_t*)p)[r]; break;
        case 2: v[j] = ((const u16*)p)[r]; break;
        case 4: v[j] = ((const u32*)p)[r]; break;
        default: v[j] = ((const u64*)p)[r]; break;
      }
    }
  }
}

__device__ __forceinline__ void store4(void* p, u32 bank, u64 i, bool full, u64 n, const u64 v[4]) {
  if (full) {
    switch (bank) {
      case 1: {
        u32 x = (u32)(v[0] & 0xff) | ((u32)(v[1] & 0xff) << 8) | ((u32)(v[2] & 0xff) << 16) | ((u32)(v[3] & 0xff) << 24);
        *(u32*)((uint8_t*)p + i) = x;
        break;
      }
      case 2: {
        uint2 x;
        x.x = (u32)(v[0] & 0xffff) | ((u32)(v[1] & 0xffff) << 16);
        x.y = (u32)(v[2] & 0xffff) | ((u32)(v[3] & 0xffff) << 16);
        *(uint2*)((u16*)p + i) = x;
        break;
      }
      case 4: {
        uint4 x = make_uint4((u32)v[0], (u32)v[1], (u32)v[2], (u32)v[3]);
        *(uint4*)((u32*)p + i) = x;
        break;
      }
      default: {
        ulonglong2* q = (ulonglong2*)((u64*)p + i);
        q[0] = make_ulonglong2(v[0], v[1]);
        q[1] = make_ulonglong2(v[2], v[3]);
        break;
      }
    }
  } else {
    for (int j = 0; j < 4; ++j) {
      u64 r = i + j;
      if (r >= n) continue;
      switch (bank) {
        case 1: ((uint8_t*)p)[r] = (uint8_t)v[j]; break;
        case 2: ((u16*)p)[r] = (u16)v[j]; break;
        case 4: ((u32*)p)[r] = (u32)v[j]; break;
        default: ((u64*)p)[r] = v[j]; break;
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_stitch(StitchArgs a) {
  u64 base = ((u64)blockIdx.x * blockDim.x + threadIdx.x) * 4;
  if (base >= a.n) return;
  bool full = base + 4 <= a.n;
  for (u32 c = 0; c < a.ndst; ++c) {
    u64 acc[4] = {0, 0, 0, 0};
    for (u32 e = a.ent_begin[c]; e < a.ent_begin[c + 1]; ++e) {
      u32 s = a.e_src[e], w = a.e_width[e], sh = a.e_shift[e];
      u64 v[4];
      load4(a.src[s], a.src_bank[s], base, full, a.n, v);
      u64 m = w >= 64 ? ~0ull : ((1ull << w) - 1);
      // a descending source: its bits inverted within their width, so the
      // ascending sort of the new column orders it from the largest value
      const u64 flip = ((a.invert >> s) & 1u) ? m : 0ull;
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[j] = (w >= 64 ? 0 : (acc[j] << w)) | (((v[j] >> sh) & m) ^ flip);
    }
    store4(a.dst[c], a.dst_bank[c], base, full, a.n, acc);
  }
}

// ---------------------------------------------------------------------------
// Row records of several payload columns for the gather by one order
// (BATproject of each column, src/composer.cpp:608-628).  Column loops run
// over the uniform descriptor; records are assembled with static word
// indices only (no register array indexed at run time).  The gather writes
// four consecutive rows a thread: column pointers are 16-byte aligned
// (checked by the host), so a column's four values leave as one 8-, 16- or
// 32-byte vector.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void col_store4(void* p, u32 wd, u64 i0, u32 m, const u64 (&v)[4]) {
  if (m == 4) {
    if (wd == 2) {
      *(uint2*)((u16*)p + i0) = make_uint2((u32)v[0] | (u32)v[1] << 16, (u32)v[2] | (u32)v[3] << 16);
    } else if (wd == 4) {
      *(uint4*)((u32*)p + i0) = make_uint4((u32)v[0], (u32)v[1], (u32)v[2], (u32)v[3]);
    } else {
      uint4* q = (uint4*)((u64*)p + i0);
      q[0] = make_uint4((u32)v[0], (u32)(v[0] >> 32), (u32)v[1], (u32)(v[1] >> 32));
      q[1] = make_uint4((u32)v[2], (u32)(v[2] >> 32), (u32)v[3], (u32)(v[3] >> 32));
    }
  } else {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if ((u32)r >= m) continue;
      if (wd == 2) ((u16*)p)[i0 + r] = (u16)v[r];
      else if (wd == 4) ((u32*)p)[i0 + r] = (u32)v[r];
      else ((u64*)p)[i0 + r] = v[r];
    }
  }
}

// Rows b*1024 + wave*256 + r*64 + lane (r < 4): every column load and
// record store of a wave covers 64 consecutive rows.
template <int R>
__global__ __launch_bounds__(256) void k_rec_pack(RecCols c, void* __restrict__ rec, u64 n) {
  constexpr int NW = R / 4;
  const u64 base = (u64)blockIdx.x * 1024 + (threadIdx.x >> 6) * 256 + (threadIdx.x & 63);
  u32 w[4][NW];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int j = 0; j < NW; ++j) w[r][j] = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if ((u32)k >= c.ncol) break;
    const u32 wd = c.width[k], q = c.off[k] >> 2, sh = (c.off[k] & 2) * 8;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const u64 row = base + (u64)r * 64;
      if (row >= n) continue;
      const u64 v = wd == 2 ? (u64)((const u16*)c.src[k])[row]
                            : wd == 4 ? (u64)((const u32*)c.src[k])[row] : ((const u64*)c.src[k])[row];
#pragma unroll
      for (int j = 0; j < NW; ++j) {
        if ((u32)j == q) w[r][j] |= (u32)v << sh;
        if (wd == 8 && (u32)j == q + 1) w[r][j] |= (u32)(v >> 32);
      }
    }
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const u64 row = base + (u64)r * 64;
    if (row >= n) continue;
    if constexpr (R == 16)
      ((uint4*)rec)[row] = make_uint4(w[r][0], w[r][1], w[r][2], w[r][3]);
    else
      ((uint2*)rec)[row] = make_uint2(w[r][0], w[r][1]);
  }
}

// outs[k][i] = cols[k][order[i]]: one random record load a row (REC_ROWS
// consecutive rows a thread, all their loads in flight), then every
// column's values as vector stores of four rows.
template <int R>
__global__ __launch_bounds__(256) void k_rec_gather(RecCols c, const void* __restrict__ rec,
                                                    const u32* __restrict__ order, u64 n) {
  constexpr int NW = R / 4, RT = REC_ROWS;
  const u64 i0 = ((u64)blockIdx.x * blockDim.x + threadIdx.x) * RT;
  if (i0 >= n) return;
  const u32 m = (u32)(n - i0 < RT ? n - i0 : RT);
  u32 o[RT];
  if (m == RT) {
#pragma unroll
    for (int g = 0; g < RT / 4; ++g) {
      const uint4 t = *(const uint4*)(order + i0 + 4 * g);
      o[4 * g] = t.x; o[4 * g + 1] = t.y; o[4 * g + 2] = t.z; o[4 * g + 3] = t.w;
    }
  } else {
#pragma unroll
    for (int r = 0; r < RT; ++r) o[r] = order[i0 + ((u32)r < m ? r : 0)];
  }
  u32 w[RT][NW];
#pragma unroll
  for (int r = 0; r < RT; ++r) {
    if constexpr (R == 16) {
      const uint4 t = ((const uint4*)rec)[o[r]];
      w[r][0] = t.x; w[r][1] = t.y; w[r][2] = t.z; w[r][3] = t.w;
    } else {
      const uint2 t = ((const uint2*)rec)[o[r]];
      w[r][0] = t.x; w[r][1] = t.y;
    }
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if ((u32)k >= c.ncol) break;
    const u32 wd = c.width[k], q = c.off[k] >> 2, sh = (c.off[k] & 2) * 8;
    const u32 mask = wd == 2 ? 0xffffu : 0xffffffffu;
#pragma unroll
    for (int g = 0; g < RT / 4; ++g) {
      if ((u32)(4 * g) >= m) break;
      u64 v[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        v[r] = 0;
#pragma unroll
        for (int j = 0; j < NW; ++j) {
          if ((u32)j == q) v[r] |= (w[4 * g + r][j] >> sh) & mask;
          if (wd == 8 && (u32)j == q + 1) v[r] |= (u64)w[4 * g + r][j] << 32;
        }
      }
      col_store4(c.dst[k], wd, i0 + 4 * g, m - 4 * g < 4 ? m - 4 * g : 4, v);
    }
  }
}

// Bit records (declared widths summing to at most 64 bits): fields at bit
// offsets of one u64.  The pack takes rows b*2048 + wave*512 + r*64 + lane
// (r < 8: eight loads of a column in flight a thread); the gather maps rows
// as the byte-record gather does.
__global__ __launch_bounds__(256) void k_bitrec_pack(RecCols c, u64* __restrict__ rec, u64 n, u32* __restrict__ err) {
  const u64 base = (u64)blockIdx.x * 2048 + (threadIdx.x >> 6) * 512 + (threadIdx.x & 63);
  u64 w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  bool bad = false;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if ((u32)k >= c.ncol) break;
    const u32 wd = c.width[k], b = c.bits[k], off = c.off[k];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const u64 row = base + (u64)r * 64;
      if (row >= n) continue;
      const u64 v = wd == 2 ? (u64)((const u16*)c.src[k])[row]
                            : wd == 4 ? (u64)((const u32*)c.src[k])[row] : ((const u64*)c.src[k])[row];
      bad = bad || (b < 64 && (v >> b) != 0);
      w[r] |= v << off;
    }
  }
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const u64 row = base + (u64)r * 64;
    if (row < n) rec[row] = w[r];
  }
  if (bad) atomicOr(err, 1u);
}

__global__ __launch_bounds__(256) void k_bitrec_gather(RecCols c, const u64* __restrict__ rec,
                                                       const u32* __restrict__ order, u64 n) {
  const u64 i0 = ((u64)blockIdx.x * blockDim.x + threadIdx.x) * 4;
  if (i0 >= n) return;
  const u32 m = (u32)(n - i0 < 4 ? n - i0 : 4);
  u32 o[4];
  if (m == 4) {
    const uint4 t = *(const uint4*)(order + i0);
    o[0] = t.x; o[1] = t.y; o[2] = t.z; o[3] = t.w;
  } else {
#pragma unroll
    for (int r = 0; r < 4; ++r) o[r] = order[i0 + ((u32)r < m ? r : 0)];
  }
  u64 w[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) w[r] = rec[o[r]];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if ((u32)k >= c.ncol) break;
    const u32 b = c.bits[k], off = c.off[k];
    const u64 mask = b >= 64 ? ~0ull : ((1ull << b) - 1);
    u64 v[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = (w[r] >> off) & mask;
    col_store4(c.dst[k], c.width[k], i0, m, v);
  }
}

// ---------------------------------------------------------------------------
// a6: oid-permutation gather, out[i] = vals[order[i]] (src/composer.cpp:608-628).
// Four rows per thread: one 16-byte order load, four random loads, one
// vector store.
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_gather(const T* __restrict__ vals, const u32* __restrict__ order,
                                                T* __restrict__ out, u64 n) {
  u64 i = ((u64)blockIdx.x * blockDim.x + threadIdx.x) * 4;
  if (i + 4 <= n) {
    uint4 o = *(const uint4*)(order + i);
    T r0 = vals[o.x], r1 = vals[o.y], r2 = vals[o.z], r3 = vals[o.w];
    out[i] = r0; out[i + 1] = r1; out[i + 2] = r2; out[i + 3] = r3;
  } else {
    for (; i < n; ++i) out[i] = vals[order[i]];
  }
}

// ---------------------------------------------------------------------------
// extractGroupInfo (src/composer.cpp:630-678): dense group ids and group
// start offsets (offsets[ngroups] = n) of a sorted column, in two passes
// with no inter-tile waiting.  Row r is a head when
// r == 0 or its value >> shift (or incoming group id) differs from row r - 1
// (shift > 0: the column is the high part of a composite key).
// k_bcount counts the heads of every tile; after an exclusive scan of the
// counts, k_bwrite recomputes the heads and writes ids / offsets.  When
// hbits is set, k_bcount also writes each thread's 32 head bits there (bit
// v * PV + i for row base + (v * 256 + tid) * PV + i) and k_bwrite reads
// them instead of the rows.
// ---------------------------------------------------------------------------
// Row r of a tile is read as element i of lane l's 16-byte vector v:
// r = base + (v * 256 + tid) * PV + i, PV = 16 / sizeof(K) keys per vector.
// hm[v] gets the head bits of the lane's PV rows.  The predecessor of a
// lane's first row is the next lane down's last row; lane 0 loads it.
template <class K, int ITEMS, bool HASG>
__device__ __forceinline__ void tile_heads_g(const K* __restrict__ v, const u32* __restrict__ in_group, u64 n,
                                           u32 shift, u64 base, bool vec_ok, u32 (&hm)[ITEMS * sizeof(K) / 16]) {
  constexpr int PV = 16 / sizeof(K), NV = ITEMS / PV;
  const u32 tid = threadIdx.x, lane = tid & (WAVE - 1);
  K k[NV][PV];
  u32 g[NV][PV];
  // a whole aligned tile: straight-line vector loads (no per-vector branch)
  const bool full = vec_ok && base + (u64)ITEMS * 256 <= n;
  // the wave's first thread: lane 0's predecessor row is wave-uniform, a
  // scalar load (its wait is not the vector loads' in-order one)
  const u32 wtid = __builtin_amdgcn_readfirstlane(tid);
  // one vector of the keys (and the group ids) at element e0
  auto vec_load = [&](int vv, u64 e0) {
    unpack16<K>(*(const uint4*)(v + e0), k[vv]);
    if constexpr (HASG) {
      if constexpr (PV >= 4) {
#pragma unroll
        for (int i = 0; i < PV; i += 4) {
          const uint4 x = *(const uint4*)(in_group + e0 + i);
          g[vv][i] = x.x;
          g[vv][i + 1] = x.y;
          g[vv][i + 2] = x.z;
          g[vv][i + 3] = x.w;
        }
      } else {
        const uint2 x = *(const uint2*)(in_group + e0);
        g[vv][0] = x.x;
        g[vv][1] = x.y;
      }
    }
  };
  if (full) {
#pragma unroll
    for (int vv = 0; vv < NV; ++vv) vec_load(vv, base + ((u64)vv * 256 + tid) * PV);
  } else {
#pragma unroll
    for (int vv = 0; vv < NV; ++vv) {
      const u64 e0 = base + ((u64)vv * 256 + tid) * PV;
      if (vec_ok && e0 + PV <= n) {
        vec_load(vv, e0);
      } else {
#pragma unroll
        for (int i = 0; i < PV; ++i) {
          k[vv][i] = e0 + i < n ? v[e0 + i] : (K)0;
          if constexpr (HASG) g[vv][i] = e0 + i < n ? in_group[e0 + i] : 0u;
        }
      }
    }
  }
  if constexpr (!HASG) {
#pragma unroll
    for (int vv = 0; vv < NV; ++vv)
#pragma unroll
      for (int i = 0; i < PV; ++i) g[vv][i] = 0;
  }
#pragma unroll
  for (int vv = 0; vv < NV; ++vv) {
    const u64 e0 = base + ((u64)vv * 256 + tid) * PV;
    u64 pv = (u64)__shfl_up((unsigned long long)k[vv][PV - 1], 1);
    u32 pg = __shfl_up(g[vv][PV - 1], 1);
    const u64 ew = base + ((u64)vv * 256 + wtid) * PV;  // lane 0's first row
    const u64 ep = ew > 0 && ew < n ? ew - 1 : 0;
    const u64 pvw = (u64)v[ep];
    u32 pgw = 0;
    if constexpr (HASG) pgw = in_group[ep];
    if (lane == 0 && e0 > 0 && e0 < n) {
      pv = pvw;
      pg = pgw;
    }
    u32 m = 0;
#pragma unroll
    for (int i = 0; i < PV; ++i) {
      const u64 pk = i ? (u64)k[vv][i - 1] : pv;
      const u32 pgi = i ? g[vv][i - 1] : pg;
      const bool head = (e0 + i < n) && (e0 + i == 0 || ((u64)k[vv][i] >> shift) != (pk >> shift) || g[vv][i] != pgi);
      m |= (head ? 1u : 0u) << i;
    }
    hm[vv] = m;
  }
}

// the group ids' presence as a template flag: no branch between the loads
template <class K, int ITEMS>
__device__ __forceinline__ void tile_heads(const K* __restrict__ v, const u32* __restrict__ in_group, u64 n,
                                           u32 shift, u64 base, bool vec_ok, u32 (&hm)[ITEMS * sizeof(K) / 16]) {
  if (in_group)
    tile_heads_g<K, ITEMS, true>(v, in_group, n, shift, base, vec_ok, hm);
  else
    tile_heads_g<K, ITEMS, false>(v, in_group, n, shift, base, vec_ok, hm);
}

template <class K, int ITEMS>
__global__ __launch_bounds__(256) void k_bcount(const K* __restrict__ v, const u32* __restrict__ in_group, u64 n,
                                                u32 shift, u32* __restrict__ tcnt, u32* __restrict__ hbits) {
  constexpr int NT = 256, NW = NT / WAVE, TILE = NT * ITEMS, NV = ITEMS * sizeof(K) / 16, PV = 16 / sizeof(K);
  static_assert(NV * PV == 32, "32 head bits a thread");
  __shared__ u32 s_w[NW];
  const u32 tid = threadIdx.x;
  const bool vec_ok = ((uintptr_t)v & 15) == 0 && ((uintptr_t)in_group & 15) == 0;
  u32 hm[NV];
  tile_heads<K, ITEMS>(v, in_group, n, shift, (u64)blockIdx.x * TILE, vec_ok, hm);
  if (hbits) {
    u32 m = 0;
#pragma unroll
    for (int vv = 0; vv < NV; ++vv) m |= hm[vv] << (vv * PV);
    hbits[(u64)blockIdx.x * NT + tid] = m;
  }
  u32 c = 0;
#pragma unroll
  for (int vv = 0; vv < NV; ++vv) c += (u32)__popc(hm[vv]);
  c = wave_incl_scan(c);  // lane 63 holds the wave's total
  if ((tid & (WAVE - 1)) == WAVE - 1) s_w[tid / WAVE] = c;
  __syncthreads();
  if (tid == 0) {
    u32 t = 0;
    for (int w = 0; w < NW; ++w) t += s_w[w];
    tcnt[blockIdx.x] = t;
  }
}

template <class K, int ITEMS>
__global__ __launch_bounds__(256) void k_bwrite(const K* __restrict__ v, const u32* __restrict__ in_group, u64 n,
                                                u32 shift, const u32* __restrict__ tpre, u32* __restrict__ out_group,
                                                u32* __restrict__ seg_off, u32* __restrict__ ngroups_out,
                                                const u32* __restrict__ hbits) {
  constexpr int NT = 256, NW = NT / WAVE, TILE = NT * ITEMS, PV = 16 / sizeof(K), NV = ITEMS / PV;
  __shared__ u32 s_cnt[NV][NW];
  const u32 tid = threadIdx.x, wv = tid / WAVE, lane = tid & (WAVE - 1);
  const u32 tile = blockIdx.x;
  const u64 base = (u64)tile * TILE;
  const bool vec_ok = ((uintptr_t)v & 15) == 0 && ((uintptr_t)in_group & 15) == 0;
  u32 hm[NV], lex[NV];
  if (hbits) {
    const u32 m = hbits[(u64)tile * NT + tid];
#pragma unroll
    for (int vv = 0; vv < NV; ++vv) hm[vv] = (m >> (vv * PV)) & ((1u << PV) - 1u);
  } else {
    tile_heads<K, ITEMS>(v, in_group, n, shift, base, vec_ok, hm);
  }
#pragma unroll
  for (int vv = 0; vv < NV; ++vv) {
    const u32 c = (u32)__popc(hm[vv]);
    const u32 inc = wave_incl_scan(c);
    lex[vv] = inc - c;
    if (lane == WAVE - 1) s_cnt[vv][wv] = inc;
  }
  __syncthreads();
  u32 run = tpre[tile];
  const bool ovec = out_group && ((uintptr_t)out_group & 15) == 0;
#pragma unroll
  for (int vv = 0; vv < NV; ++vv) {
    u32 before_w = 0, tot_v = 0;
    for (int w = 0; w < NW; ++w) {
      const u32 c = s_cnt[vv][w];
      before_w += (w < (int)wv) ? c : 0;
      tot_v += c;
    }
    const u64 e0 = base + ((u64)vv * 256 + tid) * PV;
    u32 p = run + before_w + lex[vv];  // heads strictly before row e0
    u32 ids[PV];
#pragma unroll
    for (int i = 0; i < PV; ++i) {
      const bool head = (hm[vv] >> i) & 1u;
      if (head && seg_off) seg_off[p] = (u32)(e0 + i);
      ids[i] = head ? p : p - 1u;
      p += head ? 1u : 0u;
    }
    if (out_group) {
      if (ovec && e0 + PV <= n) {
        if constexpr (PV >= 4) {
#pragma unroll
          for (int i = 0; i < PV; i += 4) *(uint4*)(out_group + e0 + i) = make_uint4(ids[i], ids[i + 1], ids[i + 2], ids[i + 3]);
        } else {
          *(uint2*)(out_group + e0) = make_uint2(ids[0], ids[1]);
        }
      } else {
#pragma unroll
        for (int i = 0; i < PV; ++i)
          if (e0 + i < n) out_group[e0 + i] = ids[i];
      }
    }
    run += tot_v;
  }
  if (base + TILE >= n && tid == 0) {
    *ngroups_out = run;
    if (seg_off) seg_off[run] = (u32)n;
  }
}

// Windowed gather: out[i] = vals[order[i]] only for the rows whose oid lies
// in [wlo, whi).  Passes over consecutive windows of the source read each
// window while it stays in the memory-side cache: a window's random reads
// hit there instead of fetching a full DRAM burst per 8-byte value, and the
// rows of one large group (ascending oids) write contiguous runs per pass.
template <class T>
__global__ __launch_bounds__(256) void k_gather_win(const T* __restrict__ vals, const u32* __restrict__ order,
                                                    T* __restrict__ out, u64 n, u32 wlo, u32 wlen) {
  u64 i = ((u64)blockIdx.x * blockDim.x + threadIdx.x) * 4;
  if (i + 4 <= n) {
    const uint4 o = *(const uint4*)(order + i);
    if (o.x - wlo < wlen) out[i] = vals[o.x];
    if (o.y - wlo < wlen) out[i + 1] = vals[o.y];
    if (o.z - wlo < wlen) out[i + 2] = vals[o.z];
    if (o.w - wlo < wlen) out[i + 3] = vals[o.w];
  } else {
    for (; i < n; ++i) {
      const u32 o = order[i];
      if (o - wlo < wlen) out[i] = vals[o];
    }
  }
}

// ---------------------------------------------------------------------------
// Radix histograms.
// k_hist_all: one pass over all keys computing the digit histogram of every
// LSD pass at once (npasses <= 8, digits split by digit_width/digit_shift).
// A key with a bit in bad sets hist[npasses * RADIX].
// (kd_seg_hist, fmcs_dd.hip: per-segment and per-tile histograms of one
// digit over tiles that never cross a segment.)
// ---------------------------------------------------------------------------
template <class K>
__global__ __launch_bounds__(256) void k_hist_all(const K* __restrict__ keys, u64 n, u32 npasses, u32 keybits,
                                                  u64 bad, u32* __restrict__ hist) {
  __shared__ u32 sh[8][RADIX];
  for (u32 i = threadIdx.x; i < 8 * RADIX; i += blockDim.x) (&sh[0][0])[i] = 0;
  __syncthreads();
  // digit_width / digit_shift unrolled: the first `narrow` digits are wb
  // bits wide, the rest wb + 1 (no per-pass arrays, which would live in
  // scratch memory under a runtime loop)
  const u32 wb = keybits / npasses, narrow = npasses - keybits % npasses;
  u64 over = 0;
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) {
    u64 k = (u64)keys[i];
    over |= k & bad;
    u32 s = 0;
    for (u32 p = 0; p < npasses; ++p) {
      const u32 w = wb + (p >= narrow ? 1u : 0u);
      atomicAdd(&sh[p][(u32)(k >> s) & ((1u << w) - 1)], 1u);
      s += w;
    }
  }
  if (__ballot(over != 0) && __lane_id() == 0) atomicOr(&hist[npasses * RADIX], 1u);
  __syncthreads();
  for (u32 p = 0; p < npasses; ++p) {
    u32 c = sh[p][threadIdx.x];
    if (c) atomicAdd(&hist[p * RADIX + threadIdx.x], c);
  }
}

__global__ __launch_bounds__(256) void k_digit_base(const u32* __restrict__ hist, const u32* __restrict__ lo,
                                                    u32* __restrict__ digit_base, u32* __restrict__ bstart,
                                                    u32* __restrict__ bsize) {
  __shared__ u32 tmp[8];
  u32 s = blockIdx.x, d = threadIdx.x;
  u32 c = hist[(u64)s * RADIX + d];
  u32 tot;
  u32 ex = block_excl_scan<256>(c, tmp, &tot);
  u32 b = (lo ? lo[s] : 0u) + ex;
  digit_base[(u64)s * RADIX + d] = b;
  if (bstart) {
    bstart[(u64)s * RADIX + d] = b;
    bsize[(u64)s * RADIX + d] = c;
  }
}

// ---------------------------------------------------------------------------
// Scatter of one radix pass (LSD over the whole array, or one MSD level over
// a list of segments).  The per-tile digit counts were written by the
// histogram kernel and scanned per segment (kd_tscan_local/_carry), so a
// tile's global offset of digit d is
//   digit_base[seg][d] + tile_pre[tile][d] + (carry_in[chunk][d] if the
//   segment started before the tile's scan chunk),
// known before the tile starts: no tile waits on another.  A tile ranks its
// keys stably with 64-lane ballot matching into per-wave LDS histograms and
// scatters through LDS so the global writes are contiguous runs.
// Tile layout: wave w owns rows [row0 + w*64*ITEMS, +64*ITEMS); item j of
// lane l is row wave_base + j*64 + l, so the ranking order (j, lane) is the
// row order and the pass is stable.
// ---------------------------------------------------------------------------
// A payload column P (pin -> pout, none when P is NoPay) moves with the
// pairs: each row's payload lands where its (key, value) lands.  It goes
// through the same LDS reorder after the pairs are written, in the LDS that
// held them.  A payload with a bit in pbad (above its declared width) sets
// *err.  FUSE: the key and the payload leave as one u64 composite
// (key << fshift) | payload in pout, and no keys are written to kout.
// Without a payload, kside (if set) also gets every key >> fshift as a u32:
// the high part of composite keys, for a boundary pass over that part alone.
// LSD: one segment [seg_lo[0], seg_hi[0]) tiled from tile 0 (no search for
// a tile's segment).  Else tile_seg (if set) holds every tile's segment;
// without it the segment is found by search over tile_base.  KO: the type
// the keys leave in (narrower than K: the low bits of every key, which an
// MSD level leaves to sort when they fit KO).
template <class K, int ITEMS, int RB, class P, bool FUSE, bool LSD, class KO>
__global__ __launch_bounds__(256) void k_onesweep(const K* __restrict__ kin, const u32* __restrict__ vin,
                                                  KO* __restrict__ kout, u32* __restrict__ vout,
                                                  const u32* __restrict__ seg_lo, const u32* __restrict__ seg_hi,
                                                  const u32* __restrict__ tile_base, u32 nsegs,
                                                  const u32* __restrict__ digit_base, const u32* __restrict__ tile_pre,
                                                  const u32* __restrict__ carry_in, u32 shift, u32 dbits,
                                                  const u32* __restrict__ nsegs_d, const u32* __restrict__ ntiles_d,
                                                  const P* __restrict__ pin, PayOut<P, FUSE>* __restrict__ pout,
                                                  u32 fshift, u64 pbad, u32* __restrict__ err,
                                                  u32* __restrict__ kside, uint8_t* __restrict__ dnext,
                                                  u32 nshift, const u32* __restrict__ tile_seg) {
  using PO = PayOut<P, FUSE>;
  // RB-bit digits (R of them); every thread owns DPT consecutive digits
  constexpr int NT = 256, NW = NT / WAVE, TILE = NT * ITEMS, WITEMS = WAVE * ITEMS;
  constexpr int R = 1 << RB, DPT = R / NT;
  constexpr bool HASP = !std::is_same<P, NoPay>::value;
  constexpr int KVB = TILE * (int)(sizeof(K) + sizeof(u32)), PB = HASP ? TILE * (int)sizeof(PO) : 0;
  __shared__ u32 s_whist[NW][R];
  __shared__ u32 s_goff[R];
  __shared__ __align__(16) unsigned char s_raw[KVB > PB ? KVB : PB];
  K* s_keys = (K*)s_raw;
  u32* s_vals = (u32*)(s_raw + TILE * sizeof(K));
  // the block scan's scratch and the tile's segment: the staging area is
  // idle until the scan's closing barrier
  u32* s_tmp = (u32*)s_raw;
  u32& s_seg = *((u32*)s_raw + 8);

  u32 tid = threadIdx.x, wv = tid / WAVE, lane = tid & (WAVE - 1);
  if (nsegs_d) nsegs = *nsegs_d;
  const u32 ntiles = ntiles_d ? *ntiles_d : tile_base[nsegs];
  // Persistent and XCD-aware: workgroups are dealt to the XCDs round-robin
  // (block b runs on XCD b % XCDS), so XCD x takes the x-th contiguous
  // eighth of the tiles and its blocks walk it together.  Neighbouring tiles
  // write neighbouring parts of every digit's output run, so the partial
  // lines at run ends meet in one XCD's L2 instead of going out to memory
  // from several.  The grid is a multiple of XCDS.
  const u32 xcd = blockIdx.x % XCDS, per_x = gridDim.x / XCDS;
  const u32 t_hi = (u32)((u64)ntiles * (xcd + 1) / XCDS);
  for (u32 tile = (u32)((u64)ntiles * xcd / XCDS) + blockIdx.x / XCDS; tile < t_hi; tile += per_x) {
  const u32 tseg = (!LSD && tile_seg) ? tile_seg[tile] : 0u;
  for (u32 i = tid; i < NW * R; i += NT) (&s_whist[0][0])[i] = 0;
  if (!LSD && !tile_seg && tid == 0) s_seg = upper_bound_m1(tile_base, nsegs, tile);
  __syncthreads();
  const u32 seg = LSD ? 0u : (tile_seg ? tseg : s_seg);
  const u32 first_tile = tile_base[seg];
  const u64 row0 = (u64)seg_lo[seg] + (u64)(tile - first_tile) * TILE;
  const u64 rowend = min((u64)seg_hi[seg], row0 + TILE);
  const u32 mask = (1u << dbits) - 1;
  const u64 wbase = row0 + (u64)wv * WITEMS;
  const u64 lt = lanemask_lt();

  // Loads in the order they are needed: the keys (for the ranking), this
  // tile's global digit offsets (after the tile scan), then the oids (for
  // the staging).  Memory operations complete in order, so the ranking
  // waits for the keys only while the rest is in flight; every load is
  // unconditional (a load under a branch makes the compiler wait for all).
  // Rows past the tile's end load the last row instead (rowend > row0): no
  // branch around the loads; those rows are never ranked or written.
  const u64 rlast = rowend - 1;
  K keys[ITEMS];
  u32 vals[ITEMS];
  u32 rank[ITEMS];
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) keys[j] = kin[min(wbase + (u64)j * WAVE + lane, rlast)];
  const u32 chunk = tile / TSCAN_CT;
  const bool carry = first_tile < chunk * TSCAN_CT;
  u32 gbase[DPT];
#pragma unroll
  for (int q = 0; q < DPT; ++q) {
    const u32 d = tid * DPT + q;
    const u32 c = carry_in[(u64)chunk * R + d];
    gbase[q] = digit_base[(u64)seg * R + d] + tile_pre[(u64)tile * R + d] + (carry ? c : 0u);
  }
  if (vin) {
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) vals[j] = vin[min(wbase + (u64)j * WAVE + lane, rlast)];
  } else {
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) vals[j] = (u32)(wbase + (u64)j * WAVE + lane);  // the identity order
  }
  // Stable ranking: lanes with equal digits are matched by ballots; the
  // lowest such lane adds the match count to the wave's LDS counter with one
  // returning atomic.  The atomics of all items are issued back to back (LDS
  // executes a wave's operations in order, so item j sees the counts of
  // items < j) and their results are collected afterwards.
  // rank[j] first holds (leader lane << 8) | matches in lower lanes
  u32 old[ITEMS];
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    u64 r = wbase + (u64)j * WAVE + lane;
    bool ok = r < rowend;
    u32 dg = digit_of(keys[j], shift, mask);
    // lanes whose digit differs from this lane's in any bit: per bit, the
    // ballot xor the lane's own bit spread over the word (0 or all ones),
    // or-ed into the two halves of the mask with one 3-input op each.  All RB
    // bits, unrolled: a digit's bits above dbits are zero in every lane.
    u32 dlo = 0, dhi = 0;
#pragma unroll
    for (int b = 0; b < RB; ++b) {
      const u32 own = (u32)__builtin_amdgcn_sbfe((int)dg, b, 1);
      const u64 bb = __ballot(own != 0);
      dlo = __builtin_amdgcn_bitop3_b32(dlo, (u32)bb, own, BITOP3_OR_XOR);
      dhi = __builtin_amdgcn_bitop3_b32(dhi, (u32)(bb >> 32), own, BITOP3_OR_XOR);
    }
    u64 peers = ok ? __ballot(ok) & ~(((u64)dhi << 32) | dlo) : 0ull;
    const u32 lead = ok ? (u32)(__ffsll((long long)peers) - 1) : lane;
    rank[j] = (lead << 8) | (u32)__popcll(peers & lt);
    old[j] = 0;
    if (ok && lane == lead) old[j] = atomicAdd(&s_whist[wv][dg], (u32)__popcll(peers));
  }
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) rank[j] = __shfl(old[j], (int)(rank[j] >> 8)) + (rank[j] & 0xffu);
  __syncthreads();
  // per digit: the waves' counts and the tile's count; after the block scan
  // of the tile counts, every wave's LDS offset of the digit (the digit's
  // offset in the tile folded in)
  u32 wcnt[DPT][NW];
  u32 tcount[DPT];
  u32 tsum = 0;
#pragma unroll
  for (int q = 0; q < DPT; ++q) {
    const u32 d = tid * DPT + q;
    u32 run = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      wcnt[q][w] = s_whist[w][d];
      run += wcnt[q][w];
    }
    tcount[q] = run;
    tsum += run;
  }
  u32 ttot;
  u32 tdig = block_excl_scan<NT>(tsum, s_tmp, &ttot);
#pragma unroll
  for (int q = 0; q < DPT; ++q) {
    const u32 d = tid * DPT + q;
    u32 run = tdig;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      s_whist[w][d] = run;
      run += wcnt[q][w];
    }
    s_goff[d] = gbase[q] - tdig;
    tdig += tcount[q];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    u64 r = wbase + (u64)j * WAVE + lane;
    if (r < rowend) {
      u32 dd = digit_of(keys[j], shift, mask);
      u32 pos = s_whist[wv][dd] + rank[j];
      s_keys[pos] = keys[j];
      s_vals[pos] = vals[j];
      if (HASP) rank[j] = pos;
    }
  }
  __syncthreads();
  const u32 cnt = (u32)(rowend - row0);
  if constexpr (HASP) {
    // pairs out, keeping each LDS row's global position; then the payload
    // rows, loaded in row order, through the same LDS positions
    // payload loads without branches (rows past the end load the last
    // row), and a full tile's pair stores without per-row conditions: the
    // wait for the payload then does not also wait for those stores
    P pv[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) pv[j] = pin[min(wbase + (u64)j * WAVE + lane, rlast)];
    u32 gpos[ITEMS];
    if (cnt == (u32)TILE) {
#pragma unroll
      for (int it = 0; it < ITEMS; ++it) {
        const u32 i = it * NT + tid;
        const K k = s_keys[i];
        const u32 g = s_goff[digit_of(k, shift, mask)] + i;
        if (!FUSE) kout[g] = (KO)k;
        vout[g] = s_vals[i];
        if (dnext) dnext[g] = (uint8_t)((u64)k >> nshift);
        gpos[it] = g;
      }
    } else {
#pragma unroll
      for (int it = 0; it < ITEMS; ++it) {
        const u32 i = it * NT + tid;
        if (i < cnt) {
          const K k = s_keys[i];
          const u32 g = s_goff[digit_of(k, shift, mask)] + i;
          if (!FUSE) kout[g] = (KO)k;
          vout[g] = s_vals[i];
          if (dnext) dnext[g] = (uint8_t)((u64)k >> nshift);
          gpos[it] = g;
        }
      }
    }
    __syncthreads();
    PO* s_pay = (PO*)s_raw;
    u64 bad = 0;
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      u64 r = wbase + (u64)j * WAVE + lane;
      if (r < rowend) {
        bad |= (u64)pv[j] & pbad;
        if constexpr (FUSE)
          s_pay[rank[j]] = ((u64)keys[j] << fshift) | (u64)pv[j];
        else
          s_pay[rank[j]] = pv[j];
      }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < ITEMS; ++it) {
      const u32 i = it * NT + tid;
      if (i < cnt) pout[gpos[it]] = s_pay[i];
    }
    if (err && __ballot(bad != 0) && __lane_id() == 0) atomicOr(err, 1u);
  } else if (cnt == (u32)TILE) {
    // full tile: independent iterations, so the LDS reads of all rows
    // overlap
#pragma unroll
    for (int it = 0; it < ITEMS; ++it) {
      const u32 i = it * NT + tid;
      const K k = s_keys[i];
      const u32 g = s_goff[digit_of(k, shift, mask)] + i;
      kout[g] = (KO)k;
      vout[g] = s_vals[i];
      if (kside) kside[g] = (u32)((u64)k >> fshift);
      if (dnext) dnext[g] = (uint8_t)((u64)k >> nshift);
    }
  } else {
    for (u32 i = tid; i < cnt; i += NT) {
      K k = s_keys[i];
      u32 g = s_goff[digit_of(k, shift, mask)] + i;
      kout[g] = (KO)k;
      vout[g] = s_vals[i];
      if (kside) kside[g] = (u32)((u64)k >> fshift);
      if (dnext) dnext[g] = (uint8_t)((u64)k >> nshift);
    }
  }
  __syncthreads();
  }
}

// The window a block of the radix tile sort or the rank takes: w of the
// launch's arrays, or the whole range d_lo .. d_hi as one segment (direct).
template <class K>
__device__ __forceinline__ bool window_args(const TileSortArgs<K>& a, u32 w, TileWin& win) {
  if (a.direct) {
    win = TileWin{a.d_lo, a.d_hi, 1u, 0u};
    return true;
  }
  if (a.only && a.only[w] == 0) return false;
  win = TileWin{a.tile_lo[w], a.tile_hi[w], a.tile_s0[w], a.tile_s1[w]};
  return true;
}

template <class K, int E>
__device__ __forceinline__ void tile_sort_one(const TileSortArgs<K>& a, u32 w, TileSortLds<K, E>& L) {
  TileWin win;
  if (!window_args(a, w, win)) return;
  const u32 sel = a.direct ? 0u : a.src_sel;
  tile_sort_win<K, E>(a.kin[sel], a.vin[sel], a.kout, a.vout, a.seg_start, a.kb, win, L, a.rklist, a.rkcount, w);
}

// One tile per block, or (a.fbcount set) a persistent loop over the tiles
// the bucket sort refused: fblist[0 .. *fbcount), or (no fblist) the first
// *fbcount windows of the launch's arrays.
template <class K, int E>
__global__ __launch_bounds__(256) void k_tile_sort(TileSortArgs<K> a) {
  __shared__ TileSortLds<K, E> L;
  if (a.fbcount) {
    const u32 cnt = *a.fbcount;
    for (u32 b = blockIdx.x; b < cnt; b += gridDim.x) {
      tile_sort_one<K, E>(a, a.fblist ? a.fblist[b] : b, L);
      __syncthreads();
    }
  } else {
    tile_sort_one<K, E>(a, a.wlist ? a.wlist[blockIdx.x] : blockIdx.x, L);
  }
}

template <class K, int E>
__device__ __forceinline__ void tile_rank_one(const TileSortArgs<K>& a, u32 w, RankLds<K, E>& L) {
  TileWin win;
  if (!window_args(a, w, win)) return;
  tile_rank_win<K, E>(a.kin[a.src_sel], a.vin[a.src_sel], a.kout, a.vout, a.seg_start, win, L);
}

template <class K, int E>
__global__ __launch_bounds__(256) void k_tile_rank(TileSortArgs<K> a) {
  __shared__ RankLds<K, E> L;
  if (a.fbcount) {
    const u32 cnt = *a.fbcount;
    for (u32 b = blockIdx.x; b < cnt; b += gridDim.x) {
      tile_rank_one<K, E>(a, a.fblist ? a.fblist[b] : b, L);
      __syncthreads();
    }
  } else {
    tile_rank_one<K, E>(a, a.wlist ? a.wlist[blockIdx.x] : blockIdx.x, L);
  }
}

// The descriptor arrays come in as restrict kernel parameters (they are the
// same as a.tile_lo ... a.tile_s1): the compiler then knows that the
// kernel's stores leave them alone and reads them with scalar loads, whose
// waits do not hold the vector loads up.  A window outside the launch's
// size band (size_lo, size_hi] is another launch's and is skipped.
template <class K, int E>
__device__ __forceinline__ void load_window(const TileSortArgs<K>& a, u32 w, const u32* __restrict__ tlo,
                                            const u32* __restrict__ thi, const u32* __restrict__ ts0,
                                            const u32* __restrict__ ts1, WinRows<K, E>& r) {
  r.w = w;
  const u32 lo = tlo[w];
  u32 hi = thi[w];
  const u32 cnt = hi > lo ? hi - lo : 0u;
  if (a.size_hi && (cnt <= a.size_lo || cnt > a.size_hi)) hi = lo;
  load_rows<K, E>(a.kin[a.src_sel], a.vin[a.src_sel], a.seg_start, lo, hi, ts0[w], ts1[w], r);
}

template <class K, int E>
__device__ __forceinline__ void tile_bucket_one(const TileSortArgs<K>& a, const WinRows<K, E>& r,
                                                u32* __restrict__ fallback, BucketLds<K, E>& L) {
  if (tile_bucket_body<K, E>(r, a.seg_start, a.kb, a.kout, a.vout, a.bad, a.err, L) && threadIdx.x == 0) {
    fallback[r.w] = 1;
    if (a.fblist) a.fblist[atomicAdd(a.fbcount, 1u)] = r.w;
  }
}

// One window per block, or (a.wcount set) a persistent loop over the
// compact window list, which lists windows roughly in row order: die x
// (blocks b with b % XCDS == x) takes the x-th contiguous part, so the
// shared lines at window edges are written from one L2.  The persistent
// grid is a multiple of XCDS.  The kernel is bound by the latency of a
// window's chain of barrier phases, so what matters is how many windows a
// CU holds at once: one register set of rows (no prefetch of the next
// window) and half as many bins as rows keep it at 62 VGPRs and 24.6 KB of
// LDS for 1024-row tiles, 6 blocks a CU.
template <class K, int E>
__global__ __launch_bounds__(256) void k_tile_bucket(TileSortArgs<K> a, u32* __restrict__ fallback,
                                                     const u32* __restrict__ tlo, const u32* __restrict__ thi,
                                                     const u32* __restrict__ ts0, const u32* __restrict__ ts1,
                                                     const u32* __restrict__ wl) {
  __shared__ BucketLds<K, E> L;
  if (a.wcount) {
    const u32 cnt = *a.wcount;
    const u32 xcd = blockIdx.x % XCDS, per_x = gridDim.x / XCDS;
    const u32 hi = (u32)((u64)cnt * (xcd + 1) / XCDS);
    for (u32 b = (u32)((u64)cnt * xcd / XCDS) + blockIdx.x / XCDS; b < hi; b += per_x) {
      WinRows<K, E> r;
      load_window<K, E>(a, wl ? wl[b] : b, tlo, thi, ts0, ts1, r);
      tile_bucket_one<K, E>(a, r, fallback, L);
      __syncthreads();
    }
  } else {
    WinRows<K, E> r;
    load_window<K, E>(a, wl ? wl[blockIdx.x] : blockIdx.x, tlo, thi, ts0, ts1, r);
    tile_bucket_one<K, E>(a, r, fallback, L);
  }
}

// Window packing of small segments (size <= small_max) of a sorted,
// disjoint segment list: segment i belongs to window start[i] / W.  A large
// segment (>= W rows) that starts in a window runs past its end, so the
// small segments of a window are contiguous and precede it.  Lists of MSD
// buckets (RADIX per parent segment, parents not adjacent in general) use
// windows relative to their parent: parent_wbase[p] * wpt + (start - lo[p]) / W,
// so a window never spans the gap between two parents.  An empty segment
// is placed at start - 1 (never before its parent's first row): the empty
// buckets after a parent's last row would otherwise take the window id one
// past the parent's range (the next parent's first window, when the parent
// fills its tiles exactly) and stretch that window back over the gap.
__device__ __forceinline__ u32 window_pos(const u32* st, const u32* sz, u32 i, u32 lo) {
  const u32 s = st[i];
  return (sz[i] == 0 && s > lo) ? s - 1 : s;
}
__device__ __forceinline__ u32 window_of(const u32* st, const u32* sz, u32 i, u32 W, const u32* plo, const u32* pwb,
                                         u32 wpt) {
  if (!plo) return window_pos(st, sz, i, 0u) / W;
  u32 p = i >> RADIX_BITS;
  return pwb[p] * wpt + (window_pos(st, sz, i, plo[p]) - plo[p]) / W;
}

__global__ void k_windows(const u32* __restrict__ st, const u32* __restrict__ sz, u32 nseg, u32 W, u32 small_max,
                          const u32* __restrict__ plo, const u32* __restrict__ pwb, u32 wpt,
                          u32* __restrict__ tlo, u32* __restrict__ thi, u32* __restrict__ ts0,
                          u32* __restrict__ ts1, u32* __restrict__ wlist, u32* __restrict__ wcount) {
  u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  bool first = false;
  u32 w = 0;
  if (i < nseg) {
    u32 size = sz[i];
    if (size <= small_max) {
      w = window_of(st, sz, i, W, plo, pwb, wpt);
      first = (i == 0) || (window_of(st, sz, i - 1, W, plo, pwb, wpt) != w);
      bool last = (i + 1 == nseg) || (window_of(st, sz, i + 1, W, plo, pwb, wpt) != w) || (sz[i + 1] > small_max);
      if (first) { tlo[w] = st[i]; ts0[w] = i; }
      if (last) { thi[w] = st[i] + size; ts1[w] = i; }
    }
  }
  // append the windows that hold rows to a compact list (one atomic per wave)
  u64 m = __ballot(first);
  if (m) {
    u32 lane = __lane_id();
    u32 leader = (u32)(__ffsll((long long)m) - 1);
    u32 base = 0;
    if (lane == leader) base = atomicAdd(wcount, (u32)__popcll(m));
    base = __shfl(base, (int)leader);
    if (first) wlist[base + (u32)__popcll(m & lanemask_lt())] = w;
  }
}

// Segment list from offsets: start[i] = off[i], size[i] = off[i+1]-off[i].
__global__ void k_offsets_to_list(const u32* __restrict__ off, u32 nseg, u32* __restrict__ st, u32* __restrict__ sz) {
  u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nseg) return;
  u32 a = off[i], b = off[i + 1];
  st[i] = a;
  sz[i] = b - a;
}

// Large segments: flag, then (after a scan) compact to (lo, hi) and count
// their tiles.
__global__ void k_large_flags(const u32* __restrict__ sz, u32 nseg, u32 small_max, u32* __restrict__ flag,
                              u32* __restrict__ rows) {
  u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  u32 s = i < nseg ? sz[i] : 0u;
  bool large = s > small_max;
  if (i < nseg) flag[i] = large ? 1u : 0u;
  u32 a = large ? 0u : s, b = large ? s : 0u;
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) {
    a += __shfl_down(a, o);
    b += __shfl_down(b, o);
  }
  // one pair of global atomics per block
  __shared__ u32 s_ab[2];
  if (threadIdx.x == 0) s_ab[0] = s_ab[1] = 0;
  __syncthreads();
  if ((threadIdx.x & (WAVE - 1)) == 0) {
    if (a) atomicAdd(&s_ab[0], a);
    if (b) atomicAdd(&s_ab[1], b);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (s_ab[0]) atomicAdd(rows, s_ab[0]);
    if (s_ab[1]) atomicAdd(rows + 1, s_ab[1]);
  }
}

__global__ void k_large_compact(const u32* __restrict__ st, const u32* __restrict__ sz, u32 nseg, u32 small_max,
                                const u32* __restrict__ pos, u32 tile_rows, u32* __restrict__ llo,
                                u32* __restrict__ lhi, u32* __restrict__ ltiles) {
  u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nseg) return;
  u32 s = sz[i];
  if (s <= small_max) return;
  u32 p = pos[i];
  llo[p] = st[i];
  lhi[p] = st[i] + s;
  ltiles[p] = (s + tile_rows - 1) / tile_rows;
}

// Copy rows of segments that need no sorting (all key bits consumed).
template <class K>
__global__ void k_copy_segs(const K* __restrict__ kin, const u32* __restrict__ vin, K* __restrict__ kout,
                            u32* __restrict__ vout, const u32* __restrict__ llo, const u32* __restrict__ lhi,
                            const u32* __restrict__ tile_base, u32 nsegs, u32 tile_rows) {
  __shared__ u32 s_seg;
  if (threadIdx.x == 0) s_seg = upper_bound_m1(tile_base, nsegs, blockIdx.x);
  __syncthreads();
  u32 seg = s_seg;
  u64 r0 = (u64)llo[seg] + (u64)(blockIdx.x - tile_base[seg]) * tile_rows;
  u64 r1 = min((u64)lhi[seg], r0 + tile_rows);
  for (u64 r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
    if (kout) kout[r] = kin[r];
    vout[r] = vin[r];
  }
}

// ---------------------------------------------------------------------------
// Device scan (exclusive, u32) in one sweep with decoupled look-back; total
// written to *total.  Values must sum below 2^30.  A tile of 256 * ITEMS
// values is ITEMS / 4 slices of 1024; a thread holds 4 consecutive values of
// each slice, read and written as one 16-byte vector (lane-contiguous, so
// every load instruction covers 1 KiB without gaps), and each slice is
// scanned across the block.  HEADS scans the head flags of a first-position
// array instead of the input values: value i is 1 when in[i] == i.
// ---------------------------------------------------------------------------
template <int ITEMS, bool HEADS>
__global__ __launch_bounds__(256) void k_scan_excl(const u32* __restrict__ in, u32* __restrict__ out, u64 n,
                                                   u32* __restrict__ status, u32* __restrict__ counter,
                                                   u32* __restrict__ total) {
  constexpr int NT = 256, NW = NT / WAVE, V = ITEMS / 4, SLICE = NT * 4, TILE = NT * ITEMS;
  static_assert(ITEMS % 4 == 0, "whole 16-byte vectors per thread");
  __shared__ u32 s_tile, s_excl;
  __shared__ u32 s_wave[V][NW];
  const u32 tid = threadIdx.x, lane = tid & (WAVE - 1), wv = tid / WAVE;
  if (tid == 0) s_tile = atomicAdd(counter, 1u);
  __syncthreads();
  const u32 tile = s_tile;
  const u64 t0 = (u64)tile * TILE;
  const bool vec = ((((uintptr_t)in | (uintptr_t)out) & 15) == 0) && t0 + TILE <= n;
  u32 v[V][4], incl[V];
#pragma unroll
  for (int s = 0; s < V; ++s) {
    const u64 base = t0 + (u64)s * SLICE + tid * 4;
    if (vec) {
      const uint4 q = *reinterpret_cast<const uint4*>(in + base);
      v[s][0] = q.x; v[s][1] = q.y; v[s][2] = q.z; v[s][3] = q.w;
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) v[s][j] = base + j < n ? in[base + j] : 0u;
    }
    u32 sum = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (HEADS) v[s][j] = (base + j < n && v[s][j] == (u32)(base + j)) ? 1u : 0u;
      sum += v[s][j];
    }
    incl[s] = wave_incl_scan(sum) - sum;  // exclusive within the wave from here on
    if (lane == WAVE - 1) s_wave[s][wv] = incl[s] + sum;
  }
  __syncthreads();
  // slice offsets inside the tile and the tile's total
  u32 tot = 0;
#pragma unroll
  for (int s = 0; s < V; ++s) {
    u32 before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      const u32 c = s_wave[s][w];
      before += (u32)w < wv ? c : 0u;
      all += c;
    }
    incl[s] += tot + before;
    tot += all;
  }
  if (tid < WAVE) {
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
      if (t0 + TILE >= n && total) *total = e + tot;
    }
  }
  __syncthreads();
  const u32 e = s_excl;
#pragma unroll
  for (int s = 0; s < V; ++s) {
    const u64 base = t0 + (u64)s * SLICE + tid * 4;
    u32 o[4];
    u32 run = e + incl[s];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      o[j] = run;
      run += v[s][j];
    }
    if (vec) {
      *reinterpret_cast<uint4*>(out + base) = make_uint4(o[0], o[1], o[2], o[3]);
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (base + j < n) out[base + j] = o[j];
    }
  }
}

// ---------------------------------------------------------------------------
// Reduce-then-scan for long arrays: no block waits on another and no tile id
// comes from a global atomic (one counter serialises its atomics, about
// 11 ns each on gfx950, which caps a one-pass scan of 2^25 values at 16384
// tiles near 200 us).  Block b owns rows [b*C, min(n, (b+1)*C)), C a
// multiple of 1024; a thread reads 4 consecutive values per 1024-row step
// (one 16-byte vector when the buffers are 16-byte aligned).  HEADS: value i
// is 1 when in[i] == i (the head flags of a first-position array).
// ---------------------------------------------------------------------------
template <bool HEADS>
__device__ __forceinline__ void scan_quad(const u32* __restrict__ in, u64 i, u64 end, bool vec, u32 v[4]) {
  if (vec && i + 4 <= end) {
    const uint4 q = *reinterpret_cast<const uint4*>(in + i);
    v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
  } else {
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = i + j < end ? in[i + j] : 0u;
  }
  if (HEADS) {
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = (i + j < end && v[j] == (u32)(i + j)) ? 1u : 0u;
  }
}

template <bool HEADS>
__global__ __launch_bounds__(256) void k_scan_reduce(const u32* __restrict__ in, u64 n, u64 C, u32* __restrict__ sums) {
  const u64 b0 = (u64)blockIdx.x * C, end = min(n, b0 + C);
  const bool vec = (((uintptr_t)in) & 15) == 0;
  u32 acc = 0;
  for (u64 i = b0 + threadIdx.x * 4; i < end; i += 1024) {
    u32 v[4];
    scan_quad<HEADS>(in, i, end, vec, v);
    acc += v[0] + v[1] + v[2] + v[3];
  }
  for (int o = WAVE / 2; o > 0; o >>= 1) acc += __shfl_down(acc, o);
  __shared__ u32 s[256 / WAVE];
  if ((threadIdx.x & (WAVE - 1)) == 0) s[threadIdx.x / WAVE] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    u32 t = 0;
    for (int w = 0; w < 256 / WAVE; ++w) t += s[w];
    sums[blockIdx.x] = t;
  }
}

// Exclusive scan of G <= 4096 block sums in place (one block of 256), the
// grand total to *total if set.
__global__ __launch_bounds__(256) void k_scan_sums(u32* __restrict__ sums, u32 G, u32* __restrict__ total) {
  __shared__ u32 s_tmp[256 / WAVE + 1];
  u32 v[16];
  u32 sum = 0;
  const u32 b = threadIdx.x * 16;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    v[j] = b + j < G ? sums[b + j] : 0u;
    sum += v[j];
  }
  u32 tot;
  u32 run = block_excl_scan<256>(sum, s_tmp, &tot);
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    if (b + j < G) sums[b + j] = run;
    run += v[j];
  }
  if (threadIdx.x == 0 && total) *total = tot;
}

template <bool HEADS>
__global__ __launch_bounds__(256) void k_scan_rescan(const u32* __restrict__ in, u32* __restrict__ out, u64 n, u64 C,
                                                     const u32* __restrict__ offs) {
  constexpr int NW = 256 / WAVE;
  __shared__ u32 s_w[2][NW];
  const u32 tid = threadIdx.x, lane = tid & (WAVE - 1), wv = tid / WAVE;
  const u64 b0 = (u64)blockIdx.x * C, end = min(n, b0 + C);
  const bool vec = ((((uintptr_t)in) | ((uintptr_t)out)) & 15) == 0;
  u32 run = offs[blockIdx.x];
  int par = 0;
  // every thread runs the same number of steps (C is a multiple of 1024);
  // the wave totals alternate between two LDS rows, so one barrier a step
  for (u64 i0 = b0; i0 < end; i0 += 1024, par ^= 1) {
    const u64 i = i0 + tid * 4;
    u32 v[4];
    scan_quad<HEADS>(in, i, end, vec, v);
    const u32 sum = v[0] + v[1] + v[2] + v[3];
    const u32 ex = wave_incl_scan(sum) - sum;
    if (lane == WAVE - 1) s_w[par][wv] = ex + sum;
    __syncthreads();
    u32 before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      const u32 c = s_w[par][w];
      before += (u32)w < wv ? c : 0u;
      all += c;
    }
    u32 o[4];
    o[0] = run + before + ex;
    o[1] = o[0] + v[0];
    o[2] = o[1] + v[1];
    o[3] = o[2] + v[2];
    if (vec && i + 4 <= end) {
      *reinterpret_cast<uint4*>(out + i) = make_uint4(o[0], o[1], o[2], o[3]);
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (i + j < end) out[i + j] = o[j];
    }
    run += all;
  }
}

__global__ void k_fill_u32(u32* __restrict__ p, u64 n, u32 v) {
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_iota_u32(u32* __restrict__ p, u64 n) {
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) p[i] = (u32)i;
}

// Composite key for the single unbinned segmented sort: (group << kb) |
// key.  A key with bits above kb sets *err (and is masked, so its row stays
// in its group).
template <class K>
__global__ void k_composite(const K* __restrict__ k, const u32* __restrict__ g, u64 n, u32 kb, u64* __restrict__ out,
                            u32* __restrict__ err) {
  const u64 kmask = kb >= 64 ? ~0ull : ((1ull << kb) - 1);
  u64 bad = 0;
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) {
    const u64 key = (u64)k[i];
    bad |= key & ~kmask;
    out[i] = ((u64)g[i] << kb) | (key & kmask);
  }
  if (__ballot(bad != 0) && __lane_id() == 0) atomicOr(err, 1u);
}

// *err |= 1 if any key has a bit in bad_mask.
template <class K>
__global__ void k_key_check(const K* __restrict__ k, u64 n, u64 bad_mask, u32* __restrict__ err) {
  u64 bad = 0;
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x)
    bad |= (u64)k[i] & bad_mask;
  if (__ballot(bad != 0) && __lane_id() == 0) atomicOr(err, 1u);
}

template <class K>
__global__ void k_strip_composite(const u64* __restrict__ c, u64 n, u64 kmask, K* __restrict__ k) {
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x)
    k[i] = (K)(c[i] & kmask);
}

// ---------------------------------------------------------------------------
// a10-a12: hash grouping.  One open-addressing table over (segment, key)
// for all segments at once; a slot holds the row index of its first
// inserter (one 32-bit CAS; keys are compared through the row index, so the
// table stores no keys).  The group's id is fixed by the smallest row
// index that maps to it (atomicMin), so ids follow first appearance
// (src/hashing.h:289-385), whatever order the inserts raced in.
// ---------------------------------------------------------------------------
constexpr u32 HASH_EMPTY = 0xffffffffu;

__device__ __forceinline__ u64 mix64(u64 x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return x;
}

// table has cmask+1 >= 2n slots, so a free or matching slot always exists
// and every probe sequence ends.  The first (smallest) row of every slot is
// kept per block in a small LDS table and added to `first` with one global
// atomicMin per block and slot, so a column with few distinct values does
// not serialize millions of atomics on a handful of words; slots that do
// not fit the LDS table go straight to `first`.
constexpr u32 HASH_LDS = 1024, HASH_LDS_PROBES = 8;
template <class K>
__global__ __launch_bounds__(256) void k_hash_insert(const K* __restrict__ keys, const u32* __restrict__ seg, u64 n,
                                                     u32* __restrict__ table, u32 cmask, u32* __restrict__ slot_of,
                                                     u32* __restrict__ first) {
  __shared__ u32 s_slot[HASH_LDS], s_min[HASH_LDS];
  for (u32 j = threadIdx.x; j < HASH_LDS; j += blockDim.x) {
    s_slot[j] = HASH_EMPTY;
    s_min[j] = HASH_EMPTY;
  }
  __syncthreads();
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) {
    const K k = keys[i];
    const u32 g = seg ? seg[i] : 0u;
    u32 h = (u32)mix64((u64)k ^ ((u64)g * 0x9E3779B97F4A7C15ull)) & cmask;
    for (;;) {
      u32 cur = __hip_atomic_load(table + h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (cur == HASH_EMPTY) {
        u32 prev = atomicCAS(table + h, HASH_EMPTY, (u32)i);
        if (prev == HASH_EMPTY) break;
        cur = prev;
      }
      if (keys[cur] == k && (!seg || seg[cur] == g)) break;
      h = (h + 1) & cmask;
    }
    slot_of[i] = h;
    // block-local first row of slot h
    u32 e = (h * 0x9E3779B1u) >> 22;  // HASH_LDS = 2^10 buckets
    bool done = false;
    for (u32 p = 0; p < HASH_LDS_PROBES && !done; ++p, e = (e + 1) & (HASH_LDS - 1)) {
      u32 cur = s_slot[e];
      if (cur == HASH_EMPTY) cur = atomicCAS(&s_slot[e], HASH_EMPTY, h), cur = (cur == HASH_EMPTY) ? h : cur;
      if (cur == h) {
        atomicMin(&s_min[e], (u32)i);
        done = true;
      }
    }
    if (!done) atomicMin(first + h, (u32)i);
  }
  __syncthreads();
  for (u32 j = threadIdx.x; j < HASH_LDS; j += blockDim.x)
    if (s_slot[j] != HASH_EMPTY) atomicMin(first + s_slot[j], s_min[j]);
}

__global__ void k_hash_heads(const u32* __restrict__ slot_of, const u32* __restrict__ first, u64 n,
                             u32* __restrict__ head) {
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x)
    head[i] = (first[slot_of[i]] == (u32)i) ? 1u : 0u;
}

__global__ void k_hash_rank_slots(const u32* __restrict__ slot_of, const u32* __restrict__ head,
                                  const u32* __restrict__ prefix, u64 n, u32* __restrict__ rank_of_slot) {
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x)
    if (head[i]) rank_of_slot[slot_of[i]] = prefix[i];
}

__global__ void k_hash_assign(const u32* __restrict__ slot_of, const u32* __restrict__ rank_of_slot, u64 n,
                              u32* __restrict__ gid) {
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x)
    gid[i] = rank_of_slot[slot_of[i]];
}

// Synthetic banks and offsets for on-device calibration.
template <class K>
__global__ void k_fill_random(K* __restrict__ p, u64 n, u64 seed, u32 bits) {
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) {
    u64 z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    p[i] = (K)(bits >= 64 ? z : (z >> (64 - bits)));
  }
}

// Synthetic column (the seeded generator of the benchmarks; the reference
// writes its inputs with src/generator.h:139-227 and src/zipf.h:33-69):
// row i draws u = SplitMix64(seed, i); its rank is u mod card (uniform) or
// floor(card^x) - 1 with x = u / 2^64 (density ~ 1/(rank+1), Zipf s = 1);
// the value is (rank * step) & mask.
template <class T>
__global__ void k_fill_column(T* __restrict__ p, u64 n, u64 seed, u64 card, u32 zipf, double log2card, u64 step,
                              u64 mask) {
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) {
    u64 z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    u64 rank;
    if (zipf) {
      const double x = (double)(z >> 11) * 0x1.0p-53;
      const double v = floor(exp2(x * log2card));
      rank = v >= 1.0 ? (u64)v - 1 : 0;
      if (rank >= card) rank = card - 1;
    } else {
      rank = z % card;
    }
    p[i] = (T)((rank * step) & mask);
  }
}

__global__ void k_uniform_offsets(u32* __restrict__ off, u32 nseg, u32 size, u64 n) {
  for (u64 s = (u64)blockIdx.x * blockDim.x + threadIdx.x; s <= nseg; s += (u64)gridDim.x * blockDim.x)
    off[s] = (u32)min((u64)s * size, n);
}

// ---------------------------------------------------------------------------
// Explicit instantiations used by the launchers.
// ---------------------------------------------------------------------------
#define FMCS_INST_ONESWEEP_IF(K, I, P, F, L)                                                                    \
  template __global__ void k_onesweep<K, I, 8, P, F, L>(const K*, const u32*, K*, u32*, const u32*, const u32*,      \
                                                        const u32*, u32, const u32*, const u32*, const u32*, u32, u32, \
                                                        const u32*, const u32*, const P*, PayOut<P, F>*, u32, u64,   \
                                                        u32*, u32*, uint8_t*, u32, const u32*);
#define FMCS_INST_ONESWEEP_F(K, P, F, L) FMCS_INST_ONESWEEP_IF(K, ONESWEEP_ITEMS, P, F, L)
#define FMCS_INST_ONESWEEP(K, P) FMCS_INST_ONESWEEP_F(K, P, false, true)
#define FMCS_INST_K(K)                                                                                              \
  template __global__ void k_gather<K>(const K*, const u32*, K*, u64);                                              \
  template __global__ void k_gather_win<K>(const K*, const u32*, K*, u64, u32, u32);                                \
  template __global__ void k_bcount<K, BOUNDARY_ITEMS>(const K*, const u32*, u64, u32, u32*, u32*);                 \
  template __global__ void k_bwrite<K, BOUNDARY_ITEMS>(const K*, const u32*, u64, u32, const u32*, u32*, u32*, u32*, \
                                                       const u32*);                                                \
  template __global__ void k_hist_all<K>(const K*, u64, u32, u32, u64, u32*);                                            \
  FMCS_INST_ONESWEEP_F(K, NoPay, false, false)                                                                      \
  FMCS_INST_ONESWEEP_IF(K, WIDE_ITEMS, NoPay, false, false)                                                         \
  FMCS_INST_ONESWEEP_IF(K, WIDE_ITEMS, NoPay, false, true)                                                          \
  FMCS_INST_ONESWEEP(K, NoPay)                                                                                      \
  FMCS_INST_ONESWEEP(K, u16)                                                                                        \
  FMCS_INST_ONESWEEP(K, u32)                                                                                        \
  FMCS_INST_ONESWEEP(K, u64)                                                                                        \
  template __global__ void k_tile_sort<K, 4>(TileSortArgs<K>);                                                      \
  template __global__ void k_tile_sort<K, 2>(TileSortArgs<K>);                                                      \
  template __global__ void k_tile_rank<K, 4>(TileSortArgs<K>);                                                      \
  template __global__ void k_tile_rank<K, 2>(TileSortArgs<K>);                                                      \
  template __global__ void k_tile_bucket<K, 4>(TileSortArgs<K>, u32*, const u32*, const u32*, const u32*,          \
                                               const u32*, const u32*);                                           \
  template __global__ void k_tile_bucket<K, 2>(TileSortArgs<K>, u32*, const u32*, const u32*, const u32*,          \
                                               const u32*, const u32*);                                           \
  template __global__ void k_tile_bucket<K, 8>(TileSortArgs<K>, u32*, const u32*, const u32*, const u32*,          \
                                               const u32*, const u32*);                                           \
  template __global__ void k_tile_sort<K, 8>(TileSortArgs<K>);                                                      \
  template __global__ void k_tile_rank<K, 8>(TileSortArgs<K>);                                                      \
  template __global__ void k_copy_segs<K>(const K*, const u32*, K*, u32*, const u32*, const u32*, const u32*, u32,  \
                                          u32);                                                                     \
  template __global__ void k_composite<K>(const K*, const u32*, u64, u32, u64*, u32*);                              \
  template __global__ void k_key_check<K>(const K*, u64, u64, u32*);                                                \
  template __global__ void k_strip_composite<K>(const u64*, u64, u64, K*);                                          \
  template __global__ void k_hash_insert<K>(const K*, const u32*, u64, u32*, u32, u32*, u32*);                    \
  template __global__ void k_fill_random<K>(K*, u64, u64, u32);                                                    \
  template __global__ void k_fill_column<K>(K*, u64, u64, u64, u32, double, u64, u64);

FMCS_INST_K(u16)
FMCS_INST_K(u32)
// the MSD level that narrows u64 keys to the 32 bits still to sort
template __global__ void k_onesweep<u64, WIDE_ITEMS, 8, NoPay, false, false, u32>(
    const u64*, const u32*, u32*, u32*, const u32*, const u32*, const u32*, u32, const u32*, const u32*, const u32*,
    u32, u32, const u32*, const u32*, const NoPay*, PayOut<NoPay, false>*, u32, u64, u32*, u32*, uint8_t*, u32,
    const u32*);
FMCS_INST_ONESWEEP_F(u16, u64, true, true)
FMCS_INST_ONESWEEP_F(u32, u64, true, true)
template __global__ void k_hash_insert<uint8_t>(const uint8_t*, const u32*, u64, u32*, u32, u32*, u32*);
template __global__ void k_gather<uint8_t>(const uint8_t*, const u32*, uint8_t*, u64);
template __global__ void k_fill_random<uint8_t>(uint8_t*, u64, u64, u32);
template __global__ void k_rec_pack<8>(RecCols, void*, u64);
template __global__ void k_rec_pack<16>(RecCols, void*, u64);
template __global__ void k_rec_gather<8>(RecCols, const void*, const u32*, u64);
template __global__ void k_rec_gather<16>(RecCols, const void*, const u32*, u64);
FMCS_INST_K(u64)
template __global__ void k_scan_excl<8, false>(const u32*, u32*, u64, u32*, u32*, u32*);
template __global__ void k_scan_excl<8, true>(const u32*, u32*, u64, u32*, u32*, u32*);
template __global__ void k_scan_reduce<false>(const u32*, u64, u64, u32*);
template __global__ void k_scan_reduce<true>(const u32*, u64, u64, u32*);
template __global__ void k_scan_rescan<false>(const u32*, u32*, u64, u64, const u32*);
template __global__ void k_scan_rescan<true>(const u32*, u32*, u64, u64, const u32*);

}  // namespace fmcs
