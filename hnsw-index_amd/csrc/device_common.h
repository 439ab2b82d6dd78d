// device_common.h — wave-level building blocks shared by the gfx950 kernels:
// the exact-order row distance, sorted register lists, the LDS visited set.
//
// Every helper assumes a 64-lane wavefront.  Row distances are computed by
// "quarters": the four 16-lane DPP rows of the wave each take one vector
// row, and lane i of a quarter owns the crate's AVX2 accumulator lanes
// j = i and j = i + 16 (src/kernels/x86.rs: accumulator c, lane l is
// j = 8c + l).  The crate's fixed reduction tree then maps onto two DPP
// row rotations, one in-lane add and two DPP quad permutations.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hhd {

constexpr uint32_t kInvalid = 0xFFFFFFFFu;

// Global-address-space views of device pointers: loads through them compile
// to global_load (vector memory counter only), not flat_load.  The host
// pass of the compiler sees plain pointers (it never runs this code).
#if defined(__HIP_DEVICE_COMPILE__)
template <class T>
using gptr = __attribute__((address_space(1))) const T *;
template <class T>
using gmut = __attribute__((address_space(1))) T *;
#else
template <class T>
using gptr = const T *;
template <class T>
using gmut = T *;
#endif
template <class T>
__device__ __forceinline__ gptr<T> G(const T *p) {
  return (gptr<T>)p;
}
template <class T>
__device__ __forceinline__ gmut<T> GM(T *p) {
  return (gmut<T>)p;
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// Orders LDS traffic between the lanes of one wavefront (no block barrier:
// waves of a block run independent loops).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t rl(uint32_t v, int l) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}
__device__ __forceinline__ float rlf(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ uint64_t rl64(uint64_t v, int l) {
  return ((uint64_t)rl((uint32_t)(v >> 32), l) << 32) | rl((uint32_t)v, l);
}
__device__ __forceinline__ int popc64(uint64_t m) { return __popcll(m); }
// Number of set bits of m below the calling lane.
__device__ __forceinline__ int mbcnt(uint64_t m) {
  return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                        __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// DPP lane moves within a 16-lane row.
template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ int dppi(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}
constexpr int kRor8 = 0x128, kRor4 = 0x124, kXor1 = 0xB1, kXor2 = 0x4E;

// f32::total_cmp key (the order of src/hnsw/search.rs:36-44 and
// src/neighbor.rs:17-21).
__device__ __forceinline__ int32_t tkey(float f) {
  int32_t i = __float_as_int(f);
  return i ^ (int32_t)(((uint32_t)(i >> 31)) >> 1);
}
// Bin of a query's descent distance for the longest-first query order:
// the top 12 bits of the total-order key (an eighth of a binade).
__device__ __forceinline__ uint32_t order_bin(float f) {
  return ((uint32_t)tkey(f) ^ 0x80000000u) >> 20;
}
// Candidate order: (distance total order, node id).
__device__ __forceinline__ bool cless(float da, uint32_t ia, float db, uint32_t ib) {
  int32_t a = tkey(da), b = tkey(db);
  return a < b || (a == b && ia < ib);
}
// Neighbor order: (distance total order, key).
__device__ __forceinline__ bool nless(float da, uint64_t ka, float db, uint64_t kb) {
  int32_t a = tkey(da), b = tkey(db);
  return a < b || (a == b && ka < kb);
}

// ------------------------------------------------------------------ rows
// dim = 32T + 8U + R.  Rows are stored accumulator-major: value 32t + j at
//   (t/4)*128 + j*tw + t%4,  tw = min(4, T - 4*(t/4)),
// then the 8U + R remainder values in natural order.
struct Shape {
  int T, U, R;
  uint32_t stride;  // elements per row
};

// One quarter-lane's share of a vector: values 32t + i (lo) and
// 32t + i + 16 (hi) for t < T, the remainder values 32T + 8u + i for i < 8,
// the tail value 32T + 8U + i for i < R.
template <int TMAX, bool EXTRA>
struct QLanes {
  float lo[TMAX], hi[TMAX];
  float e[EXTRA ? 3 : 1];
  float r;
};

template <class T>
__device__ __forceinline__ float to_f(T v);
template <>
__device__ __forceinline__ float to_f<float>(float v) { return v; }
template <>
__device__ __forceinline__ float to_f<uint16_t>(uint16_t v) {
  return __uint_as_float((uint32_t)v << 16);
}

// Loads lane i's share of the row at p (f32 or bf16 elements).
// EXACT: the dimension is exactly 32 * TMAX (D = 128, 256, 1024), so every
// 4-block is full and there is no remainder: straight-line loads with no
// shape branches, which lets a caller issue several rows' loads before
// waiting for any of them.
template <int TMAX, bool EXTRA, class E, bool EXACT = false>
__device__ __forceinline__ void load_q(const E *p_, const Shape &s, int i,
                                       QLanes<TMAX, EXTRA> &x) {
  gptr<E> p = G(p_);
  if constexpr (EXACT) {
    static_assert(!EXTRA && TMAX % 4 == 0, "exact rows have no remainder");
#pragma unroll
    for (int b = 0; b < TMAX / 4; b++) {
      gptr<E> blk = p + b * 128;
      if constexpr (sizeof(E) == 4) {
        const float4 a = *(gptr<float4>)(blk + 4 * i);
        const float4 c = *(gptr<float4>)(blk + 64 + 4 * i);
        x.lo[4 * b] = a.x; x.lo[4 * b + 1] = a.y; x.lo[4 * b + 2] = a.z; x.lo[4 * b + 3] = a.w;
        x.hi[4 * b] = c.x; x.hi[4 * b + 1] = c.y; x.hi[4 * b + 2] = c.z; x.hi[4 * b + 3] = c.w;
      } else {
        const uint2 a = *(gptr<uint2>)(blk + 4 * i);
        const uint2 c = *(gptr<uint2>)(blk + 64 + 4 * i);
        x.lo[4 * b] = __uint_as_float(a.x << 16);
        x.lo[4 * b + 1] = __uint_as_float(a.x & 0xFFFF0000u);
        x.lo[4 * b + 2] = __uint_as_float(a.y << 16);
        x.lo[4 * b + 3] = __uint_as_float(a.y & 0xFFFF0000u);
        x.hi[4 * b] = __uint_as_float(c.x << 16);
        x.hi[4 * b + 1] = __uint_as_float(c.x & 0xFFFF0000u);
        x.hi[4 * b + 2] = __uint_as_float(c.y << 16);
        x.hi[4 * b + 3] = __uint_as_float(c.y & 0xFFFF0000u);
      }
    }
    return;
  }
#pragma unroll
  for (int b = 0; b < (TMAX + 3) / 4; b++) {
    const int t0 = 4 * b;
    if (t0 < s.T) {
      const int tw = s.T - t0 < 4 ? s.T - t0 : 4;
      gptr<E> blk = p + b * 128;
      if (tw == 4) {
        if constexpr (sizeof(E) == 4) {
          const float4 a = *(gptr<float4>)(blk + 4 * i);
          const float4 c = *(gptr<float4>)(blk + 64 + 4 * i);
          x.lo[t0] = a.x;
          x.hi[t0] = c.x;
          if (t0 + 1 < TMAX) { x.lo[t0 + 1] = a.y; x.hi[t0 + 1] = c.y; }
          if (t0 + 2 < TMAX) { x.lo[t0 + 2] = a.z; x.hi[t0 + 2] = c.z; }
          if (t0 + 3 < TMAX) { x.lo[t0 + 3] = a.w; x.hi[t0 + 3] = c.w; }
        } else {
          const uint2 a = *(gptr<uint2>)(blk + 4 * i);
          const uint2 c = *(gptr<uint2>)(blk + 64 + 4 * i);
          x.lo[t0] = __uint_as_float(a.x << 16);
          x.hi[t0] = __uint_as_float(c.x << 16);
          if (t0 + 1 < TMAX) {
            x.lo[t0 + 1] = __uint_as_float(a.x & 0xFFFF0000u);
            x.hi[t0 + 1] = __uint_as_float(c.x & 0xFFFF0000u);
          }
          if (t0 + 2 < TMAX) {
            x.lo[t0 + 2] = __uint_as_float(a.y << 16);
            x.hi[t0 + 2] = __uint_as_float(c.y << 16);
          }
          if (t0 + 3 < TMAX) {
            x.lo[t0 + 3] = __uint_as_float(a.y & 0xFFFF0000u);
            x.hi[t0 + 3] = __uint_as_float(c.y & 0xFFFF0000u);
          }
        }
      } else {
#pragma unroll
        for (int tt = 0; tt < 3; tt++)
          if (tt < tw && t0 + tt < TMAX) {
            x.lo[t0 + tt] = to_f<E>(blk[tw * i + tt]);
            x.hi[t0 + tt] = to_f<E>(blk[tw * (i + 16) + tt]);
          }
      }
    }
  }
  if constexpr (EXTRA) {
    gptr<E> rest = p + 32 * s.T;
#pragma unroll
    for (int u = 0; u < 3; u++)
      if (u < s.U && i < 8) x.e[u] = to_f<E>(rest[8 * u + i]);
    if (i < s.R) x.r = to_f<E>(rest[8 * s.U + i]);
  }
}

// Lane i's accumulators j = i (lo) and j = i + 16 (hi) of the AVX2 kernel
// (src/kernels/x86.rs:47-69 dot, :96-120 l2): one fused multiply-add per
// element in t order, the 8-wide remainder into accumulator lanes 0..7,
// and the unfused tail product (x86.rs:75-77, 126-129).
template <int TMAX, bool EXTRA, bool L2, bool EXACT = false>
__device__ __forceinline__ void quarter_acc(const QLanes<TMAX, EXTRA> &q,
                                            const QLanes<TMAX, EXTRA> &x, const Shape &s, int i,
                                            float &lo, float &hi, float &tail) {
  lo = 0.0f;
  hi = 0.0f;
#pragma unroll
  for (int t = 0; t < TMAX; t++) {
    if (EXACT || t < s.T) {
      if (L2) {
        const float dl = q.lo[t] - x.lo[t], dh = q.hi[t] - x.hi[t];
        lo = __builtin_fmaf(dl, dl, lo);
        hi = __builtin_fmaf(dh, dh, hi);
      } else {
        lo = __builtin_fmaf(q.lo[t], x.lo[t], lo);
        hi = __builtin_fmaf(q.hi[t], x.hi[t], hi);
      }
    }
  }
  tail = 0.0f;
  if constexpr (EXTRA) {
#pragma unroll
    for (int u = 0; u < 3; u++) {
      if (u < s.U && i < 8) {
        if (L2) {
          const float d = q.e[u] - x.e[u];
          lo = __builtin_fmaf(d, d, lo);
        } else {
          lo = __builtin_fmaf(q.e[u], x.e[u], lo);
        }
      }
    }
    if (i < s.R) {
      if (L2) {
        const float d = q.r - x.r;
        tail = __fmul_rn(d, d);
      } else {
        tail = __fmul_rn(q.r, x.r);
      }
    }
  }
}

// The crate's lane combination: s[l] = (acc0+acc1)+(acc2+acc3) — lane l
// holds acc_l (= acc0[l]) and acc_{l+16} (= acc2[l]), and acc_{l+8} /
// acc_{l+24} sit 8 lanes away — then
// h = ((s0+s4)+(s1+s5))+((s2+s6)+(s3+s7)) (src/kernels/x86.rs:21-27,70-73),
// then the tail values added one by one in order.  Every lane of a quarter
// ends with identical bits (IEEE addition commutes).  All lanes must call.
template <bool EXTRA>
__device__ __forceinline__ float quarter_reduce(float lo, float hi, float tail, int R, int quarter) {
  const float a = __fadd_rn(lo, dpp<kRor8>(lo));
  const float b = __fadd_rn(hi, dpp<kRor8>(hi));
  float s = __fadd_rn(a, b);
  s = __fadd_rn(s, dpp<kRor4>(s));
  s = __fadd_rn(s, dpp<kXor1>(s));
  s = __fadd_rn(s, dpp<kXor2>(s));
  if constexpr (EXTRA) {
    for (int r = 0; r < R; r++) s = __fadd_rn(s, __shfl(tail, (quarter << 4) | r, 64));
  }
  return s;
}

// Integer and approximate sums: any order (the same tree is used).
__device__ __forceinline__ int quarter_sum_i(int v) {
  v += dppi<kRor8>(v);
  v += dppi<kRor4>(v);
  v += dppi<kXor1>(v);
  v += dppi<kXor2>(v);
  return v;
}

__device__ __forceinline__ float finish_metric(float sum, uint32_t metric) {
  if (metric == 0) return sum;
  if (metric == 1) return -sum;
  return __fsub_rn(1.0f, sum);
}

// ------------------------------------------------------- visited set
// Open-addressing hash of node ids in LDS, exact (no false positives or
// negatives), any capacity (Lemire range reduction), linear probing.
// Returns true when id was absent and is now inserted.  (A per-lane probe
// loop: measured 2% faster in the f32 search than a wave-uniform loop with
// the atomic masked, at 1M x 128, ef=64.)
__device__ __forceinline__ bool lds_visit(uint32_t *table, uint32_t cap, uint32_t id) {
  uint32_t h = __umulhi(id * 0x9E3779B1u, cap);
  for (;;) {
    const uint32_t prev = atomicCAS(&table[h], kInvalid, id);
    if (prev == kInvalid) return true;
    if (prev == id) return false;
    h = h + 1 == cap ? 0 : h + 1;
  }
}

// Two insertions per lane with both probes in flight (kInvalid: none).  The
// first is applied first, so an id passed twice is new only once.
__device__ __forceinline__ void lds_visit2(uint32_t *table, uint32_t cap, uint32_t a, uint32_t b,
                                           bool &na, bool &nb) {
  uint32_t ha = __umulhi(a * 0x9E3779B1u, cap), hb = __umulhi(b * 0x9E3779B1u, cap);
  bool pa = a != kInvalid, pb = b != kInvalid;
  na = nb = false;
  while (__ballot(pa || pb)) {
    uint32_t ra = a, rb = b;
    if (pa) ra = atomicCAS(&table[ha], kInvalid, a);
    if (pb) rb = atomicCAS(&table[hb], kInvalid, b);
    const bool da = ra == kInvalid || ra == a, db = rb == kInvalid || rb == b;
    na = na || (pa && ra == kInvalid);
    nb = nb || (pb && rb == kInvalid);
    const uint32_t nha = ha + 1 == cap ? 0u : ha + 1, nhb = hb + 1 == cap ? 0u : hb + 1;
    ha = pa && !da ? nha : ha;
    hb = pb && !db ? nhb : hb;
    pa = pa && !da;
    pb = pb && !db;
  }
}

__device__ __forceinline__ bool bit_visit(uint32_t *bits, uint32_t id) {
  const uint32_t b = 1u << (id & 31);
  const uint32_t old = __hip_atomic_fetch_or(bits + (id >> 5), b, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
  return (old & b) == 0;
}

__device__ __forceinline__ bool bit_test(const uint32_t *bits, uint32_t id) {
  return (G(bits)[id >> 5] >> (id & 31)) & 1u;
}

}  // namespace hhd
