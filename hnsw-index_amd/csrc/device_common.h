// device_common.h — wave-level building blocks shared by the gfx950 kernels:
// the exact-order f32 row distance, sorted register lists, LDS visited set.
//
// Every helper assumes a 64-lane wavefront.  A "half" is lanes 0-31 or
// 32-63; the row-distance helpers give each half its own row.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hhd {

constexpr uint32_t kInvalid = 0xFFFFFFFFu;

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
__device__ __forceinline__ float shx(float v, int m) { return __shfl_xor(v, m, 64); }

// f32::total_cmp key (the order of src/hnsw/search.rs:36-44 and
// src/neighbor.rs:17-21).
__device__ __forceinline__ int32_t tkey(float f) {
  int32_t i = __float_as_int(f);
  return i ^ (int32_t)(((uint32_t)(i >> 31)) >> 1);
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
// A vector as one lane of a half-wavefront holds it: lane j (0..31) keeps
// values 32t + j for t < T (the AVX2 accumulator-j stream), lanes j < 8 keep
// the 8-wide remainder values 32T + 8u + j, lanes j < R the scalar tail.
template <int TMAX>
struct Lanes {
  float m[TMAX];
  float e[3];
  float r;
};

struct Shape {
  int T, U, R;       // dim = 32T + 8U + R
  uint32_t stride;   // floats per row
};

// Loads the lane's share of the row at p (accumulator-major layout).
template <int TMAX>
__device__ __forceinline__ void load_lanes(const float *__restrict__ p, const Shape &s, int j,
                                           Lanes<TMAX> &x) {
#pragma unroll
  for (int b = 0; b < (TMAX + 3) / 4; b++) {
    const int t0 = 4 * b;
    if (t0 < s.T) {
      const int tw = s.T - t0 < 4 ? s.T - t0 : 4;
      const float *blk = p + b * 128;
      if (tw == 4) {
        float4 v = *reinterpret_cast<const float4 *>(blk + 4 * j);
        x.m[t0] = v.x;
        if (t0 + 1 < TMAX) x.m[t0 + 1] = v.y;
        if (t0 + 2 < TMAX) x.m[t0 + 2] = v.z;
        if (t0 + 3 < TMAX) x.m[t0 + 3] = v.w;
      } else if (tw == 2) {
        float2 v = *reinterpret_cast<const float2 *>(blk + 2 * j);
        x.m[t0] = v.x;
        if (t0 + 1 < TMAX) x.m[t0 + 1] = v.y;
      } else if (tw == 3) {
        x.m[t0] = blk[3 * j];
        if (t0 + 1 < TMAX) x.m[t0 + 1] = blk[3 * j + 1];
        if (t0 + 2 < TMAX) x.m[t0 + 2] = blk[3 * j + 2];
      } else {
        x.m[t0] = blk[j];
      }
    }
  }
  const float *rest = p + 32 * s.T;
#pragma unroll
  for (int u = 0; u < 3; u++)
    if (u < s.U && j < 8) x.e[u] = rest[8 * u + j];
  if (j < s.R) x.r = rest[8 * s.U + j];
}

// bf16 form of load_lanes: same positions, 2-byte values widened to f32.
__device__ __forceinline__ float bf16f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

template <int TMAX>
__device__ __forceinline__ void load_lanes_bf16(const uint16_t *__restrict__ p, const Shape &s,
                                                int j, Lanes<TMAX> &x) {
#pragma unroll
  for (int b = 0; b < (TMAX + 3) / 4; b++) {
    const int t0 = 4 * b;
    if (t0 < s.T) {
      const int tw = s.T - t0 < 4 ? s.T - t0 : 4;
      const uint16_t *blk = p + b * 128;
      if (tw == 4) {
        uint2 v = *reinterpret_cast<const uint2 *>(blk + 4 * j);
        x.m[t0] = __uint_as_float(v.x << 16);
        if (t0 + 1 < TMAX) x.m[t0 + 1] = __uint_as_float(v.x & 0xFFFF0000u);
        if (t0 + 2 < TMAX) x.m[t0 + 2] = __uint_as_float(v.y << 16);
        if (t0 + 3 < TMAX) x.m[t0 + 3] = __uint_as_float(v.y & 0xFFFF0000u);
      } else {
#pragma unroll
        for (int tt = 0; tt < 4; tt++)
          if (tt < tw && t0 + tt < TMAX) x.m[t0 + tt] = bf16f(blk[tw * j + tt]);
      }
    }
  }
  const uint16_t *rest = p + 32 * s.T;
#pragma unroll
  for (int u = 0; u < 3; u++)
    if (u < s.U && j < 8) x.e[u] = bf16f(rest[8 * u + j]);
  if (j < s.R) x.r = bf16f(rest[8 * s.U + j]);
}

// Exact-order partial for one half: returns the accumulator of lane j
// (src/kernels/x86.rs:47-69 / :96-120) and the lane's tail product.
template <int TMAX, bool L2>
__device__ __forceinline__ float lane_acc(const Lanes<TMAX> &q, const Lanes<TMAX> &x,
                                          const Shape &s, int j, float &tail) {
  float acc = 0.0f;
#pragma unroll
  for (int t = 0; t < TMAX; t++) {
    if (t < s.T) {
      if (L2) {
        float d = q.m[t] - x.m[t];
        acc = __builtin_fmaf(d, d, acc);
      } else {
        acc = __builtin_fmaf(q.m[t], x.m[t], acc);
      }
    }
  }
#pragma unroll
  for (int u = 0; u < 3; u++) {
    if (u < s.U && j < 8) {
      if (L2) {
        float d = q.e[u] - x.e[u];
        acc = __builtin_fmaf(d, d, acc);
      } else {
        acc = __builtin_fmaf(q.e[u], x.e[u], acc);
      }
    }
  }
  tail = 0.0f;
  if (j < s.R) {
    if (L2) {
      float d = q.r - x.r;
      tail = __fmul_rn(d, d);
    } else {
      tail = __fmul_rn(q.r, x.r);
    }
  }
  return acc;
}

// The AVX2 lane combination s[l] = (a0+a1)+(a2+a3), then
// h = ((s0+s4)+(s1+s5))+((s2+s6)+(s3+s7)) (src/kernels/x86.rs:21-27,70-73),
// then the unfused scalar tail in order (x86.rs:75-77).  Every lane of the
// half ends with the same bits (IEEE addition is commutative).  Must be
// called by all 64 lanes.
__device__ __forceinline__ float half_reduce(float acc, float tail, int R, int half) {
  acc = __fadd_rn(acc, shx(acc, 8));
  acc = __fadd_rn(acc, shx(acc, 16));
  acc = __fadd_rn(acc, shx(acc, 4));
  acc = __fadd_rn(acc, shx(acc, 1));
  acc = __fadd_rn(acc, shx(acc, 2));
  for (int r = 0; r < R; r++) acc = __fadd_rn(acc, __shfl(tail, (half << 5) | r, 64));
  return acc;
}

__device__ __forceinline__ float finish_metric(float sum, uint32_t metric) {
  if (metric == 0) return sum;
  if (metric == 1) return -sum;
  return __fsub_rn(1.0f, sum);
}

// ------------------------------------------------------ sorted lists
// A sorted list of up to 64*S (distance, id) pairs spread over the wave:
// element i lives in lane i / S, slot i % S.  size is wave-uniform.
template <int S>
struct SList {
  float d[S];
  uint32_t id[S];
};

// Position of (nd, nid) in the first `size` elements.
template <int S>
__device__ __forceinline__ int slist_rank(const SList<S> &L, int size, float nd, uint32_t nid) {
  const int lane = lane_id();
  int p = 0;
#pragma unroll
  for (int s = 0; s < S; s++) {
    bool lt = (lane * S + s) < size && cless(L.d[s], L.id[s], nd, nid);
    p += popc64(__ballot(lt));
  }
  return p;
}

// Inserts at position p (< cap), shifting the tail right; the element at
// index cap-1 falls off when the list was full.
template <int S>
__device__ __forceinline__ void slist_insert_at(SList<S> &L, int p, float nd, uint32_t nid) {
  const int lane = lane_id();
  float upd = __shfl_up(L.d[S - 1], 1, 64);
  uint32_t upi = (uint32_t)__shfl_up((int)L.id[S - 1], 1, 64);
#pragma unroll
  for (int s = S - 1; s >= 0; s--) {
    const int idx = lane * S + s;
    float pd = s > 0 ? L.d[s > 0 ? s - 1 : 0] : upd;
    uint32_t pi = s > 0 ? L.id[s > 0 ? s - 1 : 0] : upi;
    if (idx > p) {
      L.d[s] = pd;
      L.id[s] = pi;
    } else if (idx == p) {
      L.d[s] = nd;
      L.id[s] = nid;
    }
  }
}

// Removes element 0.
template <int S>
__device__ __forceinline__ void slist_pop_front(SList<S> &L) {
  float dn = __shfl_down(L.d[0], 1, 64);
  uint32_t in = (uint32_t)__shfl_down((int)L.id[0], 1, 64);
#pragma unroll
  for (int s = 0; s < S; s++) {
    L.d[s] = s < S - 1 ? L.d[s < S - 1 ? s + 1 : 0] : dn;
    L.id[s] = s < S - 1 ? L.id[s < S - 1 ? s + 1 : 0] : in;
  }
}

// Element i (wave-uniform i).
template <int S>
__device__ __forceinline__ void slist_get(const SList<S> &L, int i, float &d, uint32_t &id) {
  const int ln = i / S, sl = i % S;
  float vd = L.d[0];
  uint32_t vi = L.id[0];
#pragma unroll
  for (int s = 1; s < S; s++)
    if (sl == s) {
      vd = L.d[s];
      vi = L.id[s];
    }
  d = rlf(vd, ln);
  id = rl(vi, ln);
}

// Number of leading elements whose distance is not IEEE-greater than w.
template <int S>
__device__ __forceinline__ int slist_count_not_above(const SList<S> &L, int size, float w) {
  const int lane = lane_id();
  int c = 0;
#pragma unroll
  for (int s = 0; s < S; s++) {
    bool keep = (lane * S + s) < size && !(L.d[s] > w);
    c += popc64(__ballot(keep));
  }
  return c;
}

// ------------------------------------------------------- visited set
// Open-addressing hash of node ids in LDS (exact: no false positives or
// negatives).  Returns true when id was not present and is now inserted.
__device__ __forceinline__ bool lds_visit(uint32_t *table, uint32_t mask, int shift, uint32_t id) {
  uint32_t h = (id * 0x9E3779B1u) >> shift;
  for (;;) {
    uint32_t prev = atomicCAS(&table[h], kInvalid, id);
    if (prev == kInvalid) return true;
    if (prev == id) return false;
    h = (h + 1) & mask;
  }
}

__device__ __forceinline__ bool bit_visit(uint32_t *bits, uint32_t id) {
  uint32_t b = 1u << (id & 31);
  uint32_t old = atomicOr(&bits[id >> 5], b);
  return (old & b) == 0;
}

__device__ __forceinline__ bool bit_test(const uint32_t *bits, uint32_t id) {
  return (bits[id >> 5] >> (id & 31)) & 1u;
}

}  // namespace hhd
