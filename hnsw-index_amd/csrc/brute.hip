// brute.hip — exact brute-force kNN on gfx950 (BruteForceIndex,
// src/brute.rs:131-180).
//
// Stage 1 (brute_partial_kernel): a wavefront holds QB queries in
// registers and streams one slice of the rows, four rows per step (one per
// 16-lane quarter), computing every distance in the crate's AVX2 order and
// keeping a sorted (distance, key) top-k list per query in registers.
// Stage 2 (brute_merge_kernel): one wavefront per query merges the slice
// lists.  Ranking is the crate's (distance total order, key).
// The f64 variant evaluates every distance in double precision: the FP64
// oracle used to measure recall.
#include <hip/hip_runtime.h>

#include "brute_kernels.h"
#include "device_common.h"

using namespace hhd;

namespace hhk {

// Sorted (distance, key) list of up to 64*S entries (element i in lane
// i/S, slot i%S); D is float or double.
template <typename D, int S>
struct KList {
  D d[S];
  uint64_t key[S];
};

template <typename D>
__device__ __forceinline__ bool kless(D da, uint64_t ka, D db, uint64_t kb);
template <>
__device__ __forceinline__ bool kless<float>(float da, uint64_t ka, float db, uint64_t kb) {
  return nless(da, ka, db, kb);
}
template <>
__device__ __forceinline__ bool kless<double>(double da, uint64_t ka, double db, uint64_t kb) {
  return da < db || (da == db && ka < kb);
}

__device__ __forceinline__ double rld(double v, int l) {
  uint64_t b = (uint64_t)__double_as_longlong(v);
  return __longlong_as_double((long long)rl64(b, l));
}
__device__ __forceinline__ float rlv(float v, int l) { return rlf(v, l); }
__device__ __forceinline__ double rlv(double v, int l) { return rld(v, l); }

template <typename D, int S>
__device__ __forceinline__ void klist_insert(KList<D, S> &L, int size, D nd, uint64_t nk) {
  const int lane = lane_id();
  int p = 0;
#pragma unroll
  for (int s = 0; s < S; s++) p += popc64(__ballot((lane * S + s) < size && kless<D>(L.d[s], L.key[s], nd, nk)));
  D upd = __shfl_up(L.d[S - 1], 1, 64);
  uint64_t upk = (uint64_t)__shfl_up((long long)L.key[S - 1], 1, 64);
#pragma unroll
  for (int s = S - 1; s >= 0; s--) {
    const int idx = lane * S + s;
    D pd = s > 0 ? L.d[s > 0 ? s - 1 : 0] : upd;
    uint64_t pk = s > 0 ? L.key[s > 0 ? s - 1 : 0] : upk;
    if (idx > p) {
      L.d[s] = pd;
      L.key[s] = pk;
    } else if (idx == p) {
      L.d[s] = nd;
      L.key[s] = nk;
    }
  }
}

template <typename D, int S>
__device__ __forceinline__ void klist_get(const KList<D, S> &L, int i, D &d, uint64_t &k) {
  const int ln = i / S, sl = i % S;
  D vd = L.d[0];
  uint64_t vk = L.key[0];
#pragma unroll
  for (int s = 1; s < S; s++)
    if (sl == s) {
      vd = L.d[s];
      vk = L.key[s];
    }
  d = rlv(vd, ln);
  k = rl64(vk, ln);
}

// Offers (d, key) to a list holding the best kk entries; size/kth are
// wave-uniform and updated in place.
template <typename D, int S>
__device__ __forceinline__ void klist_offer(KList<D, S> &L, int &size, int kk, D &kth_d,
                                            uint64_t &kth_k, D d, uint64_t key) {
  if (size < kk || kless<D>(d, key, kth_d, kth_k)) {
    klist_insert<D, S>(L, size, d, key);
    if (size < kk) size++;
    klist_get<D, S>(L, size - 1, kth_d, kth_k);
  }
}

template <int TMAX, bool EXTRA, int QB, int S>
__global__ void __launch_bounds__(256) brute_partial_kernel(const BruteArgs A) {
  const int lane = lane_id(), i = lane & 15, quarter = lane >> 4;
  const uint64_t work = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const uint64_t groups = (A.nq + QB - 1) / QB;
  if (work >= groups * A.nslices) return;
  const uint64_t g = work / A.nslices, slice = work % A.nslices;
  const uint64_t r0 = slice * A.rows_per_slice;
  const uint64_t r1 = r0 + A.rows_per_slice < A.n ? r0 + A.rows_per_slice : A.n;
  const Shape s{A.T, A.U, A.R, A.stride};
  if (A.only) {
    bool any = false;
    for (int b = 0; b < QB; b++)
      if (g * QB + b < A.nq && G(A.only)[g * QB + b]) any = true;
    if (!any) return;
  }
  QLanes<TMAX, EXTRA> q[QB];
  int qvalid[QB];
#pragma unroll
  for (int b = 0; b < QB; b++) {
#pragma unroll
    for (int t = 0; t < TMAX; t++) q[b].lo[t] = q[b].hi[t] = 0.0f;
    q[b].e[0] = q[b].r = 0.0f;
    if constexpr (EXTRA) q[b].e[1] = q[b].e[2] = 0.0f;
    const uint64_t qi = g * QB + b;
    qvalid[b] = qi < A.nq && G(A.status)[qi] == 0 && (!A.only || G(A.only)[qi]);
    if (qi < A.nq) load_q<TMAX, EXTRA, float>(A.qprep + qi * A.stride, s, i, q[b]);
  }
  KList<float, S> L[QB];
  int size[QB];
  float kth_d[QB];
  uint64_t kth_k[QB];
#pragma unroll
  for (int b = 0; b < QB; b++) {
#pragma unroll
    for (int t = 0; t < S; t++) {
      L[b].d[t] = 0.0f;
      L[b].key[t] = 0;
    }
    size[b] = 0;
    kth_d[b] = 0.0f;
    kth_k[b] = 0;
  }
  const int kk = (int)A.k;
  for (uint64_t base = r0; base < r1; base += 4) {
    const uint64_t row = base + quarter;
    const bool rv = row < r1 && (!A.allow || bit_test(A.allow, (uint32_t)row));
    QLanes<TMAX, EXTRA> x;
    uint64_t key = 0;
    if (rv) {
      load_q<TMAX, EXTRA, float>(A.rows + row * A.stride, s, i, x);
      key = G(A.keys)[row];
    }
    bool vq[4];
    uint64_t kq[4];
#pragma unroll
    for (int w = 0; w < 4; w++) {
      vq[w] = rl((uint32_t)rv, 16 * w) != 0;
      kq[w] = rl64(key, 16 * w);
    }
#pragma unroll
    for (int b = 0; b < QB; b++) {
      float lo = 0.0f, hi = 0.0f, tail = 0.0f;
      if (rv) {
        if (A.metric == 0) quarter_acc<TMAX, EXTRA, true>(q[b], x, s, i, lo, hi, tail);
        else quarter_acc<TMAX, EXTRA, false>(q[b], x, s, i, lo, hi, tail);
      }
      const float h = finish_metric(quarter_reduce<EXTRA>(lo, hi, tail, s.R, quarter), A.metric);
#pragma unroll
      for (int w = 0; w < 4; w++)
        if (vq[w]) klist_offer<float, S>(L[b], size[b], kk, kth_d[b], kth_k[b], rlf(h, 16 * w), kq[w]);
    }
  }
#pragma unroll
  for (int b = 0; b < QB; b++) {
    const uint64_t qi = g * QB + b;
    if (qi >= A.nq || !qvalid[b]) continue;
    const uint64_t o = (qi * A.nslices + slice) * (uint64_t)A.k;
#pragma unroll
    for (int t = 0; t < S; t++) {
      const int idx = lane * S + t;
      if (idx < size[b]) {
        GM(A.part_d)[o + idx] = L[b].d[t];
        GM(A.part_k)[o + idx] = L[b].key[t];
      }
    }
    if (lane == 0) GM(A.part_n)[qi * A.nslices + slice] = (uint32_t)size[b];
  }
}

template <int S>
__global__ void __launch_bounds__(256) brute_merge_kernel(const BruteArgs A) {
  const int lane = lane_id();
  const uint64_t qi = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (qi >= A.nq) return;
  if (A.only && !G(A.only)[qi]) return;
  if (G(A.status)[qi] != 0) {
    if (lane == 0) GM(A.out_count)[qi] = 0;
    return;
  }
  KList<float, S> L;
#pragma unroll
  for (int t = 0; t < S; t++) {
    L.d[t] = 0.0f;
    L.key[t] = 0;
  }
  int size = 0;
  float kth_d = 0.0f;
  uint64_t kth_k = 0;
  const int kk = (int)A.k;
  for (uint32_t sl = 0; sl < A.nslices; sl++) {
    const uint64_t o = (qi * A.nslices + sl) * (uint64_t)A.k;
    const int cnt = (int)G(A.part_n)[qi * A.nslices + sl];
    for (int i = 0; i < cnt; i++) {
      const float d = G(A.part_d)[o + i];
      const uint64_t key = G(A.part_k)[o + i];
      if (!(size < kk || nless(d, key, kth_d, kth_k))) break;  // slice lists are sorted
      klist_offer<float, S>(L, size, kk, kth_d, kth_k, d, key);
    }
  }
  const uint64_t ob = qi * (uint64_t)A.k_stride;
#pragma unroll
  for (int t = 0; t < S; t++) {
    const int idx = lane * S + t;
    if (idx < size) {
      A.out_keys[ob + idx] = L.key[t];
      A.out_dist[ob + idx] = L.d[t];
    }
  }
  if (lane == 0) A.out_count[qi] = (uint32_t)size;
}

// FP64 oracle: one query per wave and slice, f64 distances, same ranking.
template <int S>
__global__ void __launch_bounds__(256) brute_f64_kernel(const BruteArgs A) {
  const int lane = lane_id(), i = lane & 15, quarter = lane >> 4;
  const uint64_t work = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (work >= A.nq * A.nslices) return;
  const uint64_t qi = work / A.nslices, slice = work % A.nslices;
  if (G(A.status)[qi] != 0) return;
  const uint64_t r0 = slice * A.rows_per_slice;
  const uint64_t r1 = r0 + A.rows_per_slice < A.n ? r0 + A.rows_per_slice : A.n;
  gptr<float> qp = G(A.qprep) + qi * A.stride;
  KList<double, S> L;
#pragma unroll
  for (int t = 0; t < S; t++) {
    L.d[t] = 0.0;
    L.key[t] = 0;
  }
  int size = 0;
  double kth_d = 0.0;
  uint64_t kth_k = 0;
  const int kk = (int)A.k;
  for (uint64_t base = r0; base < r1; base += 4) {
    const uint64_t row = base + quarter;
    const bool rv = row < r1;
    double acc = 0.0;
    uint64_t key = 0;
    if (rv) {
      gptr<float> xp = G(A.rows) + row * A.stride;
      // rows and queries share the layout, so a position is a position
      for (uint32_t pos = i; pos < A.dim; pos += 16) {
        const double a = (double)qp[pos], c = (double)xp[pos];
        if (A.metric == 0) {
          const double d = a - c;
          acc += d * d;
        } else {
          acc += a * c;
        }
      }
      key = G(A.keys)[row];
    }
#pragma unroll
    for (int m = 1; m < 16; m <<= 1) acc += __shfl_xor(acc, m, 64);
    const double dist = A.metric == 0 ? acc : (A.metric == 1 ? -acc : 1.0 - acc);
#pragma unroll
    for (int w = 0; w < 4; w++)
      if (rl((uint32_t)rv, 16 * w))
        klist_offer<double, S>(L, size, kk, kth_d, kth_k, rld(dist, 16 * w), rl64(key, 16 * w));
  }
  const uint64_t o = (qi * A.nslices + slice) * (uint64_t)A.k;
#pragma unroll
  for (int t = 0; t < S; t++) {
    const int idx = lane * S + t;
    if (idx < size) {
      GM(A.part_d64)[o + idx] = L.d[t];
      GM(A.part_k)[o + idx] = L.key[t];
    }
  }
  if (lane == 0) GM(A.part_n)[qi * A.nslices + slice] = (uint32_t)size;
}

template <int S>
__global__ void __launch_bounds__(256) brute_merge_f64_kernel(const BruteArgs A) {
  const int lane = lane_id();
  const uint64_t qi = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (qi >= A.nq) return;
  if (G(A.status)[qi] != 0) {
    if (lane == 0) GM(A.out_count)[qi] = 0;
    return;
  }
  KList<double, S> L;
#pragma unroll
  for (int t = 0; t < S; t++) {
    L.d[t] = 0.0;
    L.key[t] = 0;
  }
  int size = 0;
  double kth_d = 0.0;
  uint64_t kth_k = 0;
  const int kk = (int)A.k;
  for (uint32_t sl = 0; sl < A.nslices; sl++) {
    const uint64_t o = (qi * A.nslices + sl) * (uint64_t)A.k;
    const int cnt = (int)G(A.part_n)[qi * A.nslices + sl];
    for (int i = 0; i < cnt; i++) {
      const double d = G(A.part_d64)[o + i];
      const uint64_t key = G(A.part_k)[o + i];
      if (!(size < kk || kless<double>(d, key, kth_d, kth_k))) break;
      klist_offer<double, S>(L, size, kk, kth_d, kth_k, d, key);
    }
  }
  const uint64_t ob = qi * (uint64_t)A.k_stride;
#pragma unroll
  for (int t = 0; t < S; t++) {
    const int idx = lane * S + t;
    if (idx < size) {
      A.out_keys[ob + idx] = L.key[t];
      A.out_dist64[ob + idx] = L.d[t];
    }
  }
  if (lane == 0) A.out_count[qi] = (uint32_t)size;
}

#define HH_BRUTE_INST(TM, EX, QB)                                                     \
  template __global__ void brute_partial_kernel<TM, EX, QB, 1>(const BruteArgs);     \
  template __global__ void brute_partial_kernel<TM, EX, QB, 2>(const BruteArgs);     \
  template __global__ void brute_partial_kernel<TM, EX, QB, 4>(const BruteArgs);     \
  template __global__ void brute_partial_kernel<TM, EX, QB, 8>(const BruteArgs);
HH_BRUTE_INST(4, false, 4)
HH_BRUTE_INST(4, true, 4)
HH_BRUTE_INST(8, false, 2)
HH_BRUTE_INST(8, true, 2)
HH_BRUTE_INST(32, false, 1)
HH_BRUTE_INST(32, true, 1)
template __global__ void brute_merge_kernel<1>(const BruteArgs);
template __global__ void brute_merge_kernel<2>(const BruteArgs);
template __global__ void brute_merge_kernel<4>(const BruteArgs);
template __global__ void brute_merge_kernel<8>(const BruteArgs);
template __global__ void brute_f64_kernel<1>(const BruteArgs);
template __global__ void brute_f64_kernel<2>(const BruteArgs);
template __global__ void brute_f64_kernel<4>(const BruteArgs);
template __global__ void brute_f64_kernel<8>(const BruteArgs);
template __global__ void brute_merge_f64_kernel<1>(const BruteArgs);
template __global__ void brute_merge_f64_kernel<2>(const BruteArgs);
template __global__ void brute_merge_f64_kernel<4>(const BruteArgs);
template __global__ void brute_merge_f64_kernel<8>(const BruteArgs);

}  // namespace hhk
