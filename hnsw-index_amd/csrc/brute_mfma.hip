// brute_mfma.hip — MFMA brute force with a fused candidate filter and an
// exact re-rank, returning exactly BruteForceIndex's results
// (src/brute.rs:131-160: top-k by (distance total order, key), distances
// bit-identical to the AVX2 kernels).
//
//   1. sample pass: approximate distances d' of every query to a strided
//      sample of the rows (v_mfma_f32_32x32x2f32, dense output), and per
//      query the r-th smallest sample value as a threshold tau;
//   2. main pass: the same MFMA tile over all rows; every (query, row) with
//      d' < tau is appended to the query's candidate buffer;
//   3. exact pass: candidates re-scored in the crate's AVX2 order, top-k by
//      (distance, key); the result is exact when tau > d_k + E, E bounding
//      |d' - d_exact| (FP32 GEMM error plus the AVX2 rounding); queries that
//      fail the test or overflow their buffer are flagged and re-run by the
//      exact streaming kernel.
//
// Tile: a wavefront holds 64 queries as MFMA A operands (2 tiles of 32,
// all of D in registers, D <= 128) and streams 32-row B tiles.  The K index
// is permuted so that each lane's B operand comes from one float4 per 4
// MFMA steps: step s = 4g + t uses element 8g + 4h + t, h = lane / 32; A and
// B use the same permutation, so the dot product is unchanged.
#include <hip/hip_runtime.h>

#include "brute_kernels.h"
#include "device_common.h"

using namespace hhd;

namespace hhk {

typedef float floatx16 __attribute__((ext_vector_type(16)));

// Row of the 32x32 f32 MFMA result held by lane l in accumulator v (the
// column is l % 32).
__device__ __forceinline__ int mfma_row(int v, int l) { return 8 * (v >> 2) + 4 * (l >> 5) + (v & 3); }

template <int KS>  // K steps = stride / 2 (stride <= 128)
__global__ void __launch_bounds__(256) brute_mfma_kernel(const MfmaArgs A) {
  const int lane = lane_id(), j = lane & 31, h = lane >> 5;
  const int wave = threadIdx.x >> 6;
  const uint64_t q0 = (uint64_t)blockIdx.x * 64;
  // A operands: queries q0 + qt*32 + j, elements 8g + 4h + t
  float a[2][KS];
  float tau_r[2][16];
#pragma unroll
  for (int qt = 0; qt < 2; qt++) {
    const uint64_t qi = q0 + qt * 32 + j;
    const bool ok = qi < A.nq;
    gptr<float4> qp = (gptr<float4>)(A.qprep + (ok ? qi : 0) * A.stride);
#pragma unroll
    for (int g = 0; g < KS / 4; g++) {
      float4 v = ok ? (float4)qp[2 * g + h] : make_float4(0.f, 0.f, 0.f, 0.f);
      a[qt][4 * g + 0] = v.x;
      a[qt][4 * g + 1] = v.y;
      a[qt][4 * g + 2] = v.z;
      a[qt][4 * g + 3] = v.w;
    }
#pragma unroll
    for (int v = 0; v < 16; v++) {
      const uint64_t qm = q0 + qt * 32 + mfma_row(v, lane);
      // rows of invalid queries (and past nq) never pass: -inf
      float t = -__builtin_inff();
      if (A.mode == 1 && qm < A.nq && G(A.status)[qm] == 0) {
        t = G(A.tau)[qm];
        if (A.metric == 0) t -= G(A.qq)[qm];  // compare xx - 2 dot < tau - qq
      }
      tau_r[qt][v] = t;
    }
  }
  const uint64_t tiles = (A.nrows + 31) / 32;
  for (uint64_t tile = (uint64_t)blockIdx.y * 4 + wave; tile < tiles; tile += (uint64_t)gridDim.y * 4) {
    const uint64_t rl_ = tile * 32 + j;  // row index within this pass
    const bool rv = rl_ < A.nrows;
    const uint64_t row = rv ? rl_ * A.row_step : 0;
    gptr<float4> xp = (gptr<float4>)(A.rows + row * A.stride);
    float4 b[KS / 4];
#pragma unroll
    for (int g = 0; g < KS / 4; g++) b[g] = rv ? (float4)xp[2 * g + h] : make_float4(0.f, 0.f, 0.f, 0.f);
    floatx16 acc0 = {0}, acc1 = {0};
#pragma unroll
    for (int g = 0; g < KS / 4; g++) {
      acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a[0][4 * g + 0], b[g].x, acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a[1][4 * g + 0], b[g].x, acc1, 0, 0, 0);
      acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a[0][4 * g + 1], b[g].y, acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a[1][4 * g + 1], b[g].y, acc1, 0, 0, 0);
      acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a[0][4 * g + 2], b[g].z, acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a[1][4 * g + 2], b[g].z, acc1, 0, 0, 0);
      acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a[0][4 * g + 3], b[g].w, acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a[1][4 * g + 3], b[g].w, acc1, 0, 0, 0);
    }
    if (!rv) continue;
    const float xx = A.metric == 0 ? G(A.xx)[row] : 0.0f;
#pragma unroll
    for (int qt = 0; qt < 2; qt++) {
#pragma unroll
      for (int v = 0; v < 16; v++) {
        const float dot = qt == 0 ? acc0[v] : acc1[v];
        // d' without the query's constant: L2 xx - 2 dot, IP -dot, cosine 1 - dot
        const float dp = A.metric == 0 ? xx - 2.0f * dot : (A.metric == 1 ? -dot : 1.0f - dot);
        const uint64_t qm = q0 + qt * 32 + mfma_row(v, lane);
        if (A.mode == 0) {
          // every row of the sample is written, so the threshold pass
          // never reads an unwritten value
          if (qm < A.nq) GM(A.dense)[qm * A.nrows + rl_] = A.metric == 0 ? dp + G(A.qq)[qm] : dp;
        } else if (dp < tau_r[qt][v]) {
          const uint32_t slot = atomicAdd(A.cand_count + qm, 1u);
          if (slot < A.cand_cap) GM(A.cand)[qm * A.cand_cap + slot] = (uint32_t)row;
        }
      }
    }
  }
}

// tau[q] = r-th smallest of dense[q][0..nrows) (one wave per query).
__global__ void __launch_bounds__(256) brute_tau_kernel(const MfmaArgs A) {
  const int lane = lane_id();
  const uint64_t qi = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (qi >= A.nq) return;
  if (G(A.status)[qi] != 0) return;
  constexpr int R = 8;
  float best[R];
#pragma unroll
  for (int i = 0; i < R; i++) best[i] = __builtin_inff();
  gptr<float> row = G(A.dense) + qi * A.nrows;
  for (uint64_t s = lane; s < A.nrows; s += 64) {
    float v = row[s];
#pragma unroll
    for (int i = 0; i < R; i++) {
      const float lo = fminf(best[i], v);
      v = fmaxf(best[i], v);
      best[i] = lo;
    }
  }
  // r-th smallest over the wave: pop the wave minimum r times
  int head = 0;
  float kth = __builtin_inff();
  for (int it = 0; it < R; it++) {
    float mine = __builtin_inff();
#pragma unroll
    for (int i = 0; i < R; i++)
      if (i == head) mine = best[i];
    float m = mine;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) m = fminf(m, __shfl_xor(m, o, 64));
    kth = m;
    const uint64_t owner = __ballot(mine == m && m < __builtin_inff());
    if (owner) {
      const int first = __ffsll((unsigned long long)owner) - 1;
      head += lane == first ? 1 : 0;
    }
  }
  if (lane == 0) GM(A.tau)[qi] = kth;
}

// Exact re-rank of each query's candidates; flags queries whose candidate
// set cannot be shown to contain the exact top-k.
template <int TMAX, bool EXTRA, int S>
__global__ void __launch_bounds__(256) brute_rerank_kernel(const MfmaArgs A, const BruteArgs B) {
  const int lane = lane_id(), i = lane & 15, quarter = lane >> 4;
  const uint64_t qi = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (qi >= A.nq) return;
  if (G(A.status)[qi] != 0) {
    if (lane == 0) {
      GM(B.out_count)[qi] = 0;
      GM(A.redo)[qi] = 0;
    }
    return;
  }
  const uint32_t cnt = G(A.cand_count)[qi];
  const int kk = (int)B.k;
  bool redo = cnt > A.cand_cap || cnt < (uint32_t)kk;
  const Shape s{B.T, B.U, B.R, B.stride};
  QLanes<TMAX, EXTRA> q;
#pragma unroll
  for (int t = 0; t < TMAX; t++) q.lo[t] = q.hi[t] = 0.0f;
  q.e[0] = q.r = 0.0f;
  if constexpr (EXTRA) q.e[1] = q.e[2] = 0.0f;
  load_q<TMAX, EXTRA, float>(B.qprep + qi * B.stride, s, i, q);
  float kd[S];
  uint64_t kk_[S];
#pragma unroll
  for (int t = 0; t < S; t++) {
    kd[t] = 0.0f;
    kk_[t] = 0;
  }
  int size = 0;
  float kth_d = 0.0f;
  uint64_t kth_k = 0;
  if (!redo) {
    gptr<uint32_t> cand = G(A.cand) + qi * A.cand_cap;
    for (uint32_t base = 0; base < cnt; base += 4) {
      const uint32_t c = base + quarter;
      const bool cv = c < cnt;
      const uint32_t row = cv ? cand[c] : 0;
      QLanes<TMAX, EXTRA> x;
      float lo = 0.0f, hi = 0.0f, tail = 0.0f;
      uint64_t key = 0;
      if (cv) {
        load_q<TMAX, EXTRA, float>(B.rows + (uint64_t)row * B.stride, s, i, x);
        if (B.metric == 0) quarter_acc<TMAX, EXTRA, true>(q, x, s, i, lo, hi, tail);
        else quarter_acc<TMAX, EXTRA, false>(q, x, s, i, lo, hi, tail);
        key = G(B.keys)[row];
      }
      const float d = finish_metric(quarter_reduce<EXTRA>(lo, hi, tail, s.R, quarter), B.metric);
#pragma unroll
      for (int w = 0; w < 4; w++) {
        if (!rl((uint32_t)cv, 16 * w)) continue;
        const float dw = rlf(d, 16 * w);
        const uint64_t kw = rl64(key, 16 * w);
        if (size < kk || nless(dw, kw, kth_d, kth_k)) {
          // sorted insert into the distributed list
          int p = 0;
#pragma unroll
          for (int t = 0; t < S; t++) p += popc64(__ballot((lane * S + t) < size && nless(kd[t], kk_[t], dw, kw)));
          const float upd = __shfl_up(kd[S - 1], 1, 64);
          const uint64_t upk = (uint64_t)__shfl_up((long long)kk_[S - 1], 1, 64);
#pragma unroll
          for (int t = S - 1; t >= 0; t--) {
            const int idx = lane * S + t;
            const float pd = t > 0 ? kd[t > 0 ? t - 1 : 0] : upd;
            const uint64_t pk = t > 0 ? kk_[t > 0 ? t - 1 : 0] : upk;
            if (idx > p) {
              kd[t] = pd;
              kk_[t] = pk;
            } else if (idx == p) {
              kd[t] = dw;
              kk_[t] = kw;
            }
          }
          if (size < kk) size++;
          const int ln = (size - 1) / S, sl = (size - 1) % S;
          float vd = kd[0];
          uint64_t vk = kk_[0];
#pragma unroll
          for (int t = 1; t < S; t++)
            if (sl == t) {
              vd = kd[t];
              vk = kk_[t];
            }
          kth_d = rlf(vd, ln);
          kth_k = rl64(vk, ln);
        }
      }
    }
    // exactness: every row outside the candidates has d' >= tau, hence an
    // exact distance >= tau - E > d_k
    const float tau = G(A.tau)[qi];
    double qq = 0.0;
    for (int t = 0; t < TMAX; t++) qq += (double)q.lo[t] * q.lo[t] + (double)q.hi[t] * q.hi[t];
    if constexpr (EXTRA) {
      for (int u = 0; u < 3; u++) qq += (double)q.e[u] * q.e[u];
      qq += (double)q.r * q.r;
    }
    for (int m = 1; m < 64; m <<= 1) qq += __shfl_xor(qq, m, 64);
    qq = qq * 0.25;  // every value is held by the 4 quarters
    const double qn = sqrt(qq), M = A.max_norm;
    const double gam = 4.0 * (A.dim + 3) * 5.960464477539063e-08;
    const double E = B.metric == 0 ? gam * (qn + M) * (qn + M) : gam * qn * M;
    if (size < kk || !((double)tau > (double)kth_d + E)) redo = true;
  }
  const uint64_t ob = qi * (uint64_t)B.k_stride;
  if (!redo) {
#pragma unroll
    for (int t = 0; t < S; t++) {
      const int idx = lane * S + t;
      if (idx < size) {
        GM(B.out_keys)[ob + idx] = kk_[t];
        GM(B.out_dist)[ob + idx] = kd[t];
      }
    }
    if (lane == 0) GM(B.out_count)[qi] = (uint32_t)size;
  }
  if (lane == 0) GM(A.redo)[qi] = redo ? 1u : 0u;
}

// Per-row squared norms (f32) for the L2 form of the GEMM distance.
__global__ void row_norms_kernel(const float *rows, uint64_t n, uint32_t stride, float *xx) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  gptr<float> p = G(rows) + r * stride;
  float s = 0.0f;
  for (uint32_t e = 0; e < stride; e++) s = __builtin_fmaf(p[e], p[e], s);
  GM(xx)[r] = s;
}

// Squared norms of the prepared queries (for the L2 GEMM distance).
__global__ void query_norms_kernel(const float *qprep, uint64_t nq, uint32_t stride, float *qq) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nq) return;
  gptr<float> p = G(qprep) + r * stride;
  float s = 0.0f;
  for (uint32_t e = 0; e < stride; e++) s = __builtin_fmaf(p[e], p[e], s);
  GM(qq)[r] = s;
}

__global__ void fill_u32_kernel(uint32_t *p, uint64_t n, uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    GM(p)[i] = v;
}

template __global__ void brute_mfma_kernel<8>(const MfmaArgs);
template __global__ void brute_mfma_kernel<16>(const MfmaArgs);
template __global__ void brute_mfma_kernel<24>(const MfmaArgs);
template __global__ void brute_mfma_kernel<32>(const MfmaArgs);
template __global__ void brute_mfma_kernel<40>(const MfmaArgs);
template __global__ void brute_mfma_kernel<48>(const MfmaArgs);
template __global__ void brute_mfma_kernel<56>(const MfmaArgs);
template __global__ void brute_mfma_kernel<64>(const MfmaArgs);
#define HH_RR_INST(TM, EX)                                                          \
  template __global__ void brute_rerank_kernel<TM, EX, 1>(const MfmaArgs, const BruteArgs); \
  template __global__ void brute_rerank_kernel<TM, EX, 2>(const MfmaArgs, const BruteArgs); \
  template __global__ void brute_rerank_kernel<TM, EX, 4>(const MfmaArgs, const BruteArgs); \
  template __global__ void brute_rerank_kernel<TM, EX, 8>(const MfmaArgs, const BruteArgs);
HH_RR_INST(4, false)
HH_RR_INST(4, true)

}  // namespace hhk
