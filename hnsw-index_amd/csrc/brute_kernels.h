// brute_kernels.h — argument block and template instances of the
// brute-force kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hhk {

struct BruteArgs {
  const float *rows;        // n x stride, accumulator-major
  const uint64_t *keys;     // n
  const uint32_t *allow;    // bitmap over rows, or null
  uint64_t n;
  uint32_t dim, metric;
  int T, U, R;
  uint32_t stride;
  const float *qprep;
  const int32_t *status;
  uint64_t nq;
  uint32_t k;               // clamped to n
  uint32_t k_stride;
  uint32_t nslices;
  uint64_t rows_per_slice;  // multiple of 4
  float *part_d;            // nq x nslices x k
  double *part_d64;
  uint64_t *part_k;
  uint32_t *part_n;         // nq x nslices
  uint64_t *out_keys;
  float *out_dist;
  double *out_dist64;
  uint32_t *out_count;
  const uint32_t *only;     // if set, queries with only[q] == 0 are skipped entirely
};

// MFMA shortlist passes (brute_mfma.hip).
struct MfmaArgs {
  const float *rows;
  uint64_t n;
  uint32_t stride, dim, metric;
  const float *qprep;
  const int32_t *status;
  uint64_t nq;
  const float *xx;          // per-row squared norms (L2)
  const float *qq;          // per-query squared norms (L2)
  uint32_t mode;            // 0: dense sample distances, 1: candidate filter
  uint64_t nrows, row_step; // rows of this pass: r * row_step, r < nrows
  float *dense;             // mode 0: nq x nrows
  float *tau;               // per-query threshold
  uint32_t *cand;           // mode 1: nq x cand_cap row ids
  uint32_t *cand_count;
  uint32_t cand_cap;
  double max_norm;          // largest row norm (error bound)
  uint32_t *redo;           // per query: 1 = re-run exactly
};

template <int TMAX, bool EXTRA, int QB, int S>
__global__ void brute_partial_kernel(const BruteArgs A);
template <int S>
__global__ void brute_merge_kernel(const BruteArgs A);
template <int S>
__global__ void brute_f64_kernel(const BruteArgs A);
template <int S>
__global__ void brute_merge_f64_kernel(const BruteArgs A);
template <int KS>
__global__ void brute_mfma_kernel(const MfmaArgs A);
__global__ void brute_tau_kernel(const MfmaArgs A);
template <int TMAX, bool EXTRA, int S>
__global__ void brute_rerank_kernel(const MfmaArgs A, const BruteArgs B);
__global__ void row_norms_kernel(const float *rows, uint64_t n, uint32_t stride, float *xx);
__global__ void query_norms_kernel(const float *qprep, uint64_t nq, uint32_t stride, float *qq);
__global__ void fill_u32_kernel(uint32_t *p, uint64_t n, uint32_t v);

}  // namespace hhk
