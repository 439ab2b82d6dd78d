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
};

template <int TMAX, bool EXTRA, int QB, int S>
__global__ void brute_partial_kernel(const BruteArgs A);
template <int S>
__global__ void brute_merge_kernel(const BruteArgs A);
template <int S>
__global__ void brute_f64_kernel(const BruteArgs A);
template <int S>
__global__ void brute_merge_f64_kernel(const BruteArgs A);

}  // namespace hhk
