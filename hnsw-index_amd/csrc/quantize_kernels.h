// Kernels of the device scalar quantizer (src/quantization.rs:32-81).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hhk {

struct SqArgs {
  const float *rows;   // n x stride, accumulator-major layout
  uint64_t n;
  uint32_t stride;     // floats per row (a multiple of 16)
  uint32_t dim;
  uint32_t T;          // full 32-value blocks of the layout
  uint32_t cstride;    // code bytes per row (a multiple of 128)
  float *offset;       // dim offsets, natural order
  float *offset_pos;   // stride offsets in row-position order (0 on padding)
  uint32_t *max_bits;  // bit pattern of max |x - offset|, cleared by the caller
  int8_t *codes;       // n x cstride, zero-filled by the caller
};

// grid: stride / 16 blocks of 256 threads; needs n >= 1
__global__ void sq_mean_kernel(SqArgs a);
// any grid of 256-thread blocks
__global__ void sq_max_kernel(SqArgs a);
__global__ void sq_encode_kernel(SqArgs a);
// Per-dimension min/max int8 codes (the engine's own quantizer; no crate
// counterpart).  Per dimension d: lo_d, hi_d the minimum and maximum over
// all rows, scale_d = (hi_d - lo_d) / 254 (1 when not positive), mid_d =
// (lo_d + hi_d) * 0.5 + 0 (a zero sum gives +0), all f32; code =
// clamp(round_half_away((x - mid_d) * (1 / scale_d)), -127, 127); the row
// term of the L2 expansion rnorm = sum_d (scale_d * code_d)^2 in f64 in
// dimension order, rounded to f32 once.
struct PdArgs {
  const float *rows;   // accumulator-major layout, stride floats per row
  uint64_t first, n;   // rows [first, n) are encoded
  uint32_t stride, dim, T, cstride, metric;
  int32_t *lo_key, *hi_key;  // stride total-order keys per position
  float *mid, *scale;        // dim each, natural order
  int8_t *codes;             // rows x cstride, padding written as zero
  float *rnorm;              // rows (L2), or null
};
// lo_key / hi_key must hold INT_MAX / INT_MIN; any grid of 256-thread blocks
__global__ void pd_minmax_kernel(PdArgs a);
// one thread per dimension
__global__ void pd_params_kernel(PdArgs a);
// one wavefront per row, any grid of 256-thread blocks
__global__ void pd_encode_kernel(PdArgs a);

// Packs rows x dim code bytes from rows of cstride bytes into dst (rows x
// dim, dense): one 64-thread block per row, any grid.
__global__ void pack_codes_kernel(const int8_t *src, uint32_t cstride, uint32_t dim,
                                  uint64_t rows, int8_t *dst);

}  // namespace hhk
