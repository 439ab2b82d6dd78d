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
// Packs rows x dim code bytes from rows of cstride bytes into dst (rows x
// dim, dense): one 64-thread block per row, any grid.
__global__ void pack_codes_kernel(const int8_t *src, uint32_t cstride, uint32_t dim,
                                  uint64_t rows, int8_t *dst);

}  // namespace hhk
