// quantize.hip — the crate's scalar quantizer trained and applied on the
// GPU over the resident rows (ScalarQuantizer::train / encode_into,
// src/quantization.rs:32-81), bit-identical to the host quantizer:
//
//   offset[d] = (f32)(Σ_r (f64)x[r][d] / rows)   L2 only, the sum taken over
//                                                the rows in order, in f64;
//                                                0 for the other metrics
//   max_abs   = max over every x of |x − offset| (f32)
//   scale     = max_abs / 127, or 1 when max_abs is 0 or not finite
//   code      = clamp(round_half_away((x − offset) · (1/scale)), ±127)
//
// The per-dimension sum keeps the crate's sequential f64 rounding: one lane
// per dimension adds its column in row order.  That chain of n dependent
// adds is the kernel's critical path, so the rows are spread over many
// blocks by position (16 row positions = 64 bytes of each row per block) and
// all 256 threads of a block load, widen to f64 and stage the next tile of
// rows while 16 lanes add the current one (one LDS read and one add per
// row on the chain).  The maximum is order-free (a grid-wide max over
// float bit patterns, which order like the non-negative values) and
// encoding is elementwise; both stream the rows once.
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "quantize_kernels.h"

using namespace hhd;

namespace hhk {

// Dimension stored at row position p: the inverse of the accumulator-major
// row layout (value 32t + j lives at (t/4)*128 + j*tw + t%4 with
// tw = min(4, T − 4(t/4)); from 32T on the values are in natural order and
// positions past the dimension are zero padding).
__device__ __forceinline__ uint32_t dim_at(uint32_t p, uint32_t T) {
  if (p >= 32u * T) return p;
  const uint32_t b = p >> 7, w = p & 127;
  const uint32_t tw = T - 4 * b < 4 ? T - 4 * b : 4;
  const uint32_t j = w / tw, t = 4 * b + w % tw;
  return 32 * t + j;
}

constexpr int kPos = 16;     // row positions per block of sq_mean_kernel
constexpr int kTile = 256;   // rows per staged tile (as f64: 32 KiB, two buffers)
constexpr int kLoads = kTile * (kPos / 4) / 256;  // float4 loads per thread per tile

// Load k of one tile of rows r0..: 4 threads per row, kTile rows in
// kLoads loads per thread; rows past the end read row n-1 (never summed:
// the sum stops at n; n >= 1).
__device__ __forceinline__ float4 sq_tile_load(const SqArgs &a, uint32_t p0, uint32_t t,
                                               uint64_t r0, int k) {
  const uint32_t i = t + 256 * k, r = i >> 2, c = i & 3;
  const uint64_t row = r0 + r < a.n ? r0 + r : a.n - 1;
  return *(gptr<float4>)(a.rows + row * a.stride + p0 + 4 * c);
}

// The loading threads also widen to f64, so the summing lanes issue one
// LDS read and one add per row.
__device__ __forceinline__ void sq_tile_store(double *buf, uint32_t t, int k, float4 x) {
  const uint32_t i = t + 256 * k, r = i >> 2, c = i & 3;
  double2 *dst = reinterpret_cast<double2 *>(buf + r * kPos + 4 * c);
  dst[0] = make_double2((double)x.x, (double)x.y);
  dst[1] = make_double2((double)x.z, (double)x.w);
}

__global__ void __launch_bounds__(256) sq_mean_kernel(SqArgs a) {
  static_assert(kLoads == 4, "one register per load below");
  __shared__ double tile[2][kTile * kPos];
  const uint32_t p0 = blockIdx.x * kPos;
  const uint32_t t = threadIdx.x;
  float4 x0, x1, x2, x3;
#define HH_SQ_FETCH(R0)               \
  x0 = sq_tile_load(a, p0, t, R0, 0); \
  x1 = sq_tile_load(a, p0, t, R0, 1); \
  x2 = sq_tile_load(a, p0, t, R0, 2); \
  x3 = sq_tile_load(a, p0, t, R0, 3);
#define HH_SQ_STORE(BUF)           \
  sq_tile_store(BUF, t, 0, x0);    \
  sq_tile_store(BUF, t, 1, x1);    \
  sq_tile_store(BUF, t, 2, x2);    \
  sq_tile_store(BUF, t, 3, x3);
  double sum = 0.0;
  HH_SQ_FETCH(0)
  HH_SQ_STORE(tile[0])
  if ((uint64_t)kTile < a.n) {
    HH_SQ_FETCH((uint64_t)kTile)
  }
  __syncthreads();
  // Iteration j: lanes 0..15 add tile j (buffer j&1) while every thread
  // stores tile j+1 (held in registers) into the other buffer and loads
  // tile j+2.  The buffer written in iteration j was last read in
  // iteration j-1, before the barrier that ends it.
  uint32_t j = 0;
  for (uint64_t r0 = 0; r0 < a.n; r0 += kTile, j++) {
    if (r0 + kTile < a.n) {
      HH_SQ_STORE(tile[(j + 1) & 1])
      if (r0 + 2 * kTile < a.n) {
        HH_SQ_FETCH(r0 + 2 * kTile)
      }
    }
    if (t < kPos) {
      const double *buf = tile[j & 1];
      const uint32_t nr = a.n - r0 < (uint64_t)kTile ? (uint32_t)(a.n - r0) : (uint32_t)kTile;
      uint32_t r = 0;
      // 16 reads in flight, then the 16 dependent adds of the chain
      for (; r + 16 <= nr; r += 16) {
        double v[16];
#pragma unroll
        for (int u = 0; u < 16; u++) v[u] = buf[(r + u) * kPos + t];
#pragma unroll
        for (int u = 0; u < 16; u++) sum = __dadd_rn(sum, v[u]);
      }
      for (; r < nr; r++) sum = __dadd_rn(sum, buf[r * kPos + t]);
    }
    __syncthreads();
  }
#undef HH_SQ_FETCH
#undef HH_SQ_STORE
  if (t < kPos) {
    const uint32_t p = p0 + t;
    const uint32_t d = dim_at(p, a.T);
    const float o = d < a.dim ? (float)(sum / (double)a.n) : 0.0f;
    GM(a.offset_pos)[p] = o;
    if (d < a.dim) GM(a.offset)[d] = o;
  }
}

// Row r, position group c (4 positions) of a row-chunked flat loop: blocks
// take chunks of kChunk rows so the index arithmetic stays in 32 bits.
constexpr uint32_t kChunk = 1024;

__global__ void __launch_bounds__(256) sq_max_kernel(SqArgs a) {
  __shared__ uint32_t red[4];
  const uint32_t s4 = a.stride / 4;
  uint32_t m = 0;
  for (uint64_t r0 = (uint64_t)blockIdx.x * kChunk; r0 < a.n; r0 += (uint64_t)gridDim.x * kChunk) {
    const uint32_t nr = a.n - r0 < kChunk ? (uint32_t)(a.n - r0) : kChunk;
    const uint32_t units = nr * s4;
    gptr<float4> base = (gptr<float4>)(a.rows + r0 * a.stride);
    for (uint32_t u = threadIdx.x; u < units; u += 256) {
      const uint32_t c = u % s4;
      const float4 x = base[u];
      const float4 o = *(gptr<float4>)(a.offset_pos + 4 * c);
      // padding positions hold 0 with offset 0: |0 - 0| leaves the max alone
      m = max(m, __float_as_uint(fabsf(__fsub_rn(x.x, o.x))));
      m = max(m, __float_as_uint(fabsf(__fsub_rn(x.y, o.y))));
      m = max(m, __float_as_uint(fabsf(__fsub_rn(x.z, o.z))));
      m = max(m, __float_as_uint(fabsf(__fsub_rn(x.w, o.w))));
    }
  }
  for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t b = max(max(red[0], red[1]), max(red[2], red[3]));
    atomicMax(a.max_bits, b);
  }
}

__device__ __forceinline__ float sq_scale_of(uint32_t max_bits) {
  const float max_abs = __uint_as_float(max_bits);
  return (max_abs > 0.0f && isfinite(max_abs)) ? max_abs / 127.0f : 1.0f;
}

__device__ __forceinline__ int8_t sq_code(float x, float o, float inv) {
  float c = roundf(__fmul_rn(__fsub_rn(x, o), inv));
  c = fminf(127.0f, fmaxf(-127.0f, c));
  return (int8_t)c;
}

// codes[r*cstride + d] for d < dim, natural order; the caller zero-fills
// the buffer first so the padding to cstride stays zero.
__global__ void __launch_bounds__(256) sq_encode_kernel(SqArgs a) {
  const float inv = 1.0f / sq_scale_of(*G(a.max_bits));
  const uint32_t s4 = a.stride / 4;
  for (uint64_t r0 = (uint64_t)blockIdx.x * kChunk; r0 < a.n; r0 += (uint64_t)gridDim.x * kChunk) {
    const uint32_t nr = a.n - r0 < kChunk ? (uint32_t)(a.n - r0) : kChunk;
    const uint32_t units = nr * s4;
    gptr<float4> base = (gptr<float4>)(a.rows + r0 * a.stride);
    for (uint32_t u = threadIdx.x; u < units; u += 256) {
      const uint32_t r = u / s4, c = u - r * s4;
      const float4 x = base[u];
      const float4 o = *(gptr<float4>)(a.offset_pos + 4 * c);
      const float xs[4] = {x.x, x.y, x.z, x.w}, os[4] = {o.x, o.y, o.z, o.w};
      int8_t *row = a.codes + (r0 + r) * a.cstride;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const uint32_t d = dim_at(4 * c + k, a.T);
        if (d < a.dim) GM(row)[d] = sq_code(xs[k], os[k], inv);
      }
    }
  }
}

// ------------------------------------------- per-dimension min/max codes

// Monotone int key of the f32 total order (-0 below +0).
__device__ __forceinline__ int32_t pd_key(float f) {
  const int32_t i = __float_as_int(f);
  return i ^ (int32_t)(((uint32_t)(i >> 31)) >> 1);
}
__device__ __forceinline__ float pd_unkey(int32_t k) {
  return __int_as_float(k ^ (int32_t)(((uint32_t)(k >> 31)) >> 1));
}
// Row position of dimension d (the inverse of dim_at).
__device__ __forceinline__ uint32_t pos_of(uint32_t d, uint32_t T) {
  if (d >= 32u * T) return d;
  const uint32_t t = d >> 5, j = d & 31, b = t >> 2;
  const uint32_t tw = T - 4 * b < 4 ? T - 4 * b : 4;
  return b * 128 + j * tw + (t & 3);
}

__global__ void __launch_bounds__(256) pd_minmax_kernel(PdArgs a) {
  // thread t of a block owns one position (group) over a chunk of rows
  const uint32_t s4 = a.stride / 4;
  for (uint64_t r0 = (uint64_t)blockIdx.x * kChunk; r0 < a.n; r0 += (uint64_t)gridDim.x * kChunk) {
    const uint32_t nr = a.n - r0 < kChunk ? (uint32_t)(a.n - r0) : kChunk;
    gptr<float4> base = (gptr<float4>)(a.rows + r0 * a.stride);
    for (uint32_t c = threadIdx.x; c < s4; c += 256) {
      int32_t lo[4] = {INT32_MAX, INT32_MAX, INT32_MAX, INT32_MAX};
      int32_t hi[4] = {INT32_MIN, INT32_MIN, INT32_MIN, INT32_MIN};
      for (uint32_t r = 0; r < nr; r++) {
        const float4 x = base[(size_t)r * s4 + c];
        const int32_t k[4] = {pd_key(x.x), pd_key(x.y), pd_key(x.z), pd_key(x.w)};
#pragma unroll
        for (int u = 0; u < 4; u++) {
          lo[u] = min(lo[u], k[u]);
          hi[u] = max(hi[u], k[u]);
        }
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        atomicMin(a.lo_key + 4 * c + u, lo[u]);
        atomicMax(a.hi_key + 4 * c + u, hi[u]);
      }
    }
  }
}

__global__ void __launch_bounds__(256) pd_params_kernel(PdArgs a) {
  const uint32_t d = blockIdx.x * 256 + threadIdx.x;
  if (d >= a.dim) return;
  const uint32_t p = pos_of(d, a.T);
  const float lo = pd_unkey(G(a.lo_key)[p]), hi = pd_unkey(G(a.hi_key)[p]);
  const float s = __fdiv_rn(__fsub_rn(hi, lo), 254.0f);
  GM(a.scale)[d] = s > 0.0f ? s : 1.0f;
  GM(a.mid)[d] = __fadd_rn(__fmul_rn(__fadd_rn(lo, hi), 0.5f), 0.0f);
}

// One wavefront per row: lane l takes dimensions l, l + 64, ...; the code
// row is written whole (padding zero); the L2 row term sums the lanes'
// f64 squares in dimension order by a uniform chain.
__global__ void __launch_bounds__(256) pd_encode_kernel(PdArgs a) {
  const int lane = (int)(threadIdx.x & 63);
  const uint64_t waves = (uint64_t)gridDim.x * 4;
  for (uint64_t r = a.first + (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6); r < a.n; r += waves) {
    gptr<float> row = G(a.rows) + r * a.stride;
    gmut<int8_t> out = GM(a.codes) + r * a.cstride;
    double acc = 0.0;
    for (uint32_t d0 = 0; d0 < a.cstride; d0 += 64) {
      const uint32_t d = d0 + (uint32_t)lane;
      int8_t c = 0;
      if (d < a.dim) {
        const float x = row[pos_of(d, a.T)];
        const float inv = __fdiv_rn(1.0f, G(a.scale)[d]);
        float v = roundf(__fmul_rn(__fsub_rn(x, G(a.mid)[d]), inv));
        v = fminf(127.0f, fmaxf(-127.0f, v));
        c = (int8_t)v;
      }
      out[d] = c;
      if (a.rnorm && d0 < a.dim) {
        const double y = d < a.dim ? __dmul_rn((double)G(a.scale)[d], (double)c) : 0.0;
        const double sq = __dmul_rn(y, y);
        const int cnt = a.dim - d0 < 64 ? (int)(a.dim - d0) : 64;
        const int lo = __double2loint(sq), hi = __double2hiint(sq);
        for (int l = 0; l < cnt; l++)
          acc = __dadd_rn(acc, __hiloint2double(__builtin_amdgcn_readlane(hi, l),
                                                __builtin_amdgcn_readlane(lo, l)));
      }
    }
    if (a.rnorm) {
      const float rn = (float)acc;
      if (lane == 0) GM(a.rnorm)[r] = rn;
    }
  }
}

__global__ void __launch_bounds__(64) pack_codes_kernel(const int8_t *src, uint32_t cstride,
                                                        uint32_t dim, uint64_t rows, int8_t *dst) {
  for (uint64_t r = blockIdx.x; r < rows; r += gridDim.x)
    for (uint32_t c = threadIdx.x; c < dim; c += 64) dst[r * dim + c] = src[r * cstride + c];
}

}  // namespace hhk
