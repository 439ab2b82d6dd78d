// search_kernels.h — argument blocks of the gfx950 kernels and their
// template instances (shared by the kernel and host translation units).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/hnsw_amd.h"

namespace hhk {

struct SearchArgs {
  // index (HBM)
  const float *rows;            // n x stride, accumulator-major f32 rows
  const uint16_t *bf16;         // n x stride bf16 rows, or null
  const int8_t *codes;          // n x cstride int8 codes (natural order), or null
  const uint32_t *adj0;         // n x m0 layer-0 ids, unused slots 0xFFFFFFFF
  const uint64_t *upper_off;    // n
  const uint32_t *upper;        // crate layout: count then m ids per layer
  const uint64_t *keys;         // n
  const uint32_t *deleted_bits; // ceil(n/32) words, or null
  uint32_t n, m, m0, metric;
  uint32_t entry;
  int32_t top;
  uint64_t live;
  int T, U, R;
  uint32_t stride, cstride;
  float sq_scale;
  const int8_t *pd_codes;       // n x cstride per-dimension codes, or null
  const float *pd_rnorm;        // n row terms of the L2 expansion (L2 only), or null
  // queries (prepared by prepare_queries_kernel)
  const float *qprep;
  const int8_t *qcode;          // crate codes (precision 1) or per-dimension weights (3)
  const float2 *qpd;            // per query: (sigma, qconst) of precision 3
  const int32_t *status;
  uint64_t nq;
  const hh_query_spec *specs;   // null = uniform k/ef/precision/allow
  uint32_t k, ef, precision;
  const uint32_t *allow;
  const uint32_t *filters;
  uint32_t filter_words;
  uint32_t k_stride;
  // outputs
  uint64_t *out_keys;
  float *out_dist;
  uint32_t *out_count;
  // workspace
  uint32_t *counter;
  uint32_t *spill;              // per-wave visited bitsets (spill_words each)
  uint64_t spill_words;
  uint64_t *cpool;              // per-wave candidate pools (cpool_cap each)
  uint32_t cpool_cap;
  uint32_t *spill_log;          // per-wave ids marked in the bitset (spill_log_cap each)
  uint32_t spill_log_cap;
  unsigned long long *stats;    // kStatLines lines of kStatStride counters (9 used)
  // Queries whose candidate pool dropped an entry are listed here (count in
  // redo_count) and run again by a phase-3 launch whose per-wave pools hold
  // cpool_cap >= n entries, so no query returns other results than the
  // crate's unbounded heap gives.  Null: no list (the pool cannot overflow).
  uint32_t *redo;
  uint32_t *redo_count;
  // Two-pass launches (phase 1 then 2) take long queries first: phase 0
  // runs each query whole; phase 1 only descends the upper layers, writing
  // seed[qi] = (best distance bits << 32) | layer-0 entry; phase 2 runs
  // layer 0 from seed[], taking queries in the order of `order` (largest
  // descent distance first); phase 3 re-runs layer 0 of the queries in
  // redo[] from their seeds.
  uint32_t phase;
  uint64_t *seed;
  const uint32_t *order;
  // construction candidates: node ids instead of keys in out_keys
  uint32_t ids_out;
  uint32_t hash_cap, hash_limit, lds_words;
  // id/distance staging entries per wave (64, or a beam step's slots
  // rounded up to 64, at most 256)
  uint32_t st_cap;
  // candidates a layer-0 step of the bf16 and per-dimension int8
  // traversals expands (1: the crate's one-at-a-time rule; 2..8 with
  // beam * m0 <= 256: one batch per step, every new neighbour tested against
  // the results as they stood at the step's start)
  uint32_t beam;
};

struct PrepArgs {
  const float *queries;
  uint64_t nq;
  uint32_t dim, metric;
  int T;
  uint32_t stride, cstride;
  float *qprep;
  int8_t *qcode;
  const float *sq_offset;
  float sq_scale;
  const float *pd_mid, *pd_scale;  // per-dimension quantizer, or null
  float2 *qpd;
  uint32_t precision;
  const hh_query_spec *specs;
  uint32_t k_stride, nfilters, ef_cap;
  int has_codes, has_bf16, has_pd;
  int32_t *status;
  uint32_t *counter;
};

constexpr uint32_t kOrderBins = 4096;

// The search counters are kept in kStatLines copies, one 128-byte line
// each, summed when read: a wave adds into line (its global index mod
// kStatLines), so the atomics of a draining launch do not all meet on one
// cache line.
constexpr uint32_t kStatStride = 16, kStatLines = 64;

// Precision classes of the search kernel instances: bf16 rows and the
// per-dimension int8 codes (beam steps), every precision (mixed batches,
// construction, beam width 1), f32 rows alone, or the crate's int8 codes
// alone.
constexpr int kPrecBf16 = 1, kPrecCodes = 2, kPrecAny = 3, kPrecF32 = 4, kPrecI8 = 5;
// Instances that traverse on int8 code rows only.
constexpr bool codes_class(int pc) { return pc == kPrecCodes || pc == kPrecI8; }
// Instances launched only with beam steps (the bf16 and per-dimension
// traversals at beam width 2 or more): the crate-rule loop is not compiled.
constexpr bool beam_class(int pc) { return pc == kPrecBf16 || pc == kPrecCodes; }

template <int TMAX, bool EXTRA, bool EXACT, int SW, int PC>
__global__ void hnsw_search_kernel(const SearchArgs P);
// First pass of a two-pass launch for m <= LPQ (16 or 32 lanes per query):
// descent_threads / LPQ queries per block (grid ceil(nq / that)), each
// descent's seed written as by phase 1.  EXACT: D = 32 * TMAX.
// Per query slot it holds the prepared query (row stride <= 32 * TMAX + 32
// floats for any dim of that TMAX) and its codes (cstride <= that rounded
// up to 128 bytes).
template <int TMAX>
constexpr int descent_threads() { return TMAX <= 8 ? 256 : 64; }
template <int TMAX>
constexpr int descent_stride() { return 32 * TMAX + 32; }
template <int TMAX>
constexpr int descent_code_words() { return (32 * TMAX + 32 + 127) / 128 * 16; }
template <int TMAX, int LPQ, bool EXACT>
__global__ void descent_kernel(const SearchArgs P);
// One block of 1024 threads: order[] = the nq queries by descending
// order_bin of their seed distance (with specs, a mixed batch: by
// descending estimated work, ef over the filter's pass rate sampled from
// the first 4096 words of each of up to 256 bitmaps); resets the search
// kernel's query counter for the next pass.
__global__ void order_queries_kernel(const uint64_t *seed, const hh_query_spec *specs, uint64_t nq,
                                     const uint32_t *filters, uint32_t filter_words,
                                     uint32_t nfilters, uint32_t *order, uint32_t *counter);
__global__ void prepare_queries_kernel(const PrepArgs A);

}  // namespace hhk
