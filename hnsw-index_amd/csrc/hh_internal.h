// hh_internal.h — declarations shared by the host translation units of the
// engine (not part of the public C ABI).
#pragma once

#include <cstddef>
#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "../../include/hnsw_amd.h"

namespace hh {

// Records `msg` as the calling thread's last error and returns `status`.
int fail(int status, const std::string &msg);
const char *last_error();

// Metric::distance with the CPU kernels Kernels::best() would pick.
float distance(uint32_t metric, const float *a, const float *b, size_t n);
// check + prepare_into (src/vector.rs:22-63); out may alias v.
int prepare_vector(uint32_t metric, uint32_t dim, const float *v, float *out);
void sq_train(uint32_t metric, uint32_t dim, uint64_t rows, const float *vectors,
              std::vector<float> &offset, float &scale);
void sq_encode(const float *offset, float scale, uint32_t dim, const float *v, int8_t *out);
uint64_t mix64(uint64_t z);
int synthetic(uint32_t kind, uint64_t n, uint64_t nq, uint32_t dim, uint32_t clusters,
              double sigma, double sigma_q, uint64_t seed, int threads, float *base,
              float *queries, uint32_t *categories);
int graph_view(const hh_graph *g, hh_graph_desc *d);

// Facts about an uploaded index used by the serving loop.
int index_device(const hh_index *idx);
uint32_t index_dim(const hh_index *idx);
uint64_t index_nodes(const hh_index *idx);
// Wavefronts the search kernel keeps resident on the whole GPU at the
// launch shape of ef's list-size bucket (ef <= 64, 128, 256, 512).
uint32_t index_resident_waves(hh_index *idx, uint32_t ef = 64);
int index_reserve(hh_index *idx, uint64_t nq, uint32_t count);
// Serving loops open on the index (changes of the index are refused while
// any is): +1 at open, -1 at close.
void index_serving(hh_index *idx, int delta);
// hh_search_batch_spec_device with the filter table's own stride in words
// (a serving loop's table), which must cover the index's nodes.
int search_spec_device_strided(hh_index *idx, const float *queries, uint64_t nq,
                               const hh_query_spec *specs, uint32_t k_stride, uint32_t ef_max,
                               const uint32_t *filters, uint32_t nfilters, uint64_t filter_words,
                               uint64_t *out_keys, float *out_dist, uint32_t *out_count,
                               int32_t *status, void *stream);

// Row layout in HBM.  A dimension splits as D = 32*T + 8*U + R (the block
// structure of the crate's AVX2 kernel, src/kernels/x86.rs:47-77).  The
// first 32*T values are stored "accumulator-major": value 32*t + j (the one
// the AVX2 kernel adds into accumulator lane j) goes to
//   blk(t)*128 + j*tw(t) + (t % 4),  blk(t) = t / 4,  tw(t) = min(4, T - 4*blk(t))
// so lane j of a half-wavefront reads its own t = 4b..4b+3 run as one
// 16-byte load and the 32 lanes together read 512 contiguous bytes.  The
// remaining 8*U + R values follow in natural order.  Rows are padded to a
// multiple of 16 floats (64 bytes).
struct RowLayout {
  uint32_t dim, T, U, R, stride;
  explicit RowLayout(uint32_t d = 0)
      : dim(d), T(d / 32), U((d % 32) / 8), R(d % 8), stride((d + 15) / 16 * 16) {}
  uint32_t pos(uint32_t e) const {
    if (e < 32 * T) {
      uint32_t t = e / 32, j = e % 32, b = t / 4;
      uint32_t tw = T - 4 * b < 4 ? T - 4 * b : 4;
      return b * 128 + j * tw + (t % 4);
    }
    return e;
  }
};

}  // namespace hh
