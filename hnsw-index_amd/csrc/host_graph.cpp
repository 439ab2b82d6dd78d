// host_graph.cpp — host side of the engine: input validation, the f32
// kernels the CPU build uses, the scalar quantizer, synthetic data and the
// CPU graph construction that produces what the device index uploads.
//
// Graph construction restates the crate's INSERT (Algorithm 1) and
// neighbour selection (Algorithms 3-4) from src/hnsw/build.rs, with the
// adjacency layout of src/hnsw/graph.rs, so that a graph built here has the
// same structure the crate would hand over.  It is not on the query path.
#include <immintrin.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "hh_internal.h"

namespace hh {

// ------------------------------------------------------------ error text
static thread_local std::string g_last_error;
int fail(int status, const std::string &msg) {
  g_last_error = msg;
  return status;
}
const char *last_error() { return g_last_error.c_str(); }

// ------------------------------------------------------------- kernels
// AVX2 order of src/kernels/x86.rs:19-131 (what Kernels::best() selects on
// x86), used by the CPU graph build and host-side checks.
__attribute__((target("avx2,fma"))) static float hsum(__m256 v) {
  alignas(32) float l[8];
  _mm256_store_ps(l, v);
  return ((l[0] + l[4]) + (l[1] + l[5])) + ((l[2] + l[6]) + (l[3] + l[7]));
}

__attribute__((target("avx2,fma"))) static float dot_avx2(const float *a, const float *b,
                                                          size_t n) {
  __m256 a0 = _mm256_setzero_ps(), a1 = a0, a2 = a0, a3 = a0;
  size_t i = 0;
  for (; i + 32 <= n; i += 32) {
    a0 = _mm256_fmadd_ps(_mm256_loadu_ps(a + i), _mm256_loadu_ps(b + i), a0);
    a1 = _mm256_fmadd_ps(_mm256_loadu_ps(a + i + 8), _mm256_loadu_ps(b + i + 8), a1);
    a2 = _mm256_fmadd_ps(_mm256_loadu_ps(a + i + 16), _mm256_loadu_ps(b + i + 16), a2);
    a3 = _mm256_fmadd_ps(_mm256_loadu_ps(a + i + 24), _mm256_loadu_ps(b + i + 24), a3);
  }
  for (; i + 8 <= n; i += 8)
    a0 = _mm256_fmadd_ps(_mm256_loadu_ps(a + i), _mm256_loadu_ps(b + i), a0);
  float s = hsum(_mm256_add_ps(_mm256_add_ps(a0, a1), _mm256_add_ps(a2, a3)));
  for (; i < n; i++) {
    float p = a[i] * b[i];
    s += p;
  }
  return s;
}

__attribute__((target("avx2,fma"))) static float l2_avx2(const float *a, const float *b,
                                                         size_t n) {
  __m256 a0 = _mm256_setzero_ps(), a1 = a0, a2 = a0, a3 = a0;
  size_t i = 0;
  for (; i + 32 <= n; i += 32) {
    __m256 d0 = _mm256_sub_ps(_mm256_loadu_ps(a + i), _mm256_loadu_ps(b + i));
    __m256 d1 = _mm256_sub_ps(_mm256_loadu_ps(a + i + 8), _mm256_loadu_ps(b + i + 8));
    __m256 d2 = _mm256_sub_ps(_mm256_loadu_ps(a + i + 16), _mm256_loadu_ps(b + i + 16));
    __m256 d3 = _mm256_sub_ps(_mm256_loadu_ps(a + i + 24), _mm256_loadu_ps(b + i + 24));
    a0 = _mm256_fmadd_ps(d0, d0, a0);
    a1 = _mm256_fmadd_ps(d1, d1, a1);
    a2 = _mm256_fmadd_ps(d2, d2, a2);
    a3 = _mm256_fmadd_ps(d3, d3, a3);
  }
  for (; i + 8 <= n; i += 8) {
    __m256 d = _mm256_sub_ps(_mm256_loadu_ps(a + i), _mm256_loadu_ps(b + i));
    a0 = _mm256_fmadd_ps(d, d, a0);
  }
  float s = hsum(_mm256_add_ps(_mm256_add_ps(a0, a1), _mm256_add_ps(a2, a3)));
  for (; i < n; i++) {
    float d = a[i] - b[i];
    float p = d * d;
    s += p;
  }
  return s;
}

// src/kernels/portable.rs:13-60, the fallback of Kernels::best() when the
// CPU lacks AVX2+FMA.
static float reduce16(float *acc) {
  for (int w = 8; w > 0; w /= 2)
    for (int i = 0; i < w; i++) acc[i] += acc[i + w];
  return acc[0];
}
static float dot_portable(const float *a, const float *b, size_t n) {
  float acc[16] = {0};
  size_t c = n / 16;
  for (size_t k = 0; k < c; k++)
    for (int l = 0; l < 16; l++) {
      float p = a[k * 16 + l] * b[k * 16 + l];
      acc[l] += p;
    }
  float s = reduce16(acc);
  for (size_t i = c * 16; i < n; i++) {
    float p = a[i] * b[i];
    s += p;
  }
  return s;
}
static float l2_portable(const float *a, const float *b, size_t n) {
  float acc[16] = {0};
  size_t c = n / 16;
  for (size_t k = 0; k < c; k++)
    for (int l = 0; l < 16; l++) {
      float d = a[k * 16 + l] - b[k * 16 + l];
      float p = d * d;
      acc[l] += p;
    }
  float s = reduce16(acc);
  for (size_t i = c * 16; i < n; i++) {
    float d = a[i] - b[i];
    float p = d * d;
    s += p;
  }
  return s;
}

static bool have_avx2() {
  static const bool v = __builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma");
  return v;
}

// Metric::distance (src/metric.rs:35-41) with Kernels::best().
float distance(uint32_t metric, const float *a, const float *b, size_t n) {
  const bool v = have_avx2();
  switch (metric) {
    case HH_L2: return v ? l2_avx2(a, b, n) : l2_portable(a, b, n);
    case HH_INNER_PRODUCT: return -(v ? dot_avx2(a, b, n) : dot_portable(a, b, n));
    default: return 1.0f - (v ? dot_avx2(a, b, n) : dot_portable(a, b, n));
  }
}

// ------------------------------------------------------------ validation
// check + prepare_into (src/vector.rs:22-63).
static constexpr double kMaxSquaredNorm = 3.4028234663852886e38 / 8.0;

int prepare_vector(uint32_t metric, uint32_t dim, const float *v, float *out) {
  double norm_sq = 0.0;
  for (uint32_t i = 0; i < dim; i++) {
    if (!std::isfinite(v[i])) return fail(HH_NON_FINITE_VECTOR, "vector contains NaN or infinite values");
    double x = (double)v[i];
    double p = x * x;
    norm_sq += p;
  }
  if (norm_sq > kMaxSquaredNorm)
    return fail(HH_NON_FINITE_VECTOR, "vector contains NaN or infinite values");
  if (metric == HH_COSINE) {
    if (norm_sq == 0.0)
      return fail(HH_ZERO_VECTOR, "zero vector has no direction under cosine distance");
    double inv = 1.0 / std::sqrt(norm_sq);
    for (uint32_t i = 0; i < dim; i++) out[i] = (float)((double)v[i] * inv);
  } else if (out != v) {
    std::memcpy(out, v, sizeof(float) * dim);
  }
  return HH_OK;
}

// ------------------------------------------------------------- quantizer
// ScalarQuantizer::train / encode_into (src/quantization.rs:32-81).
void sq_train(uint32_t metric, uint32_t dim, uint64_t rows, const float *vectors,
              std::vector<float> &offset, float &scale) {
  offset.assign(dim, 0.0f);
  if (metric == HH_L2 && rows > 0) {
    std::vector<double> sum(dim, 0.0);
    for (uint64_t r = 0; r < rows; r++)
      for (uint32_t d = 0; d < dim; d++) sum[d] += (double)vectors[r * dim + d];
    for (uint32_t d = 0; d < dim; d++) offset[d] = (float)(sum[d] / (double)rows);
  }
  float max_abs = 0.0f;
  for (uint64_t r = 0; r < rows; r++)
    for (uint32_t d = 0; d < dim; d++)
      max_abs = std::fmax(max_abs, std::fabs(vectors[r * dim + d] - offset[d]));
  scale = (max_abs > 0.0f && std::isfinite(max_abs)) ? max_abs / 127.0f : 1.0f;
}

void sq_encode(const float *offset, float scale, uint32_t dim, const float *v, int8_t *out) {
  float inv = 1.0f / scale;
  for (uint32_t d = 0; d < dim; d++) {
    float x = std::round((v[d] - offset[d]) * inv);
    x = std::min(127.0f, std::max(-127.0f, x));
    out[d] = (int8_t)x;
  }
}

// ---------------------------------------------------------------- random
// SplitMix64 (src/synthetic.rs:22-60).
uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
struct SplitMix {
  uint64_t s;
  explicit SplitMix(uint64_t seed) : s(seed) {}
  uint64_t next() {
    s += 0x9E3779B97F4A7C15ULL;
    return mix64(s);
  }
  double f64() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  uint64_t below(uint64_t n) { return (uint64_t)(((unsigned __int128)next() * n) >> 64); }
  double gaussian() {
    double u1 = 1.0 - f64();
    double u2 = f64();
    double r = std::sqrt(-2.0 * std::log(u1));
    double c = std::cos(6.283185307179586 * u2);
    return r * c;
  }
};

// Runs body(i) for i in [0, n) on `threads` workers, dynamic chunks.
static void parallel_for(uint64_t n, int threads, uint64_t chunk,
                         const std::function<void(uint64_t)> &body) {
  if (threads <= 1 || n <= chunk) {
    for (uint64_t i = 0; i < n; i++) body(i);
    return;
  }
  std::atomic<uint64_t> next{0};
  auto work = [&]() {
    for (;;) {
      uint64_t s = next.fetch_add(chunk);
      if (s >= n) return;
      uint64_t e = std::min(n, s + chunk);
      for (uint64_t i = s; i < e; i++) body(i);
    }
  };
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; t++) pool.emplace_back(work);
  for (auto &t : pool) t.join();
}

// Generators.  Kinds 0 and 1 are the crate's (one stream, base then
// queries).  Kinds >= 2 have no crate counterpart; each row draws from its
// own SplitMix64 stream seeded with mix64(seed ^ stream_id), so rows can be
// generated in parallel and the data do not depend on the thread count.
int synthetic(uint32_t kind, uint64_t n, uint64_t nq, uint32_t dim, uint32_t clusters,
              double sigma, double sigma_q, uint64_t seed, int threads, float *base,
              float *queries, uint32_t *categories) {
  if (dim == 0) return fail(HH_INVALID_DIMENSION, "dim must be positive");
  if (kind == 0) {
    SplitMix r(seed);
    for (uint64_t i = 0; i < n * dim; i++) base[i] = (float)r.gaussian();
    for (uint64_t i = 0; i < nq * dim; i++) queries[i] = (float)r.gaussian();
    return HH_OK;
  }
  if (kind == 1) {
    if (clusters == 0) return fail(HH_INVALID_PARAMETER, "clusters must be positive");
    SplitMix r(seed);
    std::vector<double> c((size_t)clusters * dim);
    for (auto &x : c) x = r.gaussian();
    for (int pass = 0; pass < 2; pass++) {
      uint64_t cnt = pass ? nq : n;
      float *out = pass ? queries : base;
      for (uint64_t i = 0; i < cnt; i++) {
        uint64_t ci = r.below(clusters);
        for (uint32_t d = 0; d < dim; d++) {
          double g = r.gaussian();
          double sp = 0.35 * g;
          out[i * dim + d] = (float)(c[ci * dim + d] + sp);
        }
      }
    }
    return HH_OK;
  }
  if (kind > 5) return fail(HH_INVALID_ARGUMENT, "unknown synthetic kind");
  if ((kind == 2 || kind == 5) && clusters == 0)
    return fail(HH_INVALID_PARAMETER, "clusters must be positive");
  std::vector<double> centers;
  if (kind == 2 || kind == 5) {
    SplitMix r(mix64(seed ^ 0xC3A5C85C97CB3127ULL));
    centers.resize((size_t)clusters * dim);
    for (auto &x : centers) x = kind == 2 ? 2.0 * r.f64() - 1.0 : r.gaussian();
  }
  auto row = [&](uint64_t stream, bool is_query, float *out) {
    SplitMix r(mix64(seed ^ mix64(stream + 1)));
    if (kind == 2 || kind == 5) {
      uint64_t ci = r.below(clusters);
      double s = is_query ? sigma_q : sigma;
      for (uint32_t d = 0; d < dim; d++) {
        double g = r.gaussian();
        double sp = s * g;
        out[d] = (float)(centers[ci * dim + d] + sp);
      }
    } else if (kind == 3) {
      std::vector<double> v(dim);
      double ss = 0.0;
      for (uint32_t d = 0; d < dim; d++) {
        v[d] = r.f64();
        double p = v[d] * v[d];
        ss += p;
      }
      double inv = ss > 0.0 ? 1.0 / std::sqrt(ss) : 0.0;
      for (uint32_t d = 0; d < dim; d++) out[d] = (float)(v[d] * inv);
    } else {  // kind 4
      double g = r.gaussian();
      double scale = std::exp(sigma * g);
      for (uint32_t d = 0; d < dim; d++) out[d] = (float)(r.gaussian() * scale);
    }
  };
  parallel_for(n, threads, 4096, [&](uint64_t i) { row(i, false, base + i * dim); });
  parallel_for(nq, threads, 4096,
               [&](uint64_t i) { row((1ULL << 62) + i, true, queries + i * dim); });
  if (categories) {
    SplitMix r(mix64(seed ^ 0x5851F42D4C957F2DULL));
    for (uint64_t i = 0; i < n; i++) categories[i] = (uint32_t)r.below(clusters ? clusters : 1);
  }
  return HH_OK;
}

// ------------------------------------------------------------------ graph
// Candidate order (src/hnsw/search.rs:36-44): f32::total_cmp, then id.
static inline int32_t total_key(float f) {
  int32_t i;
  std::memcpy(&i, &f, 4);
  return i ^ (int32_t)(((uint32_t)(i >> 31)) >> 1);
}
struct Cand {
  float d;
  uint32_t id;
};
static inline bool cand_less(const Cand &a, const Cand &b) {
  int32_t ka = total_key(a.d), kb = total_key(b.d);
  return ka != kb ? ka < kb : a.id < b.id;
}
struct MaxCmp {  // std heap with cand_less => max-heap
  bool operator()(const Cand &a, const Cand &b) const { return cand_less(a, b); }
};
struct MinCmp {
  bool operator()(const Cand &a, const Cand &b) const { return cand_less(b, a); }
};

static constexpr uint8_t kMaxLevel = 31;          // src/hnsw/mod.rs:26
static constexpr uint64_t kNoEntry = ~0ULL;       // src/hnsw/mod.rs:29
static constexpr uint64_t kMaxNodes = 0xFFFFFFFFULL - 1;  // src/lib.rs:74

// sample_level (src/hnsw/build.rs:19-28).
uint8_t sample_level(uint64_t seed, uint32_t id, double ml) {
  uint64_t h = mix64(seed ^ mix64((uint64_t)id + 0x9E3779B97F4A7C15ULL));
  double u = (double)((h >> 11) + 1) * (1.0 / 9007199254740992.0);
  double level = std::floor(-std::log(u) * ml);
  return level >= (double)kMaxLevel ? kMaxLevel : (uint8_t)level;
}

// Per-thread search buffers (src/hnsw/scratch.rs:13-57).
struct Scratch {
  std::vector<uint32_t> marks;
  uint32_t epoch = 0;
  std::vector<Cand> cands, res;
  void reset(size_t n) {
    if (marks.size() < n) marks.resize(n, 0);
    if (++epoch == 0) {
      std::fill(marks.begin(), marks.end(), 0);
      epoch = 1;
    }
  }
  bool insert(uint32_t id) {
    if (marks[id] == epoch) return false;
    marks[id] = epoch;
    return true;
  }
};

}  // namespace hh

using namespace hh;

struct hh_graph {
  uint32_t dim, metric;
  hh_params p;
  std::vector<float> vectors;
  std::vector<uint64_t> keys;
  std::vector<uint8_t> levels, deleted;
  // nodes whose link lists changed since hh_graph_take_dirty last cleared
  // them (set by any thread that writes a block)
  std::vector<uint8_t> dirty;
  // Links (src/hnsw/graph.rs:26-33): layer 0 = m0+1 words per node; upper
  // blocks of level*(m+1) words per node, located by upper_off.  Words are
  // read with relaxed atomics and counts with acquire, as the crate does.
  std::vector<uint32_t> layer0;
  std::vector<uint64_t> upper_off;
  std::vector<uint32_t> upper;
  std::unique_ptr<std::mutex[]> locks;
  size_t locks_cap = 0;
  std::atomic<uint64_t> entry{kNoEntry};
  std::mutex promotion;
  std::unordered_map<uint64_t, uint32_t> key_to_id;
  bool quantized = false;
  std::vector<float> sq_offset;
  float sq_scale = 1.0f;
  std::vector<int8_t> codes;

  size_t n() const { return keys.size(); }
  const float *vec(uint32_t id) const { return vectors.data() + (size_t)id * dim; }
  float dist(const float *a, const float *b) const { return hh::distance(metric, a, b, dim); }

  uint32_t capacity(int layer) const { return layer == 0 ? p.m0 : p.m; }
  uint32_t *block(uint32_t id, int layer) {
    if (layer == 0) return layer0.data() + (size_t)id * (p.m0 + 1);
    return upper.data() + upper_off[id] + (size_t)(layer - 1) * (p.m + 1);
  }
  // Links::for_each (src/hnsw/graph.rs:88-94)
  template <class F>
  void for_each(uint32_t id, int layer, F f) {
    uint32_t *b = block(id, layer);
    uint32_t c = __atomic_load_n(&b[0], __ATOMIC_ACQUIRE);
    c = std::min(c, capacity(layer));
    for (uint32_t i = 0; i < c; i++) f(__atomic_load_n(&b[1 + i], __ATOMIC_RELAXED));
  }
  void read_into(uint32_t id, int layer, std::vector<uint32_t> &out) {
    out.clear();
    for_each(id, layer, [&](uint32_t x) { out.push_back(x); });
  }
  // Links::set / try_push (src/hnsw/graph.rs:112-135)
  void mark_dirty(uint32_t id) { __atomic_store_n(&dirty[id], (uint8_t)1, __ATOMIC_RELAXED); }
  void set(uint32_t id, int layer, const uint32_t *ids, size_t cnt) {
    mark_dirty(id);
    uint32_t *b = block(id, layer);
    for (size_t i = 0; i < cnt; i++) __atomic_store_n(&b[1 + i], ids[i], __ATOMIC_RELAXED);
    __atomic_store_n(&b[0], (uint32_t)cnt, __ATOMIC_RELEASE);
  }
  bool try_push(uint32_t id, int layer, uint32_t nb) {
    uint32_t *b = block(id, layer);
    uint32_t c = __atomic_load_n(&b[0], __ATOMIC_RELAXED);
    if (c >= capacity(layer)) return false;
    mark_dirty(id);
    __atomic_store_n(&b[1 + c], nb, __ATOMIC_RELAXED);
    __atomic_store_n(&b[0], c + 1, __ATOMIC_RELEASE);
    return true;
  }

  // greedy_closest (src/hnsw/search.rs:49-68)
  Cand greedy_closest(Cand start, int layer, const float *q) {
    Cand best = start;
    for (;;) {
      uint32_t from = best.id;
      for_each(from, layer, [&](uint32_t nb) {
        float d = dist(q, vec(nb));
        if (d < best.d) best = Cand{d, nb};
      });
      if (best.id == from) return best;
    }
  }

  // search_layer (src/hnsw/search.rs:81-142) with admit(n) = n != self_id.
  void search_layer(Scratch &s, const std::vector<Cand> &eps, size_t ef, int layer,
                    const float *q, uint32_t self_id, std::vector<Cand> &out) {
    s.reset(n());
    auto &C = s.cands;
    auto &W = s.res;
    C.clear();
    W.clear();
    auto admit = [&](uint32_t x) { return x != self_id; };
    for (const Cand &ep : eps) {
      if (!s.insert(ep.id)) continue;
      C.push_back(ep);
      std::push_heap(C.begin(), C.end(), MinCmp());
      if (admit(ep.id)) {
        W.push_back(ep);
        std::push_heap(W.begin(), W.end(), MaxCmp());
        if (W.size() > ef) {
          std::pop_heap(W.begin(), W.end(), MaxCmp());
          W.pop_back();
        }
      }
    }
    while (!C.empty()) {
      std::pop_heap(C.begin(), C.end(), MinCmp());
      Cand cur = C.back();
      C.pop_back();
      if (W.size() >= ef && !W.empty() && cur.d > W.front().d) break;
      for_each(cur.id, layer, [&](uint32_t nb) {
        if (!s.insert(nb)) return;
        float d = dist(q, vec(nb));
        bool improves = W.size() < ef || (!W.empty() && d < W.front().d);
        if (!improves) return;
        Cand c{d, nb};
        C.push_back(c);
        std::push_heap(C.begin(), C.end(), MinCmp());
        if (admit(nb)) {
          W.push_back(c);
          std::push_heap(W.begin(), W.end(), MaxCmp());
          if (W.size() > ef) {
            std::pop_heap(W.begin(), W.end(), MaxCmp());
            W.pop_back();
          }
        }
      });
    }
    out.assign(W.begin(), W.end());
    std::sort(out.begin(), out.end(), cand_less);
  }

  // extend_candidates (src/hnsw/build.rs:202-222)
  std::vector<Cand> extend(uint32_t base_id, const float *base, const std::vector<Cand> &cands,
                           int layer) {
    std::unordered_set<uint32_t> seen;
    for (auto &c : cands) seen.insert(c.id);
    seen.insert(base_id);
    std::vector<Cand> out = cands;
    for (auto &c : cands)
      for_each(c.id, layer, [&](uint32_t nb) {
        if (seen.insert(nb).second) out.push_back(Cand{dist(base, vec(nb)), nb});
      });
    std::sort(out.begin(), out.end(), cand_less);
    return out;
  }

  // select_neighbors (src/hnsw/build.rs:152-198)
  std::vector<uint32_t> select(uint32_t base_id, const float *base, const std::vector<Cand> &cands,
                               size_t m, int layer) {
    std::vector<uint32_t> sel;
    if (p.selection == 0) {
      for (size_t i = 0; i < cands.size() && i < m; i++) sel.push_back(cands[i].id);
      return sel;
    }
    std::vector<Cand> ext;
    const std::vector<Cand> *pool = &cands;
    if (p.extend_candidates) {
      ext = extend(base_id, base, cands, layer);
      pool = &ext;
    }
    std::vector<Cand> kept, discarded;
    for (const Cand &c : *pool) {
      if (kept.size() >= m) break;
      const float *v = vec(c.id);
      bool diverse = true;
      for (const Cand &s : kept)
        if (!(dist(v, vec(s.id)) >= c.d)) {
          diverse = false;
          break;
        }
      if (diverse)
        kept.push_back(c);
      else if (p.keep_pruned)
        discarded.push_back(c);
    }
    size_t room = m - kept.size();
    for (size_t i = 0; i < discarded.size() && i < room; i++) kept.push_back(discarded[i]);
    for (auto &c : kept) sel.push_back(c.id);
    return sel;
  }

  // write_shrunk (src/hnsw/build.rs:134-148); caller holds node's lock.
  void write_shrunk(uint32_t node, const std::vector<uint32_t> &ids, int layer) {
    size_t cap = capacity(layer);
    if (ids.size() <= cap) {
      set(node, layer, ids.data(), ids.size());
      return;
    }
    const float *base = vec(node);
    std::vector<Cand> cands;
    cands.reserve(ids.size());
    for (uint32_t c : ids) cands.push_back(Cand{dist(base, vec(c)), c});
    std::sort(cands.begin(), cands.end(), cand_less);
    auto sel = select(node, base, cands, cap, layer);
    set(node, layer, sel.data(), sel.size());
  }

  // install_own_neighbors (src/hnsw/build.rs:103-117)
  void install_own(uint32_t id, const std::vector<uint32_t> &selected, int layer) {
    std::lock_guard<std::mutex> g(locks[id]);
    std::vector<uint32_t> merged;
    read_into(id, layer, merged);
    if (merged.empty()) {
      set(id, layer, selected.data(), selected.size());
      return;
    }
    for (uint32_t x : selected)
      if (std::find(merged.begin(), merged.end(), x) == merged.end()) merged.push_back(x);
    write_shrunk(id, merged, layer);
  }

  // connect (src/hnsw/build.rs:121-130)
  void connect(uint32_t node, uint32_t nw, int layer) {
    std::lock_guard<std::mutex> g(locks[node]);
    std::vector<uint32_t> cur;
    read_into(node, layer, cur);
    if (std::find(cur.begin(), cur.end(), nw) != cur.end() || try_push(node, layer, nw)) return;
    cur.push_back(nw);
    write_shrunk(node, cur, layer);
  }

  // link (src/hnsw/build.rs:41-94)
  void link(uint32_t id, Scratch &s) {
    int level = levels[id];
    const float *q = vec(id);
    uint64_t e = entry.load(std::memory_order_acquire);
    std::unique_lock<std::mutex> promo;
    if (e == kNoEntry || level > (int)(e >> 32)) {
      std::unique_lock<std::mutex> guard(promotion);
      e = entry.load(std::memory_order_acquire);
      if (e == kNoEntry) {
        entry.store(((uint64_t)level << 32) | id, std::memory_order_release);
        return;
      }
      if (level > (int)(e >> 32)) promo = std::move(guard);
    }
    uint32_t entry_id = (uint32_t)e;
    int top = (int)(e >> 32);
    Cand ep{dist(q, vec(entry_id)), entry_id};
    for (int layer = top; layer >= level + 1; layer--) ep = greedy_closest(ep, layer, q);
    size_t ef = std::max<size_t>(p.ef_construction, p.m);
    std::vector<Cand> eps{ep}, found;
    for (int layer = std::min(level, top); layer >= 0; layer--) {
      search_layer(s, eps, ef, layer, q, id, found);
      auto sel = select(id, q, found, p.m, layer);
      install_own(id, sel, layer);
      for (uint32_t nb : sel) connect(nb, id, layer);
      if (!found.empty()) eps = found;
    }
    if (promo.owns_lock()) entry.store(((uint64_t)level << 32) | id, std::memory_order_release);
  }

  // INSERT (src/hnsw/build.rs:41-96) for a level-0 node whose layer-0
  // candidates were found elsewhere (the device search from the greedy
  // descent's end point, ef = max(ef_construction, m)): the selection and
  // linking of the same function.
  void link_from(uint32_t id, std::vector<Cand> &found) {
    std::sort(found.begin(), found.end(), cand_less);
    const float *q = vec(id);
    auto sel = select(id, q, found, p.m, 0);
    install_own(id, sel, 0);
    for (uint32_t nb : sel) connect(nb, id, 0);
  }

  // append_nodes + push_nodes (src/hnsw/mod.rs:737-767)
  uint32_t append(const uint64_t *ks, const float *prepared, uint64_t cnt) {
    uint32_t first = (uint32_t)n();
    vectors.insert(vectors.end(), prepared, prepared + cnt * dim);
    if (quantized) {
      size_t old = codes.size();
      codes.resize(old + cnt * dim);
      for (uint64_t i = 0; i < cnt; i++)
        sq_encode(sq_offset.data(), sq_scale, dim, prepared + i * dim, codes.data() + old + i * dim);
    }
    size_t total = n() + cnt;
    layer0.resize(total * (p.m0 + 1), 0);
    if (locks_cap < total) {
      locks.reset(new std::mutex[total]);
      locks_cap = total;
    }
    for (uint64_t i = 0; i < cnt; i++) {
      uint32_t id = first + (uint32_t)i;
      uint8_t lv = sample_level(p.seed, id, p.level_multiplier);
      upper_off.push_back(upper.size());
      upper.resize(upper.size() + (size_t)lv * (p.m + 1), 0);
      keys.push_back(ks[i]);
      levels.push_back(lv);
      deleted.push_back(0);
      dirty.push_back(0);
      key_to_id[ks[i]] = id;
    }
    return first;
  }

  // link_parallel (src/hnsw/mod.rs:340-349)
  void link_parallel(uint32_t first, int threads) {
    uint32_t end = (uint32_t)n();
    uint32_t next = first;
    if (threads <= 1) {
      Scratch s;
      for (uint32_t id = next; id < end; id++) link(id, s);
      return;
    }
    if (next < end && entry.load() == kNoEntry) {
      Scratch s;
      link(next, s);
      next++;
    }
    std::atomic<uint32_t> cursor{next};
    auto work = [&]() {
      Scratch s;
      for (;;) {
        uint32_t st = cursor.fetch_add(8);
        if (st >= end) return;
        uint32_t en = std::min(end, st + 8);
        for (uint32_t id = st; id < en; id++) link(id, s);
      }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++) pool.emplace_back(work);
    for (auto &t : pool) t.join();
  }
};

namespace hh {
int graph_view(const hh_graph *g, hh_graph_desc *d) {
  std::memset(d, 0, sizeof *d);
  d->dim = g->dim;
  d->metric = g->metric;
  d->m = g->p.m;
  d->m0 = g->p.m0;
  d->n = g->n();
  uint64_t e = g->entry.load();
  d->entry_id = e == kNoEntry ? 0 : (uint32_t)e;
  d->top_level = e == kNoEntry ? -1 : (int32_t)(e >> 32);
  d->vectors = g->vectors.data();
  d->keys = g->keys.data();
  d->levels = g->levels.data();
  d->layer0 = g->layer0.data();
  d->upper_off = g->upper_off.data();
  d->upper = g->upper.data();
  d->deleted = g->deleted.data();
  // a quantized empty graph still reports its quantizer (non-null codes)
  static const int8_t kNoCodes = 0;
  d->codes = g->quantized ? (g->codes.empty() ? &kNoCodes : g->codes.data()) : nullptr;
  d->sq_offset = g->quantized ? g->sq_offset.data() : nullptr;
  d->sq_scale = g->sq_scale;
  d->live = g->key_to_id.size();
  return HH_OK;
}
}  // namespace hh

extern "C" {

const char *hh_last_error(void) { return hh::last_error(); }
uint32_t hh_abi_version(void) { return HH_ABI_VERSION; }

void hh_params_default(uint32_t m, hh_params *out) {
  out->m = m;
  out->m0 = m * 2;
  out->ef_construction = 200;
  out->selection = 1;
  out->extend_candidates = 0;
  out->keep_pruned = 0;
  out->level_multiplier = 1.0 / std::log((double)m);
  out->seed = 0;
}

// HnswIndex::new + HnswParams::validate (src/hnsw/mod.rs:181-187,
// src/params.rs:164-182).
int hh_graph_new(uint32_t dim, uint32_t metric, const hh_params *p, hh_graph **out) {
  if (!out || !p) return fail(HH_INVALID_ARGUMENT, "null argument");
  if (dim == 0 || dim > 65536)
    return fail(HH_INVALID_DIMENSION, "invalid dimension " + std::to_string(dim) +
                                          ": must be between 1 and 65536");
  if (metric > 2) return fail(HH_INVALID_ARGUMENT, "unknown metric");
  if (p->m < 2 || p->m > 512) return fail(HH_INVALID_PARAMETER, "invalid parameter `m`: must be between 2 and 512");
  if (p->m0 < p->m || p->m0 > 2048)
    return fail(HH_INVALID_PARAMETER, "invalid parameter `m0`: must be between m and 2048");
  if (p->ef_construction < 1 || p->ef_construction > 65536)
    return fail(HH_INVALID_PARAMETER, "invalid parameter `ef_construction`: must be between 1 and 65536");
  if (!(std::isfinite(p->level_multiplier) && p->level_multiplier > 0.0 &&
        p->level_multiplier <= 10.0))
    return fail(HH_INVALID_PARAMETER, "invalid parameter `level_multiplier`: must be finite and in (0, 10]");
  hh_graph *g = new (std::nothrow) hh_graph();
  if (!g) return fail(HH_OUT_OF_MEMORY, "allocation failed");
  g->dim = dim;
  g->metric = metric;
  g->p = *p;
  *out = g;
  return HH_OK;
}

void hh_graph_free(hh_graph *g) { delete g; }

int hh_graph_insert_batch(hh_graph *g, const uint64_t *keys, const float *vectors, uint64_t n,
                          int threads, uint64_t *bad_key) {
  if (!g || (n && (!keys || !vectors))) return fail(HH_INVALID_ARGUMENT, "null argument");
  // prepare_batch validates everything before anything is inserted
  // (src/vector.rs:69-87); then duplicate and capacity checks
  // (src/hnsw/mod.rs:318-331).
  std::vector<float> prepared(n * g->dim);
  for (uint64_t i = 0; i < n; i++) {
    int st = prepare_vector(g->metric, g->dim, vectors + i * g->dim, prepared.data() + i * g->dim);
    if (st) return st;
  }
  std::unordered_set<uint64_t> seen;
  seen.reserve(n);
  for (uint64_t i = 0; i < n; i++) {
    if (g->key_to_id.count(keys[i]) || !seen.insert(keys[i]).second) {
      if (bad_key) *bad_key = keys[i];
      return fail(HH_DUPLICATE_KEY, "key " + std::to_string(keys[i]) + " is already present");
    }
  }
  if (kMaxNodes - g->n() < n)
    return fail(HH_CAPACITY_EXCEEDED, "index capacity exceeded");
  uint32_t first = g->append(keys, prepared.data(), n);
  g->link_parallel(first, threads);
  return HH_OK;
}

int hh_graph_append(hh_graph *g, const uint64_t *keys, const float *vectors, uint64_t n,
                    uint64_t *first, uint64_t *bad_key) {
  if (!g || !first || (n && (!keys || !vectors))) return fail(HH_INVALID_ARGUMENT, "null argument");
  std::vector<float> prepared(n * g->dim);
  for (uint64_t i = 0; i < n; i++) {
    int st = prepare_vector(g->metric, g->dim, vectors + i * g->dim, prepared.data() + i * g->dim);
    if (st) return st;
  }
  std::unordered_set<uint64_t> seen;
  seen.reserve(n);
  for (uint64_t i = 0; i < n; i++) {
    if (g->key_to_id.count(keys[i]) || !seen.insert(keys[i]).second) {
      if (bad_key) *bad_key = keys[i];
      return fail(HH_DUPLICATE_KEY, "key " + std::to_string(keys[i]) + " is already present");
    }
  }
  if (kMaxNodes - g->n() < n) return fail(HH_CAPACITY_EXCEEDED, "index capacity exceeded");
  *first = g->append(keys, prepared.data(), n);
  return HH_OK;
}

int hh_graph_link_batch(hh_graph *g, uint64_t first, uint64_t count, const uint32_t *cand_ids,
                        const float *cand_d, const uint32_t *cand_counts, uint32_t cand_stride,
                        int threads) {
  if (!g) return fail(HH_INVALID_ARGUMENT, "null argument");
  if (first > g->n() || count > g->n() - first)
    return fail(HH_INVALID_ARGUMENT, "nodes outside the graph");
  if (cand_ids && (!cand_d || !cand_counts || cand_stride == 0))
    return fail(HH_INVALID_ARGUMENT, "null candidate array");
  const uint32_t begin = (uint32_t)first, end = (uint32_t)(first + count);
  // candidates apply to level-0 nodes of a graph that already has an entry
  const bool have_entry = g->entry.load() != kNoEntry;
  auto one = [&](uint32_t id, Scratch &s) {
    const uint64_t r = id - begin;
    if (cand_ids && have_entry && g->levels[id] == 0) {
      std::vector<Cand> found;
      const uint32_t c = std::min(cand_counts[r], cand_stride);
      found.reserve(c);
      for (uint32_t j = 0; j < c; j++) {
        const uint32_t nb = cand_ids[r * cand_stride + j];
        if (nb >= g->n() || nb == id) continue;
        found.push_back(Cand{cand_d[r * cand_stride + j], nb});
      }
      g->link_from(id, found);
    } else {
      g->link(id, s);
    }
  };
  if (threads <= 1) {
    Scratch s;
    for (uint32_t id = begin; id < end; id++) one(id, s);
    return HH_OK;
  }
  std::atomic<uint32_t> cursor{begin};
  auto work = [&]() {
    Scratch s;
    for (;;) {
      const uint32_t st = cursor.fetch_add(8);
      if (st >= end) return;
      for (uint32_t id = st; id < std::min(end, st + 8); id++) one(id, s);
    }
  };
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; t++) pool.emplace_back(work);
  for (auto &t : pool) t.join();
  return HH_OK;
}

int hh_graph_take_dirty(hh_graph *g, uint32_t *ids, uint64_t cap, uint64_t *count) {
  if (!g || !count) return fail(HH_INVALID_ARGUMENT, "null argument");
  uint64_t c = 0, put = 0;
  for (size_t i = 0; i < g->dirty.size(); i++) {
    if (!g->dirty[i]) continue;
    if (ids && put < cap) {
      ids[put++] = (uint32_t)i;
      g->dirty[i] = 0;
    }
    c++;
  }
  *count = c;
  return HH_OK;
}

int hh_graph_remove(hh_graph *g, uint64_t key, int *removed) {
  if (!g) return fail(HH_INVALID_ARGUMENT, "null argument");
  auto it = g->key_to_id.find(key);
  int r = 0;
  if (it != g->key_to_id.end()) {
    g->deleted[it->second] = 1;
    g->key_to_id.erase(it);
    r = 1;
  }
  if (removed) *removed = r;
  return HH_OK;
}

int hh_graph_enable_quantization(hh_graph *g) {
  if (!g) return fail(HH_INVALID_ARGUMENT, "null argument");
  sq_train(g->metric, g->dim, g->n(), g->vectors.data(), g->sq_offset, g->sq_scale);
  g->codes.resize(g->vectors.size());
  for (size_t i = 0; i < g->n(); i++)
    sq_encode(g->sq_offset.data(), g->sq_scale, g->dim, g->vec((uint32_t)i),
              g->codes.data() + i * g->dim);
  g->quantized = true;
  return HH_OK;
}

int hh_graph_set_quantizer(hh_graph *g, const float *offset, float scale, const int8_t *codes) {
  if (!g || (!offset && g->dim)) return fail(HH_INVALID_ARGUMENT, "null argument");
  // the validity rule of ScalarQuantizer::from_parts (src/quantization.rs:58-62)
  bool valid = std::isfinite(scale) && scale > 0.0f;
  for (uint32_t d = 0; d < g->dim && valid; d++) valid = std::isfinite(offset[d]);
  if (!valid) return fail(HH_INVALID_PARAMETER, "invalid quantizer parameters");
  g->sq_offset.assign(offset, offset + g->dim);
  g->sq_scale = scale;
  g->codes.resize(g->vectors.size());
  if (codes) {
    if (!g->codes.empty()) std::memcpy(g->codes.data(), codes, g->codes.size());
  } else {
    for (size_t i = 0; i < g->n(); i++)
      sq_encode(g->sq_offset.data(), g->sq_scale, g->dim, g->vec((uint32_t)i),
                g->codes.data() + i * g->dim);
  }
  g->quantized = true;
  return HH_OK;
}

int hh_graph_disable_quantization(hh_graph *g) {
  if (!g) return fail(HH_INVALID_ARGUMENT, "null argument");
  g->quantized = false;
  g->codes.clear();
  return HH_OK;
}

uint64_t hh_graph_len(const hh_graph *g) { return g ? g->key_to_id.size() : 0; }
uint64_t hh_graph_node_count(const hh_graph *g) { return g ? g->n() : 0; }

int hh_graph_view(const hh_graph *g, hh_graph_desc *out) {
  if (!g || !out) return fail(HH_INVALID_ARGUMENT, "null argument");
  return hh::graph_view(g, out);
}

int hh_synthetic(uint32_t kind, uint64_t n, uint64_t nq, uint32_t dim, uint32_t clusters,
                 double sigma, double sigma_q, uint64_t seed, int threads, float *out_base,
                 float *out_queries, uint32_t *categories) {
  if ((n && !out_base) || (nq && !out_queries)) return fail(HH_INVALID_ARGUMENT, "null argument");
  return hh::synthetic(kind, n, nq, dim, clusters, sigma, sigma_q, seed, threads, out_base,
                       out_queries, categories);
}

float hh_cpu_distance(uint32_t metric, const float *a, const float *b, uint64_t n) {
  return hh::distance(metric, a, b, n);
}

// eval::recall_at_k (src/eval.rs:25-47)
double hh_recall_at_k(uint64_t nq, const uint64_t *truth, const uint32_t *truth_count,
                      uint32_t truth_stride, const uint64_t *found, const uint32_t *found_count,
                      uint32_t found_stride, uint32_t k) {
  uint64_t hits = 0, expected = 0;
  for (uint64_t q = 0; q < nq; q++) {
    uint32_t tn = std::min(truth_count[q], k), fn = std::min(found_count[q], k);
    const uint64_t *t = truth + q * truth_stride;
    const uint64_t *f = found + q * found_stride;
    expected += tn;
    for (uint32_t i = 0; i < fn; i++)
      for (uint32_t j = 0; j < tn; j++)
        if (f[i] == t[j]) {
          hits++;
          break;
        }
  }
  return expected == 0 ? 1.0 : (double)hits / (double)expected;
}

}  // extern "C"

// ==================================================================== files
// The crate's versioned, checksummed index format, version 1
// (src/hnsw/io.rs:1-28 describes the layout; writer :103-167, loader
// :218-540).  Little-endian throughout; a CRC-32 (IEEE) of every preceding
// byte closes the file.  Loading treats the input as untrusted: the
// envelope is checked first, every column is located before anything is
// parsed, the adjacency section is walked without allocating, and the heap
// the graph would take is compared with the caller's limit before any graph
// storage is allocated.

namespace {

// CRC-32/ISO-HDLC: reflected polynomial 0xEDB88320, initial and final
// complement (src/crc32.rs:1-7), one 256-entry table.
struct CrcTable {
  uint32_t t[256];
  CrcTable() {
    for (uint32_t i = 0; i < 256; i++) {
      uint32_t c = i;
      for (int b = 0; b < 8; b++) c = (c & 1) ? (c >> 1) ^ 0xEDB88320u : c >> 1;
      t[i] = c;
    }
  }
};
const CrcTable &crc_table() {
  static const CrcTable table;
  return table;
}
uint32_t crc32_update(uint32_t crc, const uint8_t *p, size_t n) {
  const uint32_t *t = crc_table().t;
  for (size_t i = 0; i < n; i++) crc = t[(crc ^ p[i]) & 0xFF] ^ (crc >> 8);
  return crc;
}

constexpr uint8_t kMagic[8] = {'H', 'N', 'S', 'W', 'I', 'D', 'X', 0};
constexpr uint32_t kVersion = 1;
constexpr size_t kHeaderLen = 56, kTrailerLen = 4;
constexpr double kLevelSumSlack = 1024.0;     // src/hnsw/io.rs:69-71
constexpr double kUnitNormTolerance = 1e-3;   // src/hnsw/io.rs:72-73
constexpr uint64_t kMemoryFactor = 64;        // src/hnsw/io.rs:74-81
constexpr uint64_t kMemoryFloor = 64ull << 20;

struct Out {
  std::vector<uint8_t> b;
  void bytes(const void *p, size_t n) {
    const uint8_t *s = (const uint8_t *)p;
    b.insert(b.end(), s, s + n);
  }
  void u8(uint8_t v) { b.push_back(v); }
  void u32(uint32_t v) { bytes(&v, 4); }
  void u64(uint64_t v) { bytes(&v, 8); }
  void f64(double v) { bytes(&v, 8); }
};

// Bounds-checked reader over the checksummed body.
struct In {
  const uint8_t *p;
  size_t len, pos;
  bool take(size_t n, const uint8_t **out) {
    if (n > len - pos) return false;
    *out = p + pos;
    pos += n;
    return true;
  }
  template <class T>
  bool get(T *v) {
    const uint8_t *s;
    if (!take(sizeof(T), &s)) return false;
    std::memcpy(v, s, sizeof(T));
    return true;
  }
};

int format_error(hh_format_info *info, uint32_t kind, const char *what, const std::string &msg) {
  if (info) {
    info->kind = kind;
    std::snprintf(info->what, sizeof info->what, "%s", what ? what : "");
  }
  return fail(HH_FORMAT_ERROR, "invalid index file: " + msg);
}
int truncated(hh_format_info *info) {
  return format_error(info, HH_FORMAT_TRUNCATED, "", "truncated");
}
int invalid(hh_format_info *info, const char *what) {
  return format_error(info, HH_FORMAT_INVALID, what, std::string("invalid ") + what);
}

// Heap bytes of a loaded graph in this library's layout: rows, keys,
// levels, tombstones, per-node lock and upper-block offset, layer-0 block,
// key-map entry, upper blocks, int8 codes and quantizer offsets.
uint64_t graph_heap_bytes(uint64_t n, uint32_t dim, uint32_t m, uint32_t m0, uint64_t level_sum,
                          bool quantized) {
  const uint64_t per_node = 4ull * dim + 8 + 1 + 1 + sizeof(std::mutex) + 8 +
                            4ull * (m0 + 1) + 48 + (quantized ? dim : 0);
  return n * per_node + level_sum * 4ull * (m + 1) + (quantized ? 4ull * dim : 0);
}

}  // namespace

extern "C" {

uint32_t hh_crc32(const uint8_t *data, uint64_t len) {
  if (!data) len = 0;
  return ~crc32_update(~0u, data, (size_t)len);
}

int hh_graph_info(const hh_graph *g, uint32_t *dim, uint32_t *metric, hh_params *params) {
  if (!g) return fail(HH_INVALID_ARGUMENT, "null argument");
  if (dim) *dim = g->dim;
  if (metric) *metric = g->metric;
  if (params) *params = g->p;
  return HH_OK;
}

uint64_t hh_graph_memory_usage(const hh_graph *g) {
  if (!g) return 0;
  uint64_t level_sum = 0;
  for (uint8_t l : g->levels) level_sum += l;
  return graph_heap_bytes(g->n(), g->dim, g->p.m, g->p.m0, level_sum, g->quantized);
}

// HnswIndex::to_bytes / write_to (src/hnsw/mod.rs:611-646, io.rs:103-167).
int hh_graph_to_bytes(const hh_graph *g, uint8_t *buf, uint64_t cap, uint64_t *len) {
  if (!g || !len) return fail(HH_INVALID_ARGUMENT, "null argument");
  hh_graph *gm = const_cast<hh_graph *>(g);  // link reads use the graph's atomics
  Out w;
  w.bytes(kMagic, 8);
  w.u32(kVersion);
  const uint64_t e = g->entry.load(std::memory_order_acquire);
  w.u8((uint8_t)g->metric);
  w.u8(g->p.selection ? 1 : 0);
  w.u8(g->p.selection ? (uint8_t)((g->p.extend_candidates ? 1 : 0) | (g->p.keep_pruned ? 2 : 0)) : 0);
  w.u8(g->quantized ? 1 : 0);
  w.u32(g->dim);
  w.u32(g->p.m);
  w.u32(g->p.m0);
  w.u32(g->p.ef_construction);
  w.f64(g->p.level_multiplier);
  w.u64(g->p.seed);
  w.u32((uint32_t)g->n());
  w.u32(e == kNoEntry ? 0xFFFFFFFFu : (uint32_t)e);
  w.bytes(g->keys.data(), g->keys.size() * 8);
  w.bytes(g->levels.data(), g->levels.size());
  w.bytes(g->deleted.data(), g->deleted.size());
  w.bytes(g->vectors.data(), g->vectors.size() * 4);
  std::vector<uint32_t> nb;
  for (uint32_t id = 0; id < (uint32_t)g->n(); id++)
    for (int layer = 0; layer <= (int)g->levels[id]; layer++) {
      gm->read_into(id, layer, nb);
      w.u32((uint32_t)nb.size());
      w.bytes(nb.data(), nb.size() * 4);
    }
  if (g->quantized) {
    w.bytes(g->sq_offset.data(), g->sq_offset.size() * 4);
    float s = g->sq_scale;
    w.bytes(&s, 4);
  }
  const uint32_t crc = hh_crc32(w.b.data(), w.b.size());
  w.u32(crc);
  *len = w.b.size();
  if (!buf) return HH_OK;
  if (cap < w.b.size()) return fail(HH_INVALID_ARGUMENT, "buffer too small for the index file");
  std::memcpy(buf, w.b.data(), w.b.size());
  return HH_OK;
}

// HnswIndex::from_bytes / from_bytes_with_limit (src/hnsw/mod.rs:666-700,
// io.rs:218-540).  max_memory HH_DEFAULT_MEMORY_LIMIT selects the crate's
// default: 64 times the input size, at least 64 MiB.
int hh_graph_from_bytes(const uint8_t *bytes, uint64_t len, uint64_t max_memory, hh_graph **out,
                        hh_format_info *info) {
  if (!out || (!bytes && len)) return fail(HH_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (info) std::memset(info, 0, sizeof *info);
  // ---- envelope (io.rs:244-273)
  if (len < 12) {
    const size_t pre = len < 8 ? (size_t)len : 8;
    if (std::memcmp(bytes, kMagic, pre) == 0) return truncated(info);
    return format_error(info, HH_FORMAT_BAD_MAGIC, "", "missing HNSWIDX magic bytes");
  }
  if (std::memcmp(bytes, kMagic, 8) != 0)
    return format_error(info, HH_FORMAT_BAD_MAGIC, "", "missing HNSWIDX magic bytes");
  uint32_t version;
  std::memcpy(&version, bytes + 8, 4);
  if (version != kVersion) {
    if (info) info->version = version;
    return format_error(info, HH_FORMAT_UNSUPPORTED_VERSION, "",
                        "unsupported format version " + std::to_string(version));
  }
  if (len < kHeaderLen + kTrailerLen) return truncated(info);
  const size_t body = (size_t)len - kTrailerLen;
  uint32_t stored;
  std::memcpy(&stored, bytes + body, 4);
  const uint32_t computed = hh_crc32(bytes, body);
  if (stored != computed) {
    if (info) {
      info->stored = stored;
      info->computed = computed;
    }
    return format_error(info, HH_FORMAT_CHECKSUM_MISMATCH, "", "checksum mismatch");
  }
  In r{bytes, body, 12};
  // ---- header (io.rs:287-358)
  uint8_t metric, sel, flags, quant;
  r.get(&metric);
  r.get(&sel);
  r.get(&flags);
  r.get(&quant);
  if (metric > 2) return invalid(info, "metric");
  if (!((sel == 0 && flags == 0) || (sel == 1 && flags <= 3))) return invalid(info, "neighbor selection");
  if (quant > 1) return invalid(info, "quantized flag");
  uint32_t dim, m, m0, efc, n, entry_id;
  double lm;
  uint64_t seed;
  r.get(&dim);
  if (dim == 0 || dim > 65536) return invalid(info, "dimension");
  r.get(&m);
  r.get(&m0);
  r.get(&efc);
  r.get(&lm);
  r.get(&seed);
  if (m < 2 || m > 512 || m0 < m || m0 > 2048 || efc < 1 || efc > 65536 ||
      !(std::isfinite(lm) && lm > 0.0 && lm <= 10.0))
    return invalid(info, "parameters");
  r.get(&n);
  if ((uint64_t)n > kMaxNodes) return invalid(info, "node count");
  r.get(&entry_id);
  // ---- columns, located before anything is parsed (io.rs:372-402)
  const uint8_t *kcol, *lcol, *fcol, *vcol;
  if (!r.take((size_t)n * 8, &kcol) || !r.take(n, &lcol) || !r.take(n, &fcol) ||
      !r.take((size_t)n * 4 * dim, &vcol))
    return truncated(info);
  uint64_t level_sum = 0;
  for (uint32_t i = 0; i < n; i++) {
    if (lcol[i] > kMaxLevel) return invalid(info, "node level");
    level_sum += lcol[i];
  }
  for (uint32_t i = 0; i < n; i++)
    if (fcol[i] > 1) return invalid(info, "tombstone flag");
  if ((double)level_sum > 2.0 * ((double)n / std::expm1(1.0 / lm)) + kLevelSumSlack)
    return invalid(info, "implausible level distribution");
  // ---- adjacency walked without allocating (io.rs:407-420)
  {
    In s = r;
    for (uint32_t i = 0; i < n; i++)
      for (int layer = 0; layer <= (int)lcol[i]; layer++) {
        uint32_t cnt;
        if (!s.get(&cnt)) return truncated(info);
        if (cnt > (layer == 0 ? m0 : m)) return invalid(info, "neighbor count exceeds capacity");
        const uint8_t *skip;
        if (!s.take((size_t)cnt * 4, &skip)) return truncated(info);
      }
  }
  // ---- memory bound, before any graph storage (io.rs:425-450)
  const uint64_t limit = max_memory == HH_DEFAULT_MEMORY_LIMIT
                             ? std::max<uint64_t>(len * kMemoryFactor, kMemoryFloor)
                             : max_memory;
  const uint64_t required = graph_heap_bytes(n, dim, m, m0, level_sum, quant == 1);
  if (required > limit) {
    if (info) {
      info->required = required;
      info->limit = limit;
    }
    return fail(HH_MEMORY_LIMIT_EXCEEDED,
                "loading the index needs " + std::to_string(required) +
                    " bytes of memory, more than the limit of " + std::to_string(limit));
  }
  // ---- vectors: the checks insert applies (io.rs:453-475)
  for (uint32_t i = 0; i < n; i++) {
    const uint8_t *v = vcol + (size_t)i * 4 * dim;
    double norm_sq = 0.0;
    for (uint32_t e = 0; e < dim; e++) {
      float x;
      std::memcpy(&x, v + 4 * e, 4);
      if (!std::isfinite(x)) return invalid(info, "vector out of range");
      norm_sq += (double)x * (double)x;
    }
    if (norm_sq > kMaxSquaredNorm) return invalid(info, "vector out of range");
    if (metric == HH_COSINE && std::fabs(norm_sq - 1.0) > kUnitNormTolerance)
      return invalid(info, "cosine vector is not unit length");
  }
  hh_params p{};
  p.m = m;
  p.m0 = m0;
  p.ef_construction = efc;
  p.selection = sel;
  p.extend_candidates = flags & 1;
  p.keep_pruned = (flags >> 1) & 1;
  p.level_multiplier = lm;
  p.seed = seed;
  std::unique_ptr<hh_graph> g(new (std::nothrow) hh_graph());
  if (!g) return fail(HH_OUT_OF_MEMORY, "allocation failed");
  g->dim = dim;
  g->metric = metric;
  g->p = p;
  g->keys.resize(n);
  if (n) std::memcpy(g->keys.data(), kcol, (size_t)n * 8);
  g->levels.assign(lcol, lcol + n);
  g->deleted.assign(fcol, fcol + n);
  g->dirty.assign(n, 0);
  g->vectors.resize((size_t)n * dim);
  if (n) std::memcpy(g->vectors.data(), vcol, (size_t)n * 4 * dim);
  g->layer0.assign((size_t)n * (m0 + 1), 0);
  g->upper_off.resize(n);
  uint64_t off = 0;
  for (uint32_t i = 0; i < n; i++) {
    g->upper_off[i] = off;
    off += (uint64_t)lcol[i] * (m + 1);
  }
  g->upper.assign(off, 0);
  g->locks.reset(new std::mutex[std::max<uint32_t>(n, 1)]);
  g->locks_cap = n;
  // ---- links (io.rs:495-520)
  std::vector<uint32_t> ids, sorted;
  for (uint32_t id = 0; id < n; id++)
    for (int layer = 0; layer <= (int)lcol[id]; layer++) {
      uint32_t cnt;
      r.get(&cnt);
      ids.resize(cnt);
      for (uint32_t j = 0; j < cnt; j++) {
        uint32_t nb;
        r.get(&nb);
        if (!(nb < n && nb != id && (int)lcol[nb] >= layer)) return invalid(info, "neighbor id");
        ids[j] = nb;
      }
      sorted = ids;
      std::sort(sorted.begin(), sorted.end());
      if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
        return invalid(info, "duplicate neighbor id");
      g->set(id, layer, ids.data(), cnt);
    }
  // ---- quantizer (io.rs:229-236; ScalarQuantizer::from_parts,
  // src/quantization.rs:58-62)
  if (quant) {
    const uint8_t *oc;
    float scale;
    if (!r.take((size_t)dim * 4, &oc)) return truncated(info);
    g->sq_offset.resize(dim);
    std::memcpy(g->sq_offset.data(), oc, (size_t)dim * 4);
    for (float o : g->sq_offset)
      if (!std::isfinite(o)) return invalid(info, "quantizer offset");
    if (!r.get(&scale)) return truncated(info);
    if (!(std::isfinite(scale) && scale > 0.0f)) return invalid(info, "quantizer");
    g->sq_scale = scale;
    g->codes.resize((size_t)n * dim);
    for (uint32_t i = 0; i < n; i++)
      sq_encode(g->sq_offset.data(), g->sq_scale, dim, g->vec(i), g->codes.data() + (size_t)i * dim);
    g->quantized = true;
  }
  if (r.pos != r.len)
    return format_error(info, HH_FORMAT_TRAILING_BYTES, "", "trailing bytes after the index");
  // a freshly loaded graph has no pending link changes
  std::fill(g->dirty.begin(), g->dirty.end(), (uint8_t)0);
  // ---- key map and entry point (io.rs:522-540)
  g->key_to_id.reserve(n);
  for (uint32_t id = 0; id < n; id++)
    if (!g->deleted[id] && !g->key_to_id.emplace(g->keys[id], id).second)
      return invalid(info, "duplicate live key");
  if (n == 0) {
    if (entry_id != 0xFFFFFFFFu) return invalid(info, "entry point of an empty index");
    g->entry.store(kNoEntry);
  } else {
    uint8_t top = 0;
    for (uint32_t i = 0; i < n; i++) top = std::max(top, lcol[i]);
    if (entry_id >= n || lcol[entry_id] != top) return invalid(info, "entry point");
    g->entry.store(((uint64_t)top << 32) | entry_id, std::memory_order_release);
  }
  *out = g.release();
  return HH_OK;
}

}  // extern "C"
