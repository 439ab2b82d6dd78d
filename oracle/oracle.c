/*
 * oracle.c — CPU restatement of the hnsw-index query path (test
 * infrastructure only; see oracle.h).
 *
 * Compiled with -ffp-contract=off so that every `a * b + c` below rounds
 * twice, as Rust (which never contracts) does; fused operations are written
 * as fmaf() exactly where the reference calls an FMA intrinsic.
 *
 * How each part is pinned (tests/test_oracle.py, tests/test_host.py):
 *   SplitMix64 ............ Vigna KAT (src/synthetic.rs:182-187) and its
 *                           closed form
 *   gaussian / Box-Muller . first components of gaussian(10000,200,128,42)
 *                           (tests/recall.rs:130-139); Box-Muller of the
 *                           stream; moments test
 *   next_below ............ Lemire's multiply-shift over the raw stream
 *                           (a modulo reduction fails the test)
 *   clustered ............. the crate's draw order (src/synthetic.rs:102-124)
 *                           by an independent restatement; both published
 *                           100k recall tables (docs/benchmarks.md:117-121,
 *                           f32 and int8 columns) reproduced exactly
 *   kernels ............... scalar KATs (src/kernels/scalar.rs:28-33); exact on
 *                           small integers (tests/kernels.rs:83-96); every
 *                           backend within the scalar tolerance
 *                           (tests/kernels.rs:11-14); AVX2 order checked
 *                           against numpy's float32 restatement of the
 *                           intrinsic sequence and against an FMA-free case
 *   int8 kernels .......... naive i64 sums for len 0..80 and MAX_DIM extremes
 *                           (src/kernels/int8.rs:68-85)
 *   prepare ............... src/vector.rs:106-179 KATs
 *   quantizer ............. src/quantization.rs:101-145 KATs; rounding half
 *                           away from zero; approximate distance rounding
 *                           (src/quantization.rs:85-92) against a numpy
 *                           restatement (an f64 product fails the test)
 *   per-dimension codes ... closed forms of the min/max parameters; the
 *                           distance against float64 within its stated
 *                           error bound (no crate counterpart)
 *   bf16 rows ............. round-to-nearest-even against torch's bfloat16
 *   brute force ........... src/brute.rs:218-262 KATs; FP64 variant vs numpy
 *   HNSW search ........... n <= M complete graph == brute force
 *                           (src/hnsw/tests.rs:113-127); the 10k recall
 *                           table of docs/benchmarks.md:162-169 exactly;
 *                           k/ef edge cases; filter KATs
 *                           (src/hnsw/tests.rs:293-313)
 *   beam steps ............ a hand-worked six-node graph (expansions and
 *                           distances per width); complete graphs equal
 *                           brute force; recall and cost against width 1
 *                           (no crate counterpart)
 *   recall ................ src/eval.rs:59-72 KATs; hh_recall_at_k against
 *                           them and the oracle
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- random */

/* src/synthetic.rs:56-60 */
uint64_t or_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

/* src/synthetic.rs:32-35 */
uint64_t or_splitmix_next(uint64_t *state) {
  *state += 0x9E3779B97F4A7C15ULL;
  return or_mix64(*state);
}

/* src/synthetic.rs:38-40 */
double or_next_f64(uint64_t *state) {
  return (double)(or_splitmix_next(state) >> 11) * (1.0 / 9007199254740992.0);
}

/* src/synthetic.rs:43-45 (Lemire multiply-shift) */
uint64_t or_next_below(uint64_t *state, uint64_t n) {
  unsigned __int128 p = (unsigned __int128)or_splitmix_next(state) * n;
  return (uint64_t)(p >> 64);
}

/* src/synthetic.rs:48-52 (Box-Muller, cosine branch) */
double or_next_gaussian(uint64_t *state) {
  const double tau = 6.283185307179586;
  double u1 = 1.0 - or_next_f64(state);
  double u2 = or_next_f64(state);
  return sqrt(-2.0 * log(u1)) * cos(tau * u2);
}

/* src/synthetic.rs:81-92: base rows first, then queries, one stream. */
void or_dataset_gaussian(size_t base, size_t queries, size_t dim, uint64_t seed,
                         float *out_base, float *out_queries) {
  uint64_t s = seed;
  for (size_t i = 0; i < base * dim; i++) out_base[i] = (float)or_next_gaussian(&s);
  for (size_t i = 0; i < queries * dim; i++) out_queries[i] = (float)or_next_gaussian(&s);
}

/* src/synthetic.rs:102-124: N(0,1) centres in f64, then per point a centre
 * index followed by dim spread draws. */
void or_dataset_clustered(size_t base, size_t queries, size_t dim, size_t clusters,
                          uint64_t seed, float *out_base, float *out_queries) {
  const double spread = 0.35;
  uint64_t s = seed;
  double *centers = (double *)malloc(sizeof(double) * clusters * dim);
  for (size_t i = 0; i < clusters * dim; i++) centers[i] = or_next_gaussian(&s);
  for (int pass = 0; pass < 2; pass++) {
    size_t count = pass == 0 ? base : queries;
    float *out = pass == 0 ? out_base : out_queries;
    for (size_t r = 0; r < count; r++) {
      size_t c = (size_t)or_next_below(&s, clusters);
      for (size_t d = 0; d < dim; d++) {
        double g = or_next_gaussian(&s);
        out[r * dim + d] = (float)(centers[c * dim + d] + spread * g);
      }
    }
  }
  free(centers);
}

/* --------------------------------------------------------------- kernels */

/* src/kernels/scalar.rs:7-22 */
static float dot_scalar(const float *a, const float *b, size_t n) {
  float s = 0.0f;
  for (size_t i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}
static float l2_scalar(const float *a, const float *b, size_t n) {
  float s = 0.0f;
  for (size_t i = 0; i < n; i++) {
    float d = a[i] - b[i];
    s += d * d;
  }
  return s;
}

/* src/kernels/portable.rs:13-60: 16 lanes, pairwise lane reduction, tail. */
static float portable_reduce(float *acc) {
  for (int w = 8; w > 0; w /= 2)
    for (int i = 0; i < w; i++) acc[i] += acc[i + w];
  return acc[0];
}
static float dot_portable(const float *a, const float *b, size_t n) {
  float acc[16] = {0};
  size_t chunks = n / 16;
  for (size_t c = 0; c < chunks; c++)
    for (int l = 0; l < 16; l++) acc[l] += a[c * 16 + l] * b[c * 16 + l];
  float s = portable_reduce(acc);
  for (size_t i = chunks * 16; i < n; i++) s += a[i] * b[i];
  return s;
}
static float l2_portable(const float *a, const float *b, size_t n) {
  float acc[16] = {0};
  size_t chunks = n / 16;
  for (size_t c = 0; c < chunks; c++)
    for (int l = 0; l < 16; l++) {
      float d = a[c * 16 + l] - b[c * 16 + l];
      acc[l] += d * d;
    }
  float s = portable_reduce(acc);
  for (size_t i = chunks * 16; i < n; i++) {
    float d = a[i] - b[i];
    s += d * d;
  }
  return s;
}

/* src/kernels/x86.rs:19-131.  acc[c][l] is lane l of the c-th __m256
 * accumulator: 32-wide blocks feed acc0..acc3 with one FMA each, 8-wide
 * remainder blocks feed acc0, then lanes combine as
 * s[l] = (acc0+acc1)+(acc2+acc3) and h = ((s0+s4)+(s1+s5))+((s2+s6)+(s3+s7))
 * (x86.rs:26 parses left to right), and the scalar tail adds unfused
 * products (x86.rs:75-77,126-129). */
static float avx2_finish(float acc[4][8]) {
  float s[8];
  for (int l = 0; l < 8; l++) s[l] = (acc[0][l] + acc[1][l]) + (acc[2][l] + acc[3][l]);
  return ((s[0] + s[4]) + (s[1] + s[5])) + ((s[2] + s[6]) + (s[3] + s[7]));
}
static float dot_avx2(const float *a, const float *b, size_t n) {
  float acc[4][8];
  memset(acc, 0, sizeof acc);
  size_t i = 0;
  for (; i + 32 <= n; i += 32)
    for (int c = 0; c < 4; c++)
      for (int l = 0; l < 8; l++) acc[c][l] = fmaf(a[i + 8 * c + l], b[i + 8 * c + l], acc[c][l]);
  for (; i + 8 <= n; i += 8)
    for (int l = 0; l < 8; l++) acc[0][l] = fmaf(a[i + l], b[i + l], acc[0][l]);
  float s = avx2_finish(acc);
  for (; i < n; i++) s += a[i] * b[i];
  return s;
}
static float l2_avx2(const float *a, const float *b, size_t n) {
  float acc[4][8];
  memset(acc, 0, sizeof acc);
  size_t i = 0;
  for (; i + 32 <= n; i += 32)
    for (int c = 0; c < 4; c++)
      for (int l = 0; l < 8; l++) {
        float d = a[i + 8 * c + l] - b[i + 8 * c + l];
        acc[c][l] = fmaf(d, d, acc[c][l]);
      }
  for (; i + 8 <= n; i += 8)
    for (int l = 0; l < 8; l++) {
      float d = a[i + l] - b[i + l];
      acc[0][l] = fmaf(d, d, acc[0][l]);
    }
  float s = avx2_finish(acc);
  for (; i < n; i++) {
    float d = a[i] - b[i];
    s += d * d;
  }
  return s;
}

float or_dot(int backend, const float *a, const float *b, size_t len) {
  switch (backend) {
    case OR_SCALAR: return dot_scalar(a, b, len);
    case OR_PORTABLE: return dot_portable(a, b, len);
    default: return dot_avx2(a, b, len);
  }
}
float or_l2(int backend, const float *a, const float *b, size_t len) {
  switch (backend) {
    case OR_SCALAR: return l2_scalar(a, b, len);
    case OR_PORTABLE: return l2_portable(a, b, len);
    default: return l2_avx2(a, b, len);
  }
}

/* src/metric.rs:35-41 */
float or_distance(int metric, int backend, const float *a, const float *b, size_t len) {
  switch (metric) {
    case OR_L2: return or_l2(backend, a, b, len);
    case OR_IP: return -or_dot(backend, a, b, len);
    default: return 1.0f - or_dot(backend, a, b, len);
  }
}

/* src/kernels/int8.rs:13-48: 16 i32 lanes, i64 total, i64 tail.  Every
 * partial sum is exact, so the order is immaterial; the lane split is kept
 * to mirror the reference. */
int64_t or_i8_dot(const int8_t *a, const int8_t *b, size_t len) {
  int32_t acc[16] = {0};
  size_t chunks = len / 16;
  for (size_t c = 0; c < chunks; c++)
    for (int l = 0; l < 16; l++) acc[l] += (int32_t)a[c * 16 + l] * (int32_t)b[c * 16 + l];
  int64_t s = 0;
  for (int l = 0; l < 16; l++) s += acc[l];
  for (size_t i = chunks * 16; i < len; i++) s += (int64_t)((int32_t)a[i] * (int32_t)b[i]);
  return s;
}
int64_t or_i8_l2(const int8_t *a, const int8_t *b, size_t len) {
  int32_t acc[16] = {0};
  size_t chunks = len / 16;
  for (size_t c = 0; c < chunks; c++)
    for (int l = 0; l < 16; l++) {
      int32_t d = (int32_t)a[c * 16 + l] - (int32_t)b[c * 16 + l];
      acc[l] += d * d;
    }
  int64_t s = 0;
  for (int l = 0; l < 16; l++) s += acc[l];
  for (size_t i = chunks * 16; i < len; i++) {
    int32_t d = (int32_t)a[i] - (int32_t)b[i];
    s += (int64_t)(d * d);
  }
  return s;
}

/* ------------------------------------------------------------ validation */

/* src/vector.rs:16: f32::MAX as f64 / 8 */
#define OR_MAX_SQUARED_NORM (3.4028234663852886e38 / 8.0)

/* check + prepare_query (src/vector.rs:22-45, 90-104).  out may be NULL. */
int or_prepare(int metric, size_t dim, const float *v, size_t len, float *out) {
  if (len != dim) return OR_DIMENSION_MISMATCH;
  double norm_sq = 0.0;
  for (size_t i = 0; i < len; i++) {
    if (!isfinite(v[i])) return OR_NON_FINITE;
    norm_sq += (double)v[i] * (double)v[i];
  }
  if (norm_sq > OR_MAX_SQUARED_NORM) return OR_NON_FINITE;
  if (metric == OR_COSINE) {
    if (norm_sq == 0.0) return OR_ZERO_VECTOR;
    if (out) {
      double inv = 1.0 / sqrt(norm_sq);
      for (size_t i = 0; i < len; i++) out[i] = (float)((double)v[i] * inv);
    }
  } else if (out) {
    memcpy(out, v, sizeof(float) * len);
  }
  return OR_OK;
}

/* ------------------------------------------------------------- quantizer */

/* src/quantization.rs:32-56 */
void or_sq_train(int metric, size_t dim, size_t rows, const float *vectors,
                 float *offset_out, float *scale_out) {
  for (size_t d = 0; d < dim; d++) offset_out[d] = 0.0f;
  if (metric == OR_L2 && rows > 0) {
    double *sum = (double *)calloc(dim, sizeof(double));
    for (size_t r = 0; r < rows; r++)
      for (size_t d = 0; d < dim; d++) sum[d] += (double)vectors[r * dim + d];
    for (size_t d = 0; d < dim; d++) offset_out[d] = (float)(sum[d] / (double)rows);
    free(sum);
  }
  float max_abs = 0.0f;
  for (size_t r = 0; r < rows; r++)
    for (size_t d = 0; d < dim; d++) {
      float a = fabsf(vectors[r * dim + d] - offset_out[d]);
      max_abs = fmaxf(max_abs, a); /* f32::max ignores NaN, as fmaxf */
    }
  *scale_out = (max_abs > 0.0f && isfinite(max_abs)) ? max_abs / 127.0f : 1.0f;
}

/* src/quantization.rs:73-81: round half away from zero (roundf), clamp. */
void or_sq_encode(const float *offset, float scale, size_t dim, const float *v, int8_t *out) {
  float inv = 1.0f / scale;
  for (size_t d = 0; d < dim; d++) {
    float x = roundf((v[d] - offset[d]) * inv);
    if (x < -127.0f) x = -127.0f;
    if (x > 127.0f) x = 127.0f;
    out[d] = (int8_t)x;
  }
}

/* src/quantization.rs:85-92 */
float or_sq_distance(int metric, float scale, const int8_t *a, const int8_t *b, size_t dim) {
  float s2 = scale * scale;
  switch (metric) {
    case OR_L2: return (float)or_i8_l2(a, b, dim) * s2;
    case OR_IP: return -((float)or_i8_dot(a, b, dim)) * s2;
    default: return 1.0f - (float)or_i8_dot(a, b, dim) * s2;
  }
}

/* ---------------------------------------------------- bf16 traversal */

/* f32 -> bf16, round to nearest with ties to even, widened back to f32 (the
 * bf16 rows the device engine traverses; no crate counterpart). */
void or_bf16_round(const float *in, size_t n, float *out) {
  for (size_t i = 0; i < n; i++) {
    uint32_t u;
    memcpy(&u, &in[i], 4);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) {
      u |= 0x00400000u; /* NaN stays a (quiet) NaN */
    } else {
      u += 0x7FFFu + ((u >> 16) & 1u);
    }
    u &= 0xFFFF0000u;
    memcpy(&out[i], &u, 4);
  }
}

/* ------------------------------------- per-dimension min/max int8 codes */

/* The device engine's per-dimension quantizer (no crate counterpart; the
 * crate's own is the global-scale one above).  Per dimension d over every
 * row: lo_d / hi_d the f32 minimum / maximum, scale_d = (hi_d - lo_d) / 254
 * (1 when that is not positive), mid_d = (lo_d + hi_d) * 0.5 + 0 (the + 0
 * turns a zero sum into +0, so the sign of a zero minimum or maximum never
 * matters), all in f32. */
void or_pd_train(size_t dim, size_t rows, const float *vectors, float *mid, float *scale) {
  for (size_t d = 0; d < dim; d++) {
    float lo = rows ? vectors[d] : 0.0f, hi = lo;
    for (size_t r = 1; r < rows; r++) {
      float x = vectors[r * dim + d];
      if (x < lo) lo = x;
      if (x > hi) hi = x;
    }
    float s = (hi - lo) / 254.0f;
    scale[d] = s > 0.0f ? s : 1.0f;
    mid[d] = (lo + hi) * 0.5f + 0.0f;
  }
}

/* code_d = clamp(round_half_away((x_d - mid_d) * (1 / scale_d)), -127, 127). */
void or_pd_encode(const float *mid, const float *scale, size_t dim, const float *v, int8_t *out) {
  for (size_t d = 0; d < dim; d++) {
    float inv = 1.0f / scale[d];
    float x = roundf((v[d] - mid[d]) * inv);
    if (x < -127.0f) x = -127.0f;
    if (x > 127.0f) x = 127.0f;
    out[d] = (int8_t)x;
  }
}

/* sum_d (scale_d * code_d)^2, products and sum in f64 in dimension order,
 * rounded to f32 once: the row's term of the L2 expansion. */
float or_pd_rnorm(const float *scale, size_t dim, const int8_t *codes) {
  double s = 0.0;
  for (size_t d = 0; d < dim; d++) {
    double y = (double)scale[d] * (double)codes[d];
    s += y * y;
  }
  return (float)s;
}

/* Query side.  Weights w_d = (q_d - mid_d) * scale_d (L2) or q_d * scale_d
 * (inner product, cosine), in f32; sigma = max|w_d| / 127 (1 when zero);
 * qcode_d = clamp(round_half_away(w_d * (1 / sigma)), -127, 127); qconst =
 * sum_d (q_d - mid_d)^2 (L2) or sum_d q_d * mid_d (else), in f64 in order,
 * rounded to f32 once. */
void or_pd_query(int metric, size_t dim, const float *mid, const float *scale, const float *q,
                 int8_t *code, float *sigma, float *qconst) {
  float wmax = 0.0f;
  double c = 0.0;
  for (size_t d = 0; d < dim; d++) {
    float w = metric == OR_L2 ? (q[d] - mid[d]) * scale[d] : q[d] * scale[d];
    float a = fabsf(w);
    if (a > wmax) wmax = a;
    if (metric == OR_L2) {
      double e = (double)q[d] - (double)mid[d];
      c += e * e;
    } else {
      c += (double)q[d] * (double)mid[d];
    }
  }
  float sg = wmax / 127.0f;
  if (!(sg > 0.0f)) sg = 1.0f;
  float inv = 1.0f / sg;
  for (size_t d = 0; d < dim; d++) {
    float w = metric == OR_L2 ? (q[d] - mid[d]) * scale[d] : q[d] * scale[d];
    float x = roundf(w * inv);
    if (x < -127.0f) x = -127.0f;
    if (x > 127.0f) x = 127.0f;
    code[d] = (int8_t)x;
  }
  *sigma = sg;
  *qconst = (float)c;
}

/* acc = sum_d qcode_d * code_d exactly; t = sigma * (f32)acc; L2:
 * (qconst + rnorm) - (t + t); inner product: -(qconst + t); cosine:
 * 1 - (qconst + t); every step one f32 rounding. */
float or_pd_distance(int metric, float sigma, float qconst, float rnorm, const int8_t *qcode,
                     const int8_t *code, size_t dim) {
  int64_t acc = or_i8_dot(qcode, code, dim);
  float t = sigma * (float)acc;
  if (metric == OR_L2) return (qconst + rnorm) - (t + t);
  if (metric == OR_IP) return -(qconst + t);
  return 1.0f - (qconst + t);
}

/* ---------------------------------------------------------- total order */

/* f32::total_cmp */
static inline int32_t total_key(float f) {
  int32_t i;
  memcpy(&i, &f, 4);
  return i ^ (int32_t)(((uint32_t)(i >> 31)) >> 1);
}

/* src/hnsw/search.rs:36-44 */
int or_candidate_cmp(float da, uint32_t ia, float db, uint32_t ib) {
  int32_t ka = total_key(da), kb = total_key(db);
  if (ka != kb) return ka < kb ? -1 : 1;
  if (ia != ib) return ia < ib ? -1 : 1;
  return 0;
}

/* src/neighbor.rs:17-21 (distance, then key) */
static int neighbor_cmp(float da, uint64_t ka, float db, uint64_t kb) {
  int32_t xa = total_key(da), xb = total_key(db);
  if (xa != xb) return xa < xb ? -1 : 1;
  if (ka != kb) return ka < kb ? -1 : 1;
  return 0;
}

/* ----------------------------------------------------------- brute force */

typedef struct {
  double d; /* distance (f32 value widened, or f64) */
  float df;
  uint64_t key;
} ranked;

static int ranked_less(const ranked *a, const ranked *b, int f64) {
  if (f64) {
    if (a->d != b->d) return a->d < b->d;
    return a->key < b->key;
  }
  return neighbor_cmp(a->df, a->key, b->df, b->key) < 0;
}

/* max-heap of the best k (src/brute.rs:143-158) */
static void rheap_push(ranked *h, size_t *len, ranked x, int f64) {
  size_t i = (*len)++;
  h[i] = x;
  while (i > 0) {
    size_t p = (i - 1) / 2;
    if (ranked_less(&h[p], &h[i], f64)) {
      ranked t = h[p]; h[p] = h[i]; h[i] = t; i = p;
    } else break;
  }
}
static void rheap_pop(ranked *h, size_t *len, int f64) {
  h[0] = h[--(*len)];
  size_t i = 0;
  for (;;) {
    size_t l = 2 * i + 1, r = l + 1, m = i;
    if (l < *len && ranked_less(&h[m], &h[l], f64)) m = l;
    if (r < *len && ranked_less(&h[m], &h[r], f64)) m = r;
    if (m == i) break;
    ranked t = h[m]; h[m] = h[i]; h[i] = t; i = m;
  }
}
static int ranked_qcmp64(const void *a, const void *b) {
  const ranked *x = (const ranked *)a, *y = (const ranked *)b;
  if (ranked_less(x, y, 1)) return -1;
  if (ranked_less(y, x, 1)) return 1;
  return 0;
}
static int ranked_qcmp32(const void *a, const void *b) {
  const ranked *x = (const ranked *)a, *y = (const ranked *)b;
  return neighbor_cmp(x->df, x->key, y->df, y->key);
}

static double f64_distance(int metric, const float *a, const float *b, size_t n) {
  double s = 0.0;
  if (metric == OR_L2) {
    for (size_t i = 0; i < n; i++) {
      double d = (double)a[i] - (double)b[i];
      s += d * d;
    }
    return s;
  }
  for (size_t i = 0; i < n; i++) s += (double)a[i] * (double)b[i];
  return metric == OR_IP ? -s : 1.0 - s;
}

static long brute_common(int metric, int backend, size_t dim, size_t n, const uint64_t *keys,
                         const float *vectors, const float *query, size_t k, const uint8_t *allow,
                         int f64, uint64_t *out_keys, float *out_df, double *out_dd) {
  float *q = (float *)malloc(sizeof(float) * (dim ? dim : 1));
  int st = or_prepare(metric, dim, query, dim, q);
  if (st) { free(q); return -st; }
  if (k == 0) { free(q); return 0; }
  if (k > n) k = n; /* src/brute.rs:141-142 */
  ranked *h = (ranked *)malloc(sizeof(ranked) * (k + 1));
  size_t len = 0;
  for (size_t r = 0; r < n; r++) {
    if (allow && !allow[r]) continue;
    ranked c;
    c.key = keys[r];
    if (f64) {
      c.d = f64_distance(metric, q, vectors + r * dim, dim);
      c.df = (float)c.d;
    } else {
      c.df = or_distance(metric, backend, q, vectors + r * dim, dim);
      c.d = c.df;
    }
    if (len < k) rheap_push(h, &len, c, f64);
    else if (ranked_less(&c, &h[0], f64)) {
      rheap_pop(h, &len, f64);
      rheap_push(h, &len, c, f64);
    }
  }
  qsort(h, len, sizeof(ranked), f64 ? ranked_qcmp64 : ranked_qcmp32);
  for (size_t i = 0; i < len; i++) {
    out_keys[i] = h[i].key;
    if (out_df) out_df[i] = h[i].df;
    if (out_dd) out_dd[i] = h[i].d;
  }
  free(h);
  free(q);
  return (long)len;
}

/* src/brute.rs:131-160 */
long or_brute_search(int metric, int backend, size_t dim, size_t n, const uint64_t *keys,
                     const float *vectors, const float *query, size_t k, const uint8_t *allow,
                     uint64_t *out_keys, float *out_dist) {
  return brute_common(metric, backend, dim, n, keys, vectors, query, k, allow, 0, out_keys,
                      out_dist, NULL);
}

long or_brute_search_f64(int metric, size_t dim, size_t n, const uint64_t *keys,
                         const float *vectors, const float *query, size_t k, const uint8_t *allow,
                         uint64_t *out_keys, double *out_dist) {
  return brute_common(metric, OR_AVX2, dim, n, keys, vectors, query, k, allow, 1, out_keys, NULL,
                      out_dist);
}

/* ------------------------------------------------------- thread fan-out */

typedef struct {
  atomic_size_t next;
  size_t total, chunk;
  void (*fn)(void *ctx, size_t i, int worker);
  void *ctx;
} pool_job;

typedef struct {
  pool_job *job;
  int worker;
} pool_arg;

static void *pool_worker(void *arg) {
  pool_arg *pa = (pool_arg *)arg;
  pool_job *job = pa->job;
  for (;;) {
    size_t start = atomic_fetch_add(&job->next, job->chunk);
    if (start >= job->total) break;
    size_t end = start + job->chunk < job->total ? start + job->chunk : job->total;
    for (size_t i = start; i < end; i++) job->fn(job->ctx, i, pa->worker);
  }
  return NULL;
}

/* Dynamic chunking over a fixed thread count, the work split rayon's
 * par_chunks_exact gives search_batch (src/hnsw/mod.rs:532-535). */
static void run_parallel(size_t total, int threads, void (*fn)(void *, size_t, int), void *ctx) {
  pool_job job;
  atomic_init(&job.next, 0);
  job.total = total;
  job.chunk = 4;
  job.fn = fn;
  job.ctx = ctx;
  if (threads < 1) threads = 1;
  pool_arg *args = (pool_arg *)malloc(sizeof(pool_arg) * (size_t)threads);
  for (int i = 0; i < threads; i++) { args[i].job = &job; args[i].worker = i; }
  if (threads == 1) {
    pool_worker(&args[0]);
    free(args);
    return;
  }
  pthread_t *t = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
  for (int i = 0; i < threads; i++) pthread_create(&t[i], NULL, pool_worker, &args[i]);
  for (int i = 0; i < threads; i++) pthread_join(t[i], NULL);
  free(t);
  free(args);
}

typedef struct {
  int metric, backend, f64;
  size_t dim, n, nq, k;
  const uint64_t *keys;
  const float *vectors, *queries;
  uint64_t *out_keys;
  double *out_dist;
  uint32_t *out_count;
  atomic_int status;
} brute_batch_ctx;

static void brute_batch_one(void *vctx, size_t i, int worker) {
  (void)worker;
  brute_batch_ctx *c = (brute_batch_ctx *)vctx;
  long r;
  if (c->f64) {
    r = or_brute_search_f64(c->metric, c->dim, c->n, c->keys, c->vectors, c->queries + i * c->dim,
                            c->k, NULL, c->out_keys + i * c->k, c->out_dist + i * c->k);
  } else {
    float *tmp = (float *)malloc(sizeof(float) * (c->k ? c->k : 1));
    r = or_brute_search(c->metric, c->backend, c->dim, c->n, c->keys, c->vectors,
                        c->queries + i * c->dim, c->k, NULL, c->out_keys + i * c->k, tmp);
    for (long j = 0; j < r; j++) c->out_dist[i * c->k + (size_t)j] = tmp[j];
    free(tmp);
  }
  if (r < 0) {
    atomic_store(&c->status, (int)-r);
    c->out_count[i] = 0;
  } else {
    c->out_count[i] = (uint32_t)r;
  }
}

int or_brute_search_batch(int metric, int backend, size_t dim, size_t n, const uint64_t *keys,
                          const float *vectors, const float *queries, size_t nq, size_t k,
                          int f64, int threads, uint64_t *out_keys, double *out_dist,
                          uint32_t *out_count) {
  brute_batch_ctx c;
  c.metric = metric; c.backend = backend; c.f64 = f64;
  c.dim = dim; c.n = n; c.nq = nq; c.k = k;
  c.keys = keys; c.vectors = vectors; c.queries = queries;
  c.out_keys = out_keys; c.out_dist = out_dist; c.out_count = out_count;
  atomic_init(&c.status, 0);
  run_parallel(nq, threads, brute_batch_one, &c);
  return atomic_load(&c.status);
}

/* ------------------------------------------------------------ HNSW search */

typedef struct {
  float d;
  uint32_t id;
} cand;

static inline int cand_less(cand a, cand b) { return or_candidate_cmp(a.d, a.id, b.d, b.id) < 0; }

/* BinaryHeap<Candidate> (max) and BinaryHeap<Reverse<Candidate>> (min). */
typedef struct {
  cand *v;
  size_t len, cap;
  int max; /* 1: max-heap, 0: min-heap */
} cheap;

static inline int cheap_above(const cheap *h, cand a, cand b) {
  return h->max ? cand_less(b, a) : cand_less(a, b);
}
static void cheap_push(cheap *h, cand x) {
  if (h->len == h->cap) {
    h->cap = h->cap ? h->cap * 2 : 64;
    h->v = (cand *)realloc(h->v, sizeof(cand) * h->cap);
  }
  size_t i = h->len++;
  h->v[i] = x;
  while (i > 0) {
    size_t p = (i - 1) / 2;
    if (cheap_above(h, h->v[i], h->v[p])) {
      cand t = h->v[p]; h->v[p] = h->v[i]; h->v[i] = t; i = p;
    } else break;
  }
}
static cand cheap_pop(cheap *h) {
  cand top = h->v[0];
  h->v[0] = h->v[--h->len];
  size_t i = 0;
  for (;;) {
    size_t l = 2 * i + 1, r = l + 1, m = i;
    if (l < h->len && cheap_above(h, h->v[l], h->v[m])) m = l;
    if (r < h->len && cheap_above(h, h->v[r], h->v[m])) m = r;
    if (m == i) break;
    cand t = h->v[m]; h->v[m] = h->v[i]; h->v[i] = t; i = m;
  }
  return top;
}

typedef struct {
  const or_graph *g;
  const float *q;      /* prepared query */
  const int8_t *code;  /* encoded query, quantized traversal */
  int quantized;       /* traversal precision, OR_PREC_* */
  float pd_sigma, pd_qconst;
  const uint8_t *allow;
  uint32_t *marks; /* VisitedSet (src/hnsw/scratch.rs:13-47) with u32 epochs */
  uint32_t epoch;
  or_counters cnt;
} search_ctx;

static float exact_dist(search_ctx *s, uint32_t id) {
  const or_graph *g = s->g;
  return or_distance((int)g->metric, OR_AVX2, s->q, g->vectors + (size_t)id * g->dim, g->dim);
}
/* traversal distance: exact, or the int8 approximation (src/hnsw/mod.rs:481-493) */
static float trav_dist(search_ctx *s, uint32_t id) {
  const or_graph *g = s->g;
  s->cnt.n_dist++;
  switch (s->quantized) {
    case OR_PREC_INT8:
      return or_sq_distance((int)g->metric, g->sq_scale, s->code, g->codes + (size_t)id * g->dim,
                            g->dim);
    case OR_PREC_BF16:
      return or_distance((int)g->metric, OR_AVX2, s->q, g->bf16_rows + (size_t)id * g->dim,
                         g->dim);
    case OR_PREC_INT8_PERDIM:
      return or_pd_distance((int)g->metric, s->pd_sigma, s->pd_qconst,
                            g->metric == OR_L2 ? g->pd_rnorm[id] : 0.0f, s->code,
                            g->pd_codes + (size_t)id * g->dim, g->dim);
    default:
      return exact_dist(s, id);
  }
}

/* Links::for_each (src/hnsw/graph.rs:88-94) via block (:73-84) */
static const uint32_t *neighbors(const or_graph *g, uint32_t id, int layer, uint32_t *count) {
  const uint32_t *blk;
  uint32_t cap;
  if (layer == 0) {
    blk = g->layer0 + (size_t)id * (g->m0 + 1);
    cap = g->m0;
  } else {
    blk = g->upper + g->upper_off[id] + (size_t)(layer - 1) * (g->m + 1);
    cap = g->m;
  }
  *count = blk[0] < cap ? blk[0] : cap;
  return blk + 1;
}

static int admit(search_ctx *s, uint32_t id) {
  if (s->g->deleted && s->g->deleted[id]) return 0;
  if (s->allow && !s->allow[id]) return 0;
  return 1;
}

/* greedy_closest (src/hnsw/search.rs:49-68) */
static cand greedy_closest(search_ctx *s, cand start, int layer) {
  cand best = start;
  for (;;) {
    uint32_t from = best.id, cnt;
    const uint32_t *nb = neighbors(s->g, from, layer, &cnt);
    for (uint32_t i = 0; i < cnt; i++) {
      float d = trav_dist(s, nb[i]);
      s->cnt.n_upper++;
      if (d < best.d) {
        best.d = d;
        best.id = nb[i];
      }
    }
    if (best.id == from) return best;
  }
}

static int visited_insert(search_ctx *s, uint32_t id) {
  if (s->marks[id] == s->epoch) return 0;
  s->marks[id] = s->epoch;
  return 1;
}

static int cand_qcmp(const void *a, const void *b) {
  const cand *x = (const cand *)a, *y = (const cand *)b;
  return or_candidate_cmp(x->d, x->id, y->d, y->id);
}

/* search_layer (src/hnsw/search.rs:81-142) on layer 0 with one entry point.
 * Returns the sorted results in *out (malloc'd) and their count. */
static size_t search_layer0(search_ctx *s, cand ep, size_t ef, cand **out) {
  cheap cands = {NULL, 0, 0, 0}, res = {NULL, 0, 0, 1};
  s->epoch++;
  if (s->epoch == 0) {
    memset(s->marks, 0, sizeof(uint32_t) * s->g->n);
    s->epoch = 1;
  }
  if (visited_insert(s, ep.id)) {
    cheap_push(&cands, ep);
    if (admit(s, ep.id)) {
      cheap_push(&res, ep);
      if (res.len > ef) cheap_pop(&res);
    }
  }
  while (cands.len) {
    cand cur = cheap_pop(&cands);
    if (res.len >= ef && res.len > 0 && cur.d > res.v[0].d) break;
    s->cnt.n_expand++;
    uint32_t cnt;
    const uint32_t *nb = neighbors(s->g, cur.id, 0, &cnt);
    for (uint32_t i = 0; i < cnt; i++) {
      uint32_t n = nb[i];
      if (!visited_insert(s, n)) continue;
      float d = trav_dist(s, n);
      int improves = res.len < ef || (res.len > 0 && d < res.v[0].d);
      if (improves) {
        cand c = {d, n};
        cheap_push(&cands, c);
        if (cands.len > s->cnt.max_cands) s->cnt.max_cands = cands.len;
        if (admit(s, n)) {
          cheap_push(&res, c);
          if (res.len > ef) cheap_pop(&res);
        }
      }
    }
  }
  free(cands.v);
  qsort(res.v, res.len, sizeof(cand), cand_qcmp);
  *out = res.v;
  return res.len;
}

/* Beam expansion of the bf16 and per-dimension int8 traversals (the device
 * engine's own rule; the crate has neither precision).  Each step pops up
 * to `width` candidates, each only while search.rs:107-110's stop rule does
 * not hold for it (the first one failing it ends the search as there).  The
 * step's batch is every neighbour of the popped candidates not yet visited
 * (visited in pop order, list order), and each of them is tested against
 * the results as they stood at the step's start: it becomes a candidate iff
 * the results held fewer than ef or its distance is below their worst, and
 * enters the results when admitted (the results keep their ef smallest).
 * A step's outcome does not depend on the order of its batch. */
static size_t search_layer0_beam(search_ctx *s, cand ep, size_t ef, size_t width, cand **out) {
  cheap cands = {NULL, 0, 0, 0}, res = {NULL, 0, 0, 1};
  const or_graph *g = s->g;
  s->epoch++;
  if (s->epoch == 0) {
    memset(s->marks, 0, sizeof(uint32_t) * g->n);
    s->epoch = 1;
  }
  if (visited_insert(s, ep.id)) {
    cheap_push(&cands, ep);
    if (admit(s, ep.id)) cheap_push(&res, ep);
  }
  uint32_t exp_ids[8];
  cand *fresh = (cand *)malloc(sizeof(cand) * (width * g->m0 + 1));
  while (cands.len) {
    size_t ne = 0;
    while (ne < width && cands.len) {
      const cand top = cands.v[0];
      if (res.len >= ef && res.len > 0 && top.d > res.v[0].d) break;
      exp_ids[ne++] = cheap_pop(&cands).id;
    }
    if (ne == 0) break;
    s->cnt.n_expand += ne;
    size_t nf = 0;
    for (size_t j = 0; j < ne; j++) {
      uint32_t cnt;
      const uint32_t *nb = neighbors(g, exp_ids[j], 0, &cnt);
      for (uint32_t i = 0; i < cnt; i++) {
        if (!visited_insert(s, nb[i])) continue;
        fresh[nf].id = nb[i];
        fresh[nf].d = trav_dist(s, nb[i]);
        nf++;
      }
    }
    const int full = res.len >= ef && res.len > 0;
    const float worst = res.len ? res.v[0].d : 0.0f;
    for (size_t i = 0; i < nf; i++) {
      if (full && !(fresh[i].d < worst)) continue;
      cheap_push(&cands, fresh[i]);
      if (cands.len > s->cnt.max_cands) s->cnt.max_cands = cands.len;
      if (admit(s, fresh[i].id)) {
        cheap_push(&res, fresh[i]);
        if (res.len > ef) cheap_pop(&res);
      }
    }
  }
  free(fresh);
  free(cands.v);
  qsort(res.v, res.len, sizeof(cand), cand_qcmp);
  *out = res.v;
  return res.len;
}

/* Beam steps of a query with no filter on an index with no tombstones:
 * every visited node is admissible, so the candidates worth expanding are
 * exactly the results not yet expanded.  One sorted list U holds the ef
 * smallest (distance, id) seen, each with an expanded mark.  Each step pops
 * the first `width` unexpanded entries of U (none left ends the search),
 * visits their lists like search_layer0_beam, and merges every new
 * neighbour that passes the step-start test (U held fewer than ef, or its
 * distance is below U's worst) into U, which keeps its ef smallest.  It
 * differs from search_layer0_beam only where a distance equals the worst
 * result's: a neighbour that falls out of the results then stays a
 * candidate there and is dropped here. */
static size_t search_layer0_unified(search_ctx *s, cand ep, size_t ef, size_t width,
                                    cand **out) {
  const or_graph *g = s->g;
  s->epoch++;
  if (s->epoch == 0) {
    memset(s->marks, 0, sizeof(uint32_t) * g->n);
    s->epoch = 1;
  }
  cand *u = (cand *)malloc(sizeof(cand) * (ef + width * g->m0 + 1));
  uint8_t *done = (uint8_t *)malloc(ef + width * g->m0 + 1);
  cand *fresh = (cand *)malloc(sizeof(cand) * (width * g->m0 + 1));
  size_t un = 0;
  if (visited_insert(s, ep.id)) {
    u[0] = ep;
    done[0] = 0;
    un = 1;
  }
  uint32_t exp_ids[8];
  for (;;) {
    size_t ne = 0;
    for (size_t i = 0; i < un && ne < width; i++)
      if (!done[i]) {
        done[i] = 1;
        exp_ids[ne++] = u[i].id;
      }
    if (ne == 0) break;
    s->cnt.n_expand += ne;
    size_t nf = 0;
    for (size_t j = 0; j < ne; j++) {
      uint32_t cnt;
      const uint32_t *nb = neighbors(g, exp_ids[j], 0, &cnt);
      for (uint32_t i = 0; i < cnt; i++) {
        if (!visited_insert(s, nb[i])) continue;
        fresh[nf].id = nb[i];
        fresh[nf].d = trav_dist(s, nb[i]);
        nf++;
      }
    }
    const int full = un >= ef;
    const float worst = un ? u[un - 1].d : 0.0f;
    for (size_t i = 0; i < nf; i++) {
      if (full && !(fresh[i].d < worst)) continue;
      /* insertion into the sorted list, then keep the ef smallest */
      size_t p = un;
      while (p > 0 && cand_less(fresh[i], u[p - 1])) {
        u[p] = u[p - 1];
        done[p] = done[p - 1];
        p--;
      }
      u[p] = fresh[i];
      done[p] = 0;
      un++;
      if (un > s->cnt.max_cands) s->cnt.max_cands = un;
    }
    if (un > ef) un = ef;
  }
  free(fresh);
  free(done);
  *out = u;
  return un;
}

/* Candidates a beam step of the device engine expands: the requested width
 * (at most 8), at most 256 neighbour slots per step; lists of more than 128
 * slots keep the crate's rule (width 1). */
static size_t beam_steps(const or_graph *g) {
  if (g->beam_width <= 1 || g->m0 > 128 || g->m0 == 0) return 1;
  size_t w = g->beam_width > 8 ? 8 : g->beam_width;
  const size_t fit = 256 / g->m0;
  return w < fit ? w : fit;
}

typedef struct {
  float d;
  uint64_t key;
  uint32_t id;
} hit;
static int hit_cmp(const void *a, const void *b) {
  const hit *x = (const hit *)a, *y = (const hit *)b;
  return neighbor_cmp(x->d, x->key, y->d, y->key);
}

static long search_one(search_ctx *s, const float *query, size_t qlen, size_t k, size_t ef,
                       int quantized, uint64_t *out_keys, float *out_dist, uint32_t *out_ids) {
  const or_graph *g = s->g;
  float *q = (float *)malloc(sizeof(float) * (g->dim ? g->dim : 1));
  int8_t *code = NULL;
  int st = or_prepare((int)g->metric, g->dim, query, qlen, q);
  if (st) { free(q); return -st; }
  if (quantized == OR_PREC_INT8 && !g->codes) { free(q); return -OR_QUANTIZATION_DISABLED; }
  if ((quantized == OR_PREC_BF16 && !g->bf16_rows) ||
      (quantized == OR_PREC_INT8_PERDIM && !g->pd_codes)) {
    free(q);
    return -OR_QUANTIZATION_DISABLED;
  }
  size_t live = 0;
  for (uint64_t i = 0; i < g->n; i++) live += !(g->deleted && g->deleted[i]);
  if (k == 0 || live == 0) { free(q); return 0; }
  if (ef < k) ef = k;
  s->q = q;
  s->quantized = quantized;
  if (quantized == OR_PREC_INT8) {
    code = (int8_t *)malloc(g->dim ? g->dim : 1);
    or_sq_encode(g->sq_offset, g->sq_scale, g->dim, q, code);
    s->code = code;
  } else if (quantized == OR_PREC_INT8_PERDIM) {
    code = (int8_t *)malloc(g->dim ? g->dim : 1);
    or_pd_query((int)g->metric, g->dim, g->pd_mid, g->pd_scale, q, code, &s->pd_sigma,
                &s->pd_qconst);
    s->code = code;
  }
  /* search_graph (src/hnsw/search.rs:146-161) */
  cand *found = NULL;
  size_t nf = 0;
  if (g->top >= 0) {
    cand ep = {trav_dist(s, g->entry_id), g->entry_id};
    s->cnt.n_upper++;
    for (int layer = g->top; layer >= 1; layer--) ep = greedy_closest(s, ep, layer);
    const int beam = (quantized == OR_PREC_BF16 || quantized == OR_PREC_INT8_PERDIM) &&
                     beam_steps(g) > 1;
    if (beam && !s->allow && live == g->n)
      nf = search_layer0_unified(s, ep, ef, beam_steps(g), &found);
    else if (beam)
      nf = search_layer0_beam(s, ep, ef, beam_steps(g), &found);
    else
      nf = search_layer0(s, ep, ef, &found);
  }
  hit *hits = (hit *)malloc(sizeof(hit) * (nf ? nf : 1));
  for (size_t i = 0; i < nf; i++) {
    float d = found[i].d;
    if (quantized) { /* exact re-rank (src/hnsw/mod.rs:495-497) */
      d = exact_dist(s, found[i].id);
      s->cnt.n_rerank++;
    }
    hits[i].d = d;
    hits[i].id = found[i].id;
    hits[i].key = g->keys[found[i].id];
  }
  qsort(hits, nf, sizeof(hit), hit_cmp); /* src/hnsw/mod.rs:501-510 */
  size_t outn = nf < k ? nf : k;
  for (size_t i = 0; i < outn; i++) {
    out_keys[i] = hits[i].key;
    out_dist[i] = hits[i].d;
    if (out_ids) out_ids[i] = hits[i].id;
  }
  free(hits);
  free(found);
  free(code);
  free(q);
  return (long)outn;
}

/* search_impl (src/hnsw/mod.rs:460-511) */
long or_hnsw_search(const or_graph *g, const float *query, size_t qlen, size_t k, size_t ef,
                    int quantized, const uint8_t *allow, uint64_t *out_keys, float *out_dist,
                    uint32_t *out_ids, or_counters *counters) {
  search_ctx s;
  memset(&s, 0, sizeof s);
  s.g = g;
  s.allow = allow;
  s.marks = (uint32_t *)calloc(g->n ? g->n : 1, sizeof(uint32_t));
  long r = search_one(&s, query, qlen, k, ef, quantized, out_keys, out_dist, out_ids);
  free(s.marks);
  if (counters) *counters = s.cnt;
  return r;
}

typedef struct {
  const or_graph *g;
  const float *queries;
  size_t k, ef;
  int quantized;
  const uint8_t *allow;
  uint64_t *out_keys;
  float *out_dist;
  uint32_t *out_count;
  atomic_size_t first_bad;
  atomic_int status;
  atomic_ullong n_dist, n_expand, n_rerank, n_upper, max_cands;
  uint32_t **marks;  /* one VisitedSet per worker */
  uint32_t *epochs;
} hbatch_ctx;

static void hbatch_one(void *vctx, size_t i, int worker) {
  hbatch_ctx *c = (hbatch_ctx *)vctx;
  search_ctx s;
  memset(&s, 0, sizeof s);
  s.g = c->g;
  s.allow = c->allow;
  s.marks = c->marks[worker];
  s.epoch = c->epochs[worker];
  long r = search_one(&s, c->queries + i * c->g->dim, c->g->dim, c->k, c->ef, c->quantized,
                      c->out_keys + i * c->k, c->out_dist + i * c->k, NULL);
  c->epochs[worker] = s.epoch;
  atomic_fetch_add(&c->n_dist, s.cnt.n_dist);
  atomic_fetch_add(&c->n_expand, s.cnt.n_expand);
  atomic_fetch_add(&c->n_rerank, s.cnt.n_rerank);
  atomic_fetch_add(&c->n_upper, s.cnt.n_upper);
  unsigned long long mc = atomic_load(&c->max_cands);
  while (s.cnt.max_cands > mc && !atomic_compare_exchange_weak(&c->max_cands, &mc, s.cnt.max_cands)) {
  }
  if (r < 0) {
    c->out_count[i] = 0;
    size_t cur = atomic_load(&c->first_bad);
    while (i < cur && !atomic_compare_exchange_weak(&c->first_bad, &cur, i)) {
    }
    atomic_store(&c->status, (int)-r);
  } else {
    c->out_count[i] = (uint32_t)r;
  }
}

long or_hnsw_search_batch(const or_graph *g, const float *queries, size_t nq, size_t k, size_t ef,
                          int quantized, const uint8_t *allow, int threads, uint64_t *out_keys,
                          float *out_dist, uint32_t *out_count, or_counters *counters,
                          size_t *first_bad) {
  hbatch_ctx c;
  c.g = g; c.queries = queries; c.k = k; c.ef = ef; c.quantized = quantized; c.allow = allow;
  c.out_keys = out_keys; c.out_dist = out_dist; c.out_count = out_count;
  atomic_init(&c.first_bad, (size_t)-1);
  atomic_init(&c.status, 0);
  atomic_init(&c.n_dist, 0); atomic_init(&c.n_expand, 0);
  atomic_init(&c.n_rerank, 0); atomic_init(&c.n_upper, 0); atomic_init(&c.max_cands, 0);
  if (threads < 1) threads = 1;
  c.marks = (uint32_t **)malloc(sizeof(uint32_t *) * (size_t)threads);
  c.epochs = (uint32_t *)calloc((size_t)threads, sizeof(uint32_t));
  for (int t = 0; t < threads; t++) c.marks[t] = (uint32_t *)calloc(g->n ? g->n : 1, sizeof(uint32_t));
  run_parallel(nq, threads, hbatch_one, &c);
  for (int t = 0; t < threads; t++) free(c.marks[t]);
  free(c.marks);
  free(c.epochs);
  if (counters) {
    counters->n_dist = atomic_load(&c.n_dist);
    counters->n_expand = atomic_load(&c.n_expand);
    counters->n_rerank = atomic_load(&c.n_rerank);
    counters->n_upper = atomic_load(&c.n_upper);
    counters->max_cands = atomic_load(&c.max_cands);
  }
  if (first_bad) *first_bad = atomic_load(&c.first_bad);
  int st = atomic_load(&c.status);
  return st ? -st : 0;
}

/* ------------------------------------------------------------------ eval */

/* src/eval.rs:25-47 */
double or_recall_at_k(size_t nq, const uint64_t *truth, const uint32_t *truth_count,
                      size_t truth_stride, const uint64_t *found, const uint32_t *found_count,
                      size_t found_stride, size_t k) {
  size_t hits = 0, expected = 0;
  for (size_t q = 0; q < nq; q++) {
    size_t tn = truth_count[q] < k ? truth_count[q] : k;
    size_t fn = found_count[q] < k ? found_count[q] : k;
    const uint64_t *t = truth + q * truth_stride;
    const uint64_t *f = found + q * found_stride;
    expected += tn;
    for (size_t i = 0; i < fn; i++)
      for (size_t j = 0; j < tn; j++)
        if (f[i] == t[j]) { hits++; break; }
  }
  return expected == 0 ? 1.0 : (double)hits / (double)expected;
}
