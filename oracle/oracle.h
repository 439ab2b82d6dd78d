/*
 * oracle.h — CPU restatement of the hnsw-index query path, used ONLY as
 * test infrastructure: the checker in tests/, the parity check in
 * __graft_entry__.smoke() and the CPU-baseline leg of bench.py.  Nothing in
 * the product library links or calls it.
 *
 * Every function follows a file:line of the reference crate (Rust,
 * aozmekik/hnsw-index), cited at its definition in oracle.c.  The crate
 * cannot be compiled in this environment (no cargo/rustc, crates not
 * vendored), so this restatement is pinned by the crate's own known-answer
 * tests and published figures (tests/golden/reference_kats.json) and by
 * closed forms; see the header of oracle.c for which check pins what.
 */
#ifndef HNSW_ORACLE_H
#define HNSW_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Metric tags, as src/metric.rs:44-50. */
enum { OR_L2 = 0, OR_IP = 1, OR_COSINE = 2 };
/* Kernel backends, as src/kernels/mod.rs:55-66 (NEON is absent on x86). */
enum { OR_SCALAR = 0, OR_PORTABLE = 1, OR_AVX2 = 3 };
/* Status codes (same numbering as include/hnsw_amd.h). */
enum {
  OR_OK = 0,
  OR_DIMENSION_MISMATCH = 1,
  OR_NON_FINITE = 4,
  OR_ZERO_VECTOR = 5,
  OR_QUANTIZATION_DISABLED = 8,
};

/* SplitMix64 and the synthetic datasets (src/synthetic.rs). */
uint64_t or_mix64(uint64_t z);
uint64_t or_splitmix_next(uint64_t *state);
double or_next_f64(uint64_t *state);
uint64_t or_next_below(uint64_t *state, uint64_t n);
double or_next_gaussian(uint64_t *state);
void or_dataset_gaussian(size_t base, size_t queries, size_t dim, uint64_t seed,
                         float *out_base, float *out_queries);
void or_dataset_clustered(size_t base, size_t queries, size_t dim, size_t clusters,
                          uint64_t seed, float *out_base, float *out_queries);

/* f32 kernels (src/kernels/{scalar,portable,x86}.rs). */
float or_dot(int backend, const float *a, const float *b, size_t len);
float or_l2(int backend, const float *a, const float *b, size_t len);
float or_distance(int metric, int backend, const float *a, const float *b, size_t len);
/* int8 kernels (src/kernels/int8.rs). */
int64_t or_i8_dot(const int8_t *a, const int8_t *b, size_t len);
int64_t or_i8_l2(const int8_t *a, const int8_t *b, size_t len);

/* Validation / preparation (src/vector.rs). out may be NULL for check-only. */
int or_prepare(int metric, size_t dim, const float *v, size_t len, float *out);

/* Scalar quantizer (src/quantization.rs). */
void or_sq_train(int metric, size_t dim, size_t rows, const float *vectors,
                 float *offset_out, float *scale_out);
void or_sq_encode(const float *offset, float scale, size_t dim, const float *v, int8_t *out);
float or_sq_distance(int metric, float scale, const int8_t *a, const int8_t *b, size_t dim);

/* Exact brute force (src/brute.rs:131-160).  vectors are prepared rows,
 * query is raw.  allow: optional byte per row (nonzero = filter accepts).
 * Returns the number of results (<= k) or -status. */
long or_brute_search(int metric, int backend, size_t dim, size_t n, const uint64_t *keys,
                     const float *vectors, const float *query, size_t k,
                     const uint8_t *allow, uint64_t *out_keys, float *out_dist);
/* Same ranking with every distance evaluated in f64 (the FP64 oracle). */
long or_brute_search_f64(int metric, size_t dim, size_t n, const uint64_t *keys,
                         const float *vectors, const float *query, size_t k,
                         const uint8_t *allow, uint64_t *out_keys, double *out_dist);
int or_brute_search_batch(int metric, int backend, size_t dim, size_t n, const uint64_t *keys,
                          const float *vectors, const float *queries, size_t nq, size_t k,
                          int f64, int threads, uint64_t *out_keys, double *out_dist,
                          uint32_t *out_count);

/* A graph in the crate's storage layout (src/hnsw/graph.rs:26-33):
 * layer0 = n*(m0+1) words (count, then ids); upper = per node
 * levels[i]*(m+1) words starting at upper_off[i]. */
typedef struct {
  uint32_t dim, metric, m, m0;
  uint64_t n;
  uint32_t entry_id;
  int32_t top; /* -1: empty graph */
  const float *vectors;
  const uint64_t *keys;
  const uint8_t *levels;
  const uint32_t *layer0;
  const uint64_t *upper_off;
  const uint32_t *upper;
  const uint8_t *deleted; /* NULL = none */
  /* optional int8 codes (enable_quantization) */
  const int8_t *codes;
  const float *sq_offset;
  float sq_scale;
  /* bf16 traversal: the rows rounded to bf16 (held as f32), or NULL */
  const float *bf16_rows;
  /* per-dimension int8 traversal: codes n*dim, mid/scale dim each, row
   * code norms n (L2), or NULL */
  const int8_t *pd_codes;
  const float *pd_mid, *pd_scale, *pd_rnorm;
  /* beam width of the bf16 and per-dimension int8 traversals (no crate
   * counterpart): 0 or 1 keeps the crate's one-candidate-at-a-time rule,
   * 2..8 expands that many candidates per step (search_layer0_beam,
   * search_layer0_unified), at most 256 neighbour slots per step */
  uint32_t beam_width;
} or_graph;

/* Traversal precisions (same numbering as hh_precision). */
enum { OR_PREC_F32 = 0, OR_PREC_INT8 = 1, OR_PREC_BF16 = 2, OR_PREC_INT8_PERDIM = 3 };

/* bf16 rounding (round to nearest, ties to even), result widened to f32. */
void or_bf16_round(const float *in, size_t n, float *out);

/* Per-dimension min/max int8 quantizer (no crate counterpart; the device
 * engine's own design, restated here value for value). */
void or_pd_train(size_t dim, size_t rows, const float *vectors, float *mid, float *scale);
void or_pd_encode(const float *mid, const float *scale, size_t dim, const float *v, int8_t *out);
float or_pd_rnorm(const float *scale, size_t dim, const int8_t *codes);
void or_pd_query(int metric, size_t dim, const float *mid, const float *scale, const float *q,
                 int8_t *code, float *sigma, float *qconst);
float or_pd_distance(int metric, float sigma, float qconst, float rnorm, const int8_t *qcode,
                     const int8_t *code, size_t dim);

typedef struct {
  uint64_t n_dist;    /* distance evaluations (traversal metric) */
  uint64_t n_expand;  /* layer-0 expansions */
  uint64_t n_rerank;  /* exact re-scores after int8 traversal */
  uint64_t n_upper;   /* distance evaluations in the upper-layer descent */
  uint64_t max_cands; /* largest candidate heap of any query (live candidates) */
} or_counters;

/* search_impl (src/hnsw/mod.rs:460-511).  allow: optional byte per node id
 * (the host-evaluated filter).  precision: OR_PREC_* (1 is the crate's
 * SearchOptions::quantized; 2 and 3 traverse bf16 rows / per-dimension
 * codes and re-rank exactly like it).  Returns count (<= k) or -status. */
long or_hnsw_search(const or_graph *g, const float *query, size_t qlen, size_t k, size_t ef,
                    int precision, const uint8_t *allow, uint64_t *out_keys, float *out_dist,
                    uint32_t *out_ids, or_counters *counters);
/* search_batch over threads (src/hnsw/mod.rs:520-536); out arrays nq*k.
 * Returns 0 or -status of the first failing query (index in *first_bad). */
long or_hnsw_search_batch(const or_graph *g, const float *queries, size_t nq, size_t k,
                          size_t ef, int precision, const uint8_t *allow, int threads,
                          uint64_t *out_keys, float *out_dist, uint32_t *out_count,
                          or_counters *counters, size_t *first_bad);

/* eval::recall_at_k (src/eval.rs:25-47): truth/found nq*stride with counts. */
double or_recall_at_k(size_t nq, const uint64_t *truth, const uint32_t *truth_count,
                      size_t truth_stride, const uint64_t *found, const uint32_t *found_count,
                      size_t found_stride, size_t k);

/* Candidate ordering (src/hnsw/search.rs:36-44): -1, 0, 1. */
int or_candidate_cmp(float da, uint32_t ia, float db, uint32_t ib);

#ifdef __cplusplus
}
#endif
#endif
