/*
 * hnsw_amd.h — C ABI of the MI355X (gfx950) query engine for hnsw-index
 * graphs.
 *
 * This is the boundary a host binding calls (the crate's Rust FFI, or the
 * ctypes mirror in hnsw-index_amd/hnsw_amd).  Plain pointers and sizes only.
 * Functions whose pointer arguments are marked "device" take HBM pointers
 * (hipMalloc'd or from a framework allocator) and are asynchronous on the
 * given stream; every other pointer is host memory.
 *
 * Each entry point cites the reference interface it replaces
 * (aozmekik/hnsw-index, file:line).
 *
 * Status codes: 0 on success, otherwise one of hh_status.  Codes 1..10 map
 * one-to-one onto the variants of the crate's `Error` (src/error.rs:9-51);
 * codes >= 100 are device-side conditions the crate has no variant for.
 * hh_last_error() returns a message for the calling thread's last failure.
 */
#ifndef HNSW_AMD_H
#define HNSW_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HH_ABI_VERSION 1u

typedef enum hh_status {
  HH_OK = 0,
  HH_DIMENSION_MISMATCH = 1,     /* Error::DimensionMismatch */
  HH_INVALID_DIMENSION = 2,      /* Error::InvalidDimension */
  HH_INVALID_PARAMETER = 3,      /* Error::InvalidParameter */
  HH_NON_FINITE_VECTOR = 4,      /* Error::NonFiniteVector */
  HH_ZERO_VECTOR = 5,            /* Error::ZeroVector */
  HH_DUPLICATE_KEY = 6,          /* Error::DuplicateKey */
  HH_CAPACITY_EXCEEDED = 7,      /* Error::CapacityExceeded */
  HH_QUANTIZATION_DISABLED = 8,  /* Error::QuantizationDisabled */
  HH_FORMAT_ERROR = 9,           /* Error::Format: see hh_format_info */
  HH_MEMORY_LIMIT_EXCEEDED = 10, /* Error::MemoryLimitExceeded */
  HH_DEVICE_ERROR = 100,         /* a HIP call failed */
  HH_OUT_OF_MEMORY = 101,        /* device allocation refused */
  HH_UNSUPPORTED = 102,          /* shape outside what the kernels handle */
  HH_INVALID_ARGUMENT = 103,     /* null pointer, bad enum, ... */
} hh_status;

/* Metric tags: src/metric.rs:44-50. */
typedef enum hh_metric { HH_L2 = 0, HH_INNER_PRODUCT = 1, HH_COSINE = 2 } hh_metric;

/* Traversal precision of a search.  Every non-F32 mode re-ranks the final
 * ef candidates with exact f32 distances, as src/hnsw/mod.rs:483-500 does
 * for its int8 mode. */
typedef enum hh_precision {
  HH_PREC_F32 = 0,          /* exact f32 traversal (SearchOptions::new) */
  HH_PREC_INT8 = 1,         /* crate quantizer codes (SearchOptions::quantized) */
  HH_PREC_BF16 = 2,         /* bf16 rows, f32 accumulation (no crate counterpart) */
  HH_PREC_INT8_PERDIM = 3,  /* per-dimension min/max int8 codes (no crate counterpart;
                               hh_index_quantize_perdim) */
} hh_precision;

/* Last error message of the calling thread (never NULL). */
const char *hh_last_error(void);
/* Number of visible GPUs (0 when none, never an error). */
int hh_device_count(void);
uint32_t hh_abi_version(void);

/* Device memory and streams of the HIP runtime this library runs on (no
 * counterpart in the reference, which has no device).  Callers without a
 * GPU framework of their own use these to hand device buffers and streams
 * to the *_device entry points; a process that also loads another HIP
 * runtime (for example one bundled with a Python framework) must not mix
 * that runtime's streams with these. */
int hh_device_alloc(int device, uint64_t bytes, void **out);
int hh_device_free(void *p);
/* kind: 0 host->device, 1 device->host, 2 device->device; stream NULL =
 * synchronous copy. */
int hh_memcpy(void *dst, const void *src, uint64_t bytes, int kind, void *stream);
int hh_memset(void *dst, int value, uint64_t bytes);
int hh_stream_create(int device, void **out);
int hh_stream_destroy(void *stream);
int hh_stream_synchronize(void *stream);
int hh_device_synchronize(int device);

/* ------------------------------------------------------------------ */
/* Host-side graph construction.                                        */
/* The crate builds its graph on the CPU (HnswIndex::insert /          */
/* insert_batch, src/hnsw/mod.rs:290-349, src/hnsw/build.rs:41-222).   */
/* Where the crate itself is linked, its graph is handed over through  */
/* hh_graph_desc directly; these functions restate that construction   */
/* so the engine can be used, tested and benchmarked without it.        */
/* ------------------------------------------------------------------ */

/* HnswParams (src/params.rs:61-88). selection: 0 Simple, 1 Heuristic. */
typedef struct hh_params {
  uint32_t m, m0, ef_construction;
  uint32_t selection, extend_candidates, keep_pruned;
  double level_multiplier;
  uint64_t seed;
} hh_params;

/* HnswParams::new(m) defaults (src/params.rs:79-88). */
void hh_params_default(uint32_t m, hh_params *out);

typedef struct hh_graph hh_graph;

/* HnswIndex::new (src/hnsw/mod.rs:181-187). */
int hh_graph_new(uint32_t dim, uint32_t metric, const hh_params *params, hh_graph **out);
void hh_graph_free(hh_graph *g);
/* HnswIndex::insert_batch (src/hnsw/mod.rs:318-331).  threads <= 1 links
 * in id order on the calling thread, which equals repeated
 * HnswIndex::insert (src/hnsw/mod.rs:290-302) and is deterministic.
 * On error *bad_key receives the offending key (DuplicateKey). */
int hh_graph_insert_batch(hh_graph *g, const uint64_t *keys, const float *vectors, uint64_t n,
                          int threads, uint64_t *bad_key);
/* HnswIndex::remove (src/hnsw/mod.rs:357-365): *removed = 1 if it was live. */
int hh_graph_remove(hh_graph *g, uint64_t key, int *removed);
/* Construction in two steps, for linking with candidates found on the
 * device (no crate counterpart; together they are HnswIndex::insert_batch,
 * src/hnsw/mod.rs:290-349).  hh_graph_append validates and stores the
 * vectors as insert_batch does (same errors) without linking them; *first
 * receives the first new node id. */
int hh_graph_append(hh_graph *g, const uint64_t *keys, const float *vectors, uint64_t n,
                    uint64_t *first, uint64_t *bad_key);
/* Links nodes [first, first + count) (INSERT, src/hnsw/build.rs:41-96) on
 * `threads` threads.  A level-0 node of a graph that has an entry point
 * takes its layer-0 candidates from row (id - first) of cand_ids / cand_d
 * (cand_counts[r] entries, stride cand_stride; hh_index_build_candidates
 * output) and runs the crate's selection and linking on them; every other
 * node (and every node when cand_ids is NULL) is linked by the full host
 * INSERT. */
int hh_graph_link_batch(hh_graph *g, uint64_t first, uint64_t count, const uint32_t *cand_ids,
                        const float *cand_d, const uint32_t *cand_counts, uint32_t cand_stride,
                        int threads);
/* Node ids whose link lists (any layer) were written since the last call
 * (no crate counterpart: the crate's graph is never mirrored).  *count
 * receives the number of such nodes; when ids is not NULL, the
 * min(cap, count) smallest are written to ids in ascending order and
 * cleared.  Feeds hh_index_sync. */
int hh_graph_take_dirty(hh_graph *g, uint32_t *ids, uint64_t cap, uint64_t *count);
/* HnswIndex::enable_quantization (src/hnsw/mod.rs:544-547). */
int hh_graph_enable_quantization(hh_graph *g);
int hh_graph_disable_quantization(hh_graph *g);
/* Installs a trained quantizer (ScalarQuantizer::from_parts,
 * src/quantization.rs:58-62: offset finite, scale finite and > 0, else
 * HH_INVALID_PARAMETER) with the n*dim codes of the stored vectors, or
 * encodes them here when codes is NULL.  Used with hh_index_quantize to
 * give the host graph the quantizer trained on the device. */
int hh_graph_set_quantizer(hh_graph *g, const float *offset, float scale, const int8_t *codes);
/* HnswIndex::len / node_count / stats (src/hnsw/mod.rs:236,705,562-583). */
uint64_t hh_graph_len(const hh_graph *g);
/* Dimension, metric tag and parameters of a graph (HnswIndex::dim / metric
 * / params, src/hnsw/mod.rs:210-230). */
int hh_graph_info(const hh_graph *g, uint32_t *dim, uint32_t *metric, hh_params *params);
/* Heap bytes the graph holds in this library's layout (the counterpart of
 * HnswIndex::memory_usage, src/hnsw/mod.rs:590-609; the figure a file load
 * compares with its memory limit). */
uint64_t hh_graph_memory_usage(const hh_graph *g);

/* The crate's index file, version 1 (src/hnsw/io.rs:1-28).  FormatError
 * kinds (src/error.rs:56-75): */
typedef enum hh_format_kind {
  HH_FORMAT_BAD_MAGIC = 1,
  HH_FORMAT_UNSUPPORTED_VERSION = 2,
  HH_FORMAT_TRUNCATED = 3,
  HH_FORMAT_CHECKSUM_MISMATCH = 4,
  HH_FORMAT_INVALID = 5,        /* what names the field (the crate's Invalid(&str)) */
  HH_FORMAT_TRAILING_BYTES = 6,
} hh_format_kind;
typedef struct hh_format_info {
  uint32_t kind;                /* hh_format_kind for HH_FORMAT_ERROR */
  uint32_t version;             /* UnsupportedVersion */
  uint32_t stored, computed;    /* ChecksumMismatch */
  uint64_t required, limit;     /* HH_MEMORY_LIMIT_EXCEEDED */
  char what[48];                /* Invalid */
} hh_format_info;
/* HnswIndex::to_bytes / write_to (src/hnsw/mod.rs:611-646): *len receives
 * the file size; the bytes are written when buf holds at least that many. */
int hh_graph_to_bytes(const hh_graph *g, uint8_t *buf, uint64_t cap, uint64_t *len);
/* HnswIndex::from_bytes / from_bytes_with_limit (src/hnsw/mod.rs:666-700):
 * validates the untrusted input as the crate does, in the crate's order,
 * and never allocates graph storage before the file is known to contain the
 * whole graph and its heap fits max_memory (HH_DEFAULT_MEMORY_LIMIT = the
 * crate's default: 64 times the input size, at least 64 MiB).  info
 * (optional) receives the error details. */
#define HH_DEFAULT_MEMORY_LIMIT (~(uint64_t)0)
int hh_graph_from_bytes(const uint8_t *bytes, uint64_t len, uint64_t max_memory, hh_graph **out,
                        hh_format_info *info);
/* CRC-32 (IEEE 802.3, reflected 0xEDB88320) of the index files
 * (src/crc32.rs). */
uint32_t hh_crc32(const uint8_t *data, uint64_t len);
uint64_t hh_graph_node_count(const hh_graph *g);

/* ------------------------------------------------------------------ */
/* The graph as the crate stores it (src/hnsw/graph.rs:26-33,           */
/* src/hnsw/mod.rs:148-172).  Filled by hh_graph_view or by a binding.  */
/* ------------------------------------------------------------------ */
typedef struct hh_graph_desc {
  uint32_t dim, metric, m, m0;
  uint64_t n;               /* nodes, tombstones included */
  uint32_t entry_id;        /* entry point (src/hnsw/mod.rs:715-717) */
  int32_t top_level;        /* -1 for an empty graph */
  const float *vectors;     /* n*dim, prepared (unit length for cosine) */
  const uint64_t *keys;     /* n */
  const uint8_t *levels;    /* n */
  const uint32_t *layer0;   /* n*(m0+1): count, then ids (Links::layer0) */
  const uint64_t *upper_off;/* n: first word of the node's upper blocks */
  const uint32_t *upper;    /* levels[i]*(m+1) words per node at upper_off[i] */
  const uint8_t *deleted;   /* n tombstone flags, or NULL */
  const int8_t *codes;      /* n*dim int8 codes, or NULL (no quantizer) */
  const float *sq_offset;   /* dim quantizer offsets (src/quantization.rs:25-28) */
  float sq_scale;
  uint64_t live;            /* live (non-deleted) nodes */
} hh_graph_desc;

/* Borrowed view of a host graph; valid until the graph is mutated or freed. */
int hh_graph_view(const hh_graph *g, hh_graph_desc *out);

/* ------------------------------------------------------------------ */
/* Device-resident index: upload once, search many times.              */
/* ------------------------------------------------------------------ */
typedef struct hh_index hh_index;

/* Copies the graph into HBM as: ELL layer 0 (n x m0 ids, unused slots
 * 0xFFFFFFFF), upper blocks, permuted f32 rows, keys, tombstone bits and,
 * when present, int8 codes.  Replaces the crate's in-memory HnswIndex for
 * the query path.  device: HIP ordinal.  The search kernels take dims up
 * to 1024; a search of a wider index returns HH_UNSUPPORTED. */
int hh_index_upload(const hh_graph_desc *desc, int device, hh_index **out);
void hh_index_free(hh_index *idx);
/* Re-uploads the tombstone flags (after HnswIndex::remove). */
int hh_index_update_deleted(hh_index *idx, const uint8_t *deleted, uint64_t live);
/* Converts the resident f32 rows to bf16 (round to nearest even) on the
 * device, enabling HH_PREC_BF16. */
int hh_index_enable_bf16(hh_index *idx);
/* Trains the crate's scalar quantizer on the resident rows and encodes
 * them on the device, enabling HH_PREC_INT8 (ScalarQuantizer::train and
 * encode_into, src/quantization.rs:32-81, as HnswIndex::enable_quantization
 * applies them, src/hnsw/mod.rs:544-547).  The trained parameters are
 * bit-identical to hh_graph_enable_quantization's on the same vectors:
 * per-dimension f64 sums in row order (L2), the f32 maximum of |x - offset|
 * and the same rounding of every code.  offset (dim floats) and scale
 * receive the parameters when not NULL.  Replaces a resident quantizer. */
int hh_index_quantize(hh_index *idx, float *offset, float *scale);
/* Brings a resident index up to date with the graph it mirrors after
 * inserts and removals (no crate counterpart: the crate's graph is never
 * mirrored), moving only what changed: the rows, keys, codes and links of
 * the nodes added since the upload or last sync, the layer-0 lists of the
 * old nodes in `dirty` (from hh_graph_take_dirty), the upper-layer links
 * when any of them changed, the tombstones and the entry point.  Device
 * arrays grow by half their size when full.  desc must describe the same
 * graph (dim, metric, m, m0 and the resident quantizer) with at least the
 * resident number of nodes, else HH_INVALID_ARGUMENT: upload it again.
 * Every change of a resident index (this call, hh_index_update_deleted,
 * hh_index_enable_bf16, hh_index_quantize, hh_index_quantize_perdim)
 * waits for the searches in flight and excludes new ones while it runs;
 * it is refused (HH_INVALID_ARGUMENT) while a serving loop is open on the
 * index, and plans created before it fail at launch with
 * HH_INVALID_ARGUMENT (create them again).  New rows take the resident
 * per-dimension codes too. */
int hh_index_sync(hh_index *idx, const hh_graph_desc *desc, const uint32_t *dirty,
                  uint64_t ndirty);
/* Layer-0 construction candidates of the resident rows [first, first +
 * count) (no crate counterpart): each row is searched as its own query,
 * exactly as INSERT searches for a level-0 node (greedy descent from the
 * entry point, then search_layer with `ef`, src/hnsw/build.rs:66-85), with
 * tombstoned nodes as candidates.  ids / dists: count x ef, ascending
 * (distance, id); counts: entries per row.  ef in 1..512.  The rows must
 * have no links yet (appended, then hh_index_sync'd), so a row never finds
 * itself. */
int hh_index_build_candidates(hh_index *idx, uint64_t first, uint64_t count, uint32_t ef,
                              uint32_t *ids, float *dists, uint32_t *counts);
/* Copies the resident codes to host memory as n*dim bytes (the crate's
 * code layout), for hh_graph_set_quantizer.  HH_UNSUPPORTED without codes. */
int hh_index_download_codes(hh_index *idx, int8_t *codes);

/* Per-dimension min/max int8 quantizer (no crate counterpart: the crate's
 * ScalarQuantizer, src/quantization.rs:32-56, has one global scale; ADR
 * 0004 chose that for the CPU).  Trains on the resident rows and encodes
 * them on the device; inserts synced later are encoded with it.  Per
 * dimension d: lo/hi the minimum/maximum over the rows, scale_d =
 * (hi - lo) / 254 (1 when not positive), mid_d = (lo + hi) * 0.5 (+0 for a
 * zero sum), code = clamp(round_half_away((x - mid_d) / scale_d), +-127).
 * A query's weights w_d = (q_d - mid_d) * scale_d (L2) or q_d * scale_d
 * are coded the same way with one scale sigma = max|w_d| / 127, so the
 * traversal distance is one exact integer dot product acc per row: L2
 * (qconst + rnorm_row) - 2 * sigma * acc, inner product -(qconst +
 * sigma * acc), cosine 1 - (qconst + sigma * acc), where qconst =
 * sum (q_d - mid_d)^2 or sum q_d * mid_d and rnorm_row = sum (scale_d *
 * code_d)^2.  Searches with HH_PREC_INT8_PERDIM traverse on these codes and
 * re-rank the final candidates exactly, like HH_PREC_INT8.
 * mid/scale (optional) receive the dim parameters. */
int hh_index_quantize_perdim(hh_index *idx, float *mid, float *scale);
/* The resident per-dimension codes as n*dim bytes and (L2) the n row terms
 * rnorm; either pointer may be NULL. */
int hh_index_download_perdim(hh_index *idx, int8_t *codes, float *rnorm);
/* The resident bf16 rows (hh_index_enable_bf16) as n*dim values in
 * dimension order. */
int hh_index_download_bf16(hh_index *idx, uint16_t *out);
/* Query order of large batches (no counterpart in the crate; results never
 * depend on it).  longest_first != 0 (the default): a batch of more than
 * one and at most 16 queries per resident wavefront runs in two passes,
 * the upper-layer descent of every query, then layer 0 with the queries of
 * the largest descent distance first, so that long queries do not start
 * last.  By default it applies from ef 32 on rows of 128 or 256
 * dimensions with m <= 16, and above ef 64 otherwise (ef: the batch's
 * largest); hh_index_autotune replaces that rule, per ef bucket, by
 * whichever of one and two passes it measured faster.  Mixed batches
 * (hh_search_batch_spec*) of any size above one query per resident
 * wavefront are ordered by estimated work instead: ef over the filter's
 * pass rate.  The descent pass runs 4 (m <= 16) or 2 (m <= 32) queries per
 * wavefront; above m = 32 the search kernel runs it.  0: one pass in query
 * order. */
int hh_index_set_query_order(hh_index *idx, int longest_first);

/* Beam steps of the bf16 and per-dimension int8 traversals (no counterpart
 * in the crate, which has neither precision).  width 1: the crate's layer-0
 * rule (search.rs:81-142), one candidate per expansion, each new neighbour
 * tested against the results as the sequential loop leaves them.  width
 * 2..8 (the default is 4): each step pops up to `width` candidates (each
 * only while the crate's stop rule does not hold for it; at most 256
 * neighbour slots per step, so m0 = 32 takes 8, m0 = 48 five and m0 = 64
 * four; m0 > 128 keeps width 1), gathers every new neighbour of those lists
 * in one batch and tests each against the results as they stood at the
 * step's start: it becomes a candidate iff fewer than ef results were held
 * or its distance is below their worst, and enters the results when
 * admitted.  A query with no filter on an index with no tombstones keeps
 * the results alone, each marked once expanded, and pops their first
 * unexpanded entries (the same outcome except where a distance equals the
 * worst result's).  A step's outcome does not depend on the order of its
 * batch.  f32 and the crate's int8 codes always follow the crate's rule.
 * Refused while a serving loop is open; plans captured before the change
 * are refused. */
int hh_index_set_beam_width(hh_index *idx, uint32_t width);
/* Candidates per beam step this index's layout allows at its width. */
int hh_index_beam_steps(const hh_index *idx, uint32_t *steps);

/* Search statistics accumulated by the kernels since the last reset. */
typedef struct hh_search_stats {
  uint64_t queries;
  uint64_t n_dist;          /* traversal distance evaluations, layer 0 */
  uint64_t n_upper;         /* distance evaluations in the upper descent */
  uint64_t n_expand;        /* layer-0 expansions */
  uint64_t n_rerank;        /* exact re-scores after approximate traversal */
  uint64_t n_overflow;      /* candidates dropped: the per-query candidate pool was full */
  uint64_t n_visited_spill; /* queries whose visited set left LDS */
  uint64_t n_cand_spill;    /* candidates moved from registers to the HBM pool */
  uint64_t n_redo;          /* queries run again with an n-entry candidate pool because their
                               first pool (65,536 entries) dropped a candidate; their results
                               are those of the unbounded heap (src/hnsw/search.rs:125-135) */
} hh_search_stats;
int hh_index_stats(hh_index *idx, hh_search_stats *out);
int hh_index_reset_stats(hh_index *idx);
/* Autotuner (no crate counterpart): times the search of one device batch
 * (arguments as hh_search_batch_device) under each candidate visited-set
 * capacity (LDS per wavefront, hence wavefronts per CU) and wavefronts per
 * block, and keeps the fastest for this ef bucket (ef <= 64, 128, 256,
 * 512) on this index; with the longest-first order on, it then times one
 * pass against two in that shape and keeps the faster order for the bucket.
 * Synchronous.  Results of the timed runs are left in the output buffers. */
int hh_index_autotune(hh_index *idx, const float *queries, uint64_t nq, uint32_t k, uint32_t ef,
                      uint32_t precision, uint64_t *out_keys, float *out_dist,
                      uint32_t *out_count, int32_t *status, void *stream,
                      uint32_t *chosen_hash_cap, int *chosen_wpb);
/* Sets the launch shape of an ef bucket directly (for example one found by
 * hh_index_autotune in an earlier process): visited-set capacity in
 * entries (>= 256 per 64 of ef, LDS per block <= 64 KiB) and waves per
 * block (1, 2 or 4).  hash_cap 0 restores the default choice (and the
 * default query-order rule of the bucket). */
int hh_index_set_launch_shape(hh_index *idx, uint32_t ef, uint32_t hash_cap, int wpb);
/* Times the search kernel of every following launch with a pair of HIP
 * events on the launch stream (no reference counterpart: measurement). */
int hh_index_set_timing(hh_index *idx, int on);
/* Milliseconds of the search kernel of the last timed launch (waits for it). */
int hh_index_last_kernel_ms(hh_index *idx, float *ms);
/* Measurement (no reference counterpart): the rate at which the device
 * delivers the search's row gather on this index's rows with no dependent
 * work, random rows, 16 wavefronts per CU, each 16-lane group keeping four
 * rows in flight (rounds of 16 rows per wavefront).  gbps: payload bytes
 * (dim * 4 per row) per second, best of three timed launches; ms: that
 * launch's time.  Rows of up to 256 floats.  Synchronous. */
int hh_index_gather_probe(hh_index *idx, uint32_t rounds, double *gbps, float *ms);

/*
 * Batched HNSW search (HnswIndex::search_batch / search_with /
 * search_filtered, src/hnsw/mod.rs:412-536).
 *
 * device: queries (nq*dim f32, raw: validation and cosine normalisation
 * happen on the device exactly as src/vector.rs:22-45), out_keys/out_dist
 * (nq*k), out_count (nq), status (nq int32, the per-query hh_status).
 * allow_bits: device bitmap over node ids (bit i of word i/32 set = the
 * filter accepts node i), or NULL for no filter; tombstones are applied on
 * top (src/hnsw/mod.rs:478-480).
 * Results per query: up to k neighbours sorted by (distance, key), fewer
 * only when fewer admissible nodes are reachable.  ef is raised to k.
 * Asynchronous on stream (a hipStream_t, NULL = default stream).
 */
int hh_search_batch_device(hh_index *idx, const float *queries, uint64_t nq, uint32_t k,
                           uint32_t ef, uint32_t precision, const uint32_t *allow_bits,
                           uint64_t *out_keys, float *out_dist, uint32_t *out_count,
                           int32_t *status, void *stream);

/* The same search with no device allocation (no counterpart in the
 * reference: it is for callers that cannot allocate device memory, such as
 * code under HIP graph capture or a tight serving loop).  Arguments and
 * results as hh_search_batch_device, whose kernels it runs; its workspace
 * (prepared queries, statuses when status is NULL, seeds, the query order,
 * visited bitsets, candidate pools and the re-run pools) is carved from a
 * static 64 MiB device array that each device holds for the life of the
 * process.  Every call on a device shares that array and it keeps the last
 * call's contents, so calls that overlap on one device race: run them on
 * one stream (or order them) per device.  The launch takes as many
 * wavefronts as fit beside the batch (at most one per query); a batch and
 * index that do not fit one block of wavefronts (roughly: nq * (row +
 * code + 32) bytes plus, per wavefront, n/8 bytes of bitset and up to
 * 512 KiB of pool, plus 32 * n bytes of re-run pools when n > 65,536)
 * return HH_UNSUPPORTED before anything is launched. */
int hh_search_batch_device_noalloc(hh_index *idx, const float *queries, uint64_t nq, uint32_t k,
                                   uint32_t ef, uint32_t precision, const uint32_t *allow_bits,
                                   uint64_t *out_keys, float *out_dist, uint32_t *out_count,
                                   int32_t *status, void *stream);

/* Per-query parameters for one mixed launch. filter: index of a bitmap in
 * the filter table passed alongside, or 0xFFFFFFFF for none. */
typedef struct hh_query_spec {
  uint32_t k, ef, precision, filter;
} hh_query_spec;

/* Mixed batch: query i uses specs[i] (device array); its results start at
 * out_keys + i*k_stride (k_stride >= every spec's k); slots past
 * out_count[i] are left as they were.  ef_max bounds
 * max(ef, k) over the batch and sizes the launch (0 = 512, the largest
 * supported); a query above it gets HH_UNSUPPORTED.  filters: device array
 * of nfilters bitmaps of ceil(n/32) words each, or NULL. */
int hh_search_batch_spec_device(hh_index *idx, const float *queries, uint64_t nq,
                                const hh_query_spec *specs, uint32_t k_stride, uint32_t ef_max,
                                const uint32_t *filters, uint32_t nfilters, uint64_t *out_keys,
                                float *out_dist, uint32_t *out_count, int32_t *status,
                                void *stream);

/* Host-memory form of the mixed batch.  Unlike hh_search_batch, an invalid
 * query does not fail the call: its hh_status goes to status[i] (host, nq
 * entries) and it gets no results, as a serving loop needs. */
int hh_search_batch_spec(hh_index *idx, const float *queries, uint64_t nq,
                         const hh_query_spec *specs, uint32_t k_stride, const uint32_t *filters,
                         uint32_t nfilters, uint64_t *out_keys, float *out_dist,
                         uint32_t *out_count, int32_t *status);

/* A search whose launch (query preparation + the persistent search kernel)
 * is captured once in a HIP graph and replayed: for callers that issue the
 * same batch shape repeatedly on fixed device buffers.  Arguments as in
 * hh_search_batch_device (specs != NULL selects the mixed form, as
 * hh_search_batch_spec_device; then k, precision and allow_bits are ignored
 * and ef is the batch's ef_max).  The plan owns its workspace; its launches are ordered with an
 * event, so launches of one plan never overlap.  A plan captures the index
 * as it is at creation: after any change of the index (hh_index_sync,
 * hh_index_update_deleted, hh_index_enable_bf16, hh_index_quantize,
 * hh_index_quantize_perdim) hh_search_plan_launch returns
 * HH_INVALID_ARGUMENT and launches nothing; create the plan again. */
typedef struct hh_search_plan hh_search_plan;
int hh_search_plan_create(hh_index *idx, const float *queries, uint64_t nq, uint32_t k,
                          uint32_t ef, uint32_t precision, const uint32_t *allow_bits,
                          const hh_query_spec *specs, uint32_t k_stride, const uint32_t *filters,
                          uint32_t nfilters, uint64_t *out_keys, float *out_dist,
                          uint32_t *out_count, int32_t *status, hh_search_plan **out);
int hh_search_plan_launch(hh_search_plan *plan, void *stream);
void hh_search_plan_free(hh_search_plan *plan);

/* Synchronous host-memory form of hh_search_batch_device with the crate's
 * error behaviour: on an invalid query returns its status and writes its
 * index to *first_bad (src/hnsw/mod.rs:513-536); allow_bits is a host
 * bitmap or NULL. */
int hh_search_batch(hh_index *idx, const float *queries, uint64_t nq, uint32_t k, uint32_t ef,
                    uint32_t precision, const uint32_t *allow_bits, uint64_t *out_keys,
                    float *out_dist, uint32_t *out_count, uint64_t *first_bad);

/* ------------------------------------------------------------------ */
/* Serving loop (no crate counterpart: continuous batching of single     */
/* requests over a resident index).                                      */
/* ------------------------------------------------------------------ */
typedef struct hh_server hh_server;
typedef struct hh_serve_config {
  uint32_t max_batch;     /* queries admitted per launch (0: resident waves) */
  uint32_t max_inflight;  /* launches in flight at once, 1..4 (0: 2) */
  uint32_t k_max;         /* largest k a request may ask for (1..512) */
  uint32_t max_filters;   /* size of the registered-filter table */
  uint64_t max_queue;     /* waiting requests before submit refuses (0: 64 batches) */
} hh_serve_config;

/* Starts the request loop on its own host thread and streams. */
int hh_serve_open(hh_index *idx, const hh_serve_config *cfg, hh_server **out);
/* Uploads a filter bitmap (over node ids) for requests to reference by id
 * in hh_query_spec.filter. */
int hh_serve_register_filter(hh_server *s, const uint32_t *allow_bits, uint32_t *filter_id);
/* Queues one query (host memory, dim floats) with its spec; returns a
 * ticket.  HH_CAPACITY_EXCEEDED when the queue is full. */
int hh_serve_submit(hh_server *s, const float *query, const hh_query_spec *spec, uint64_t *ticket);
/* Result of a ticket: *done = 0 while pending (or blocks when wait != 0).
 * keys/dist receive up to spec.k entries; latency is submit-to-completion. */
int hh_serve_poll(hh_server *s, uint64_t ticket, int wait, uint64_t *keys, float *dist,
                  uint32_t *count, int32_t *status, double *latency_ms, int *done);
/* Requests served, launches made and requests waiting; returns the first
 * launch error seen, if any. */
int hh_serve_stats(hh_server *s, uint64_t *served, uint64_t *batches, uint64_t *queued);
/* Load generator for measurement: submits request i (queries + i*dim,
 * specs[i]) at arrival_us[i] microseconds after the call starts, from a
 * native thread, and writes each request's submit-to-result latency.
 * Result arrays (optional) hold n*k_max keys/distances and n counts. */
int hh_serve_replay(hh_server *s, const float *queries, const hh_query_spec *specs, uint64_t n,
                    const double *arrival_us, double *latency_ms, uint64_t *out_keys,
                    float *out_dist, uint32_t *out_count, int32_t *out_status);
/* Drains every queued request, then stops the loop and frees the server. */
int hh_serve_close(hh_server *s);

/* ------------------------------------------------------------------ */
/* Exact brute force (BruteForceIndex, src/brute.rs:28-180).            */
/* ------------------------------------------------------------------ */
typedef struct hh_brute hh_brute;

/* BruteForceIndex::new + insert_batch: vectors are raw; they are validated
 * and (cosine) normalised on the host as src/vector.rs:69-87.  On a
 * duplicate key *bad_key receives it. */
int hh_brute_upload(uint32_t dim, uint32_t metric, const uint64_t *keys, const float *vectors,
                    uint64_t n, int device, hh_brute **out, uint64_t *bad_key);
void hh_brute_free(hh_brute *b);

/* BruteForceIndex::search_batch / search_filtered (src/brute.rs:131-180):
 * exact top-k under (distance, key), distances bit-identical to the
 * crate's AVX2 kernels.  k is clamped to n.  device pointers as in
 * hh_search_batch_device; allow_bits is over row positions. */
int hh_brute_search_batch_device(hh_brute *b, const float *queries, uint64_t nq, uint32_t k,
                                 const uint32_t *allow_bits, uint64_t *out_keys, float *out_dist,
                                 uint32_t *out_count, int32_t *status, void *stream);
/* The FP64 oracle: same ranking with every distance evaluated in f64. */
int hh_brute_search_batch_f64_device(hh_brute *b, const float *queries, uint64_t nq, uint32_t k,
                                     uint64_t *out_keys, double *out_dist, uint32_t *out_count,
                                     int32_t *status, void *stream);

/* Synchronous host-memory forms (src/brute.rs:162-180 error behaviour: an
 * invalid query's status is returned and its index written to *first_bad).
 * out arrays hold nq*k entries; allow_bits is a host bitmap over rows. */
int hh_brute_search_batch(hh_brute *b, const float *queries, uint64_t nq, uint32_t k,
                          const uint32_t *allow_bits, uint64_t *out_keys, float *out_dist,
                          uint32_t *out_count, uint64_t *first_bad);
int hh_brute_search_batch_f64(hh_brute *b, const float *queries, uint64_t nq, uint32_t k,
                              uint64_t *out_keys, double *out_dist, uint32_t *out_count,
                              uint64_t *first_bad);
uint64_t hh_brute_len(const hh_brute *b);
/* Execution choice for f32 batches (no counterpart in the reference, whose
 * brute force is one loop): 0 = exact streaming kernel only, 1 = automatic
 * (MFMA shortlist for unfiltered batches with n >= 32768, nq >= 32,
 * D <= 128, k <= 64), 2 = MFMA shortlist whenever D <= 128 and k <= 64.
 * Every mode returns the same results: the shortlist path proves each
 * query exact or re-runs it through the streaming kernel. */
int hh_brute_set_mfma(hh_brute *b, int mode);
/* For the last batch: queries that took the shortlist path and how many of
 * them were re-run exactly.  Waits for that batch. */
int hh_brute_mfma_stats(hh_brute *b, uint64_t *mfma_queries, uint64_t *redo_queries);

/* ------------------------------------------------------------------ */
/* Host helpers that mirror public crate items.                         */
/* ------------------------------------------------------------------ */

/* synthetic::Dataset::gaussian / clustered (src/synthetic.rs:81-124) and
 * the generators of the benchmark workloads (kind >= 2, no crate
 * counterpart): 2 = uniform-centre clusters (centres U[-1,1], sigma for
 * base rows, sigma_q for queries), 3 = U(0,1) rows then L2-normalised,
 * 4 = N(0,1) rows scaled by exp(sigma*N(0,1)), 5 = N(0,1) centres with
 * sigma spread.  categories (kind 4, optional): n category ids drawn
 * uniformly below `clusters`. */
int hh_synthetic(uint32_t kind, uint64_t n, uint64_t nq, uint32_t dim, uint32_t clusters,
                 double sigma, double sigma_q, uint64_t seed, int threads, float *out_base,
                 float *out_queries, uint32_t *categories);

/* Metric::distance with the CPU kernels Kernels::best() selects
 * (src/metric.rs:35-41, src/kernels/mod.rs:174-186): the AVX2+FMA
 * intrinsic sequence of src/kernels/x86.rs on x86 CPUs that have it, the
 * portable kernels otherwise.  Used by the CPU graph build. */
float hh_cpu_distance(uint32_t metric, const float *a, const float *b, uint64_t n);

/* eval::recall_at_k (src/eval.rs:25-47) over padded result arrays. */
double hh_recall_at_k(uint64_t nq, const uint64_t *truth, const uint32_t *truth_count,
                      uint32_t truth_stride, const uint64_t *found, const uint32_t *found_count,
                      uint32_t found_stride, uint32_t k);

#ifdef __cplusplus
}
#endif
#endif
