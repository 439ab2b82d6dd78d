// runtime.hip — host runtime of the engine: device residency of an index
// (upload, layout conversion), per-index workspace, kernel selection and
// launch, and the C ABI of include/hnsw_amd.h for the device side.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <map>
#include <tuple>
#include <cmath>
#include <cstring>
#include <condition_variable>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <unordered_set>
#include <vector>

#include "brute_kernels.h"
#include "hh_internal.h"
#include "quantize_kernels.h"
#include "search_kernels.h"

using hh::fail;

namespace {

#define HH_HIP(expr)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(e_ == hipErrorOutOfMemory ? HH_OUT_OF_MEMORY : HH_DEVICE_ERROR,       \
                  std::string(#expr) + ": " + hipGetErrorString(e_));                   \
  } while (0)

template <class T>
int dalloc(T **p, size_t count) {
  *p = nullptr;
  if (count == 0) return HH_OK;
  hipError_t e = hipMalloc((void **)p, count * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return fail(HH_OUT_OF_MEMORY, "device allocation of " + std::to_string(count * sizeof(T)) +
                                      " bytes failed: " + hipGetErrorString(e));
  }
  return HH_OK;
}
template <class T>
void dfree(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

// The search's row gather without its dependent work: each 16-lane
// quarter loads 4 rows per round (lane i: float4s at 4i + 64c), the rows
// drawn by a hash of (wave, round, slot) over [0, n), all loads of a round
// issued before any is used; the sums go to out so nothing is dropped.
__global__ void __launch_bounds__(256) gather_probe_kernel(const float *rows, uint64_t n,
                                                           uint32_t stride, uint32_t rounds,
                                                           uint64_t seed, float *out) {
  const uint32_t lane = threadIdx.x & 63, i = lane & 15, quarter = lane >> 4;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  float acc = 0.0f;
  for (uint32_t r = 0; r < rounds; r++) {
    const float *row[4];
#pragma unroll
    for (int b = 0; b < 4; b++) {
      uint64_t h = (wave * rounds + r) * 16 + 4 * b + quarter + seed;
      h ^= h >> 33; h *= 0xff51afd7ed558ccdull; h ^= h >> 33; h *= 0xc4ceb9fe1a85ec53ull;
      h ^= h >> 33;
      row[b] = rows + (h % n) * stride;
    }
    float4 v[4][4];
#pragma unroll
    for (int b = 0; b < 4; b++)
#pragma unroll
      for (int c = 0; c < 4; c++)
        v[b][c] = 4 * i + 64 * c < stride ? *(const float4 *)(row[b] + 4 * i + 64 * c)
                                          : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int b = 0; b < 4; b++)
#pragma unroll
      for (int c = 0; c < 4; c++) acc += v[b][c].x + v[b][c].y + v[b][c].z + v[b][c].w;
  }
  out[(uint64_t)blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

// natural row-major rows -> accumulator-major rows (RowLayout::pos)
__global__ void permute_rows_kernel(const float *src, float *dst, uint64_t n, uint32_t dim,
                                    uint32_t stride, int T) {
  const uint64_t total = n * dim;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t r = i / dim;
    const uint32_t e = (uint32_t)(i % dim);
    uint32_t pos = e;
    if (e < 32u * (uint32_t)T) {
      uint32_t t = e >> 5, j = e & 31, b = t >> 2;
      uint32_t tw = (uint32_t)T - 4 * b < 4 ? (uint32_t)T - 4 * b : 4;
      pos = b * 128 + j * tw + (t & 3);
    }
    dst[r * stride + pos] = src[i];
  }
}

// f32 -> bf16, round to nearest even (rows are finite).
__global__ void to_bf16_kernel(const float *src, uint16_t *dst, uint64_t count) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t u = __float_as_uint(src[i]);
    u += 0x7FFFu + ((u >> 16) & 1u);
    dst[i] = (uint16_t)(u >> 16);
  }
}

// Candidates a query may hold beyond its LDS list (8 bytes each, 512 KiB
// per resident wavefront).  Only a query with more live, not-yet-expanded
// candidates than this drops some (counted in n_overflow).  A selective
// filter keeps about ef / pass-rate of them live until its results fill:
// 10,240 at ef 512 and 5%, and 8,192 entries overflowed there.
constexpr uint32_t kCandPool = 65536;
// Nodes a query may mark in its HBM visited bitset (after leaving the LDS
// table) and still clear word by word at its end; beyond, the whole bitset.
constexpr uint32_t kSpillLog = 4096;
// Re-run pass of queries whose pool dropped a candidate: pools of n
// entries, as many as fit this many bytes (and at most kBigPoolSlots
// wavefronts); allocated on the first search that could need them.
constexpr uint64_t kBigPoolBytes = 1ull << 30;
constexpr uint64_t kBigPoolSlots = 256;
// search counters (hh_search_stats)
constexpr int kStatCount = 9;

// prepare_queries_kernel: four queries (one per wavefront) per block
int prep_grid(uint64_t nq) {
  return (int)std::max<uint64_t>(1, std::min<uint64_t>((nq + 3) / 4, 16384));
}

int grid_for(uint64_t work) {
  uint64_t g = (work + 255) / 256;
  return (int)std::max<uint64_t>(1, std::min<uint64_t>(g, 65536));
}

// Host rows (natural order) -> n device rows at `rows` in the row layout.
int upload_rows_into(const float *host_rows, uint64_t n, uint32_t dim, const hh::RowLayout &lay,
                     float *rows) {
  if (n == 0) return HH_OK;
  float *tmp = nullptr;
  int st = dalloc(&tmp, n * dim);
  if (st) return st;
  hipError_t e = hipMemcpy(tmp, host_rows, n * dim * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(rows, 0, n * lay.stride * sizeof(float));
  if (e == hipSuccess) {
    permute_rows_kernel<<<grid_for(n * dim), 256>>>(tmp, rows, n, dim, lay.stride, (int)lay.T);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  dfree(tmp);
  if (e != hipSuccess)
    return fail(HH_DEVICE_ERROR, std::string("row upload: ") + hipGetErrorString(e));
  return HH_OK;
}

int upload_rows(const float *host_rows, uint64_t n, uint32_t dim, const hh::RowLayout &lay,
                float **out) {
  *out = nullptr;
  if (n == 0) return HH_OK;
  float *rows = nullptr;
  int st = dalloc(&rows, n * lay.stride);
  if (!st) st = upload_rows_into(host_rows, n, dim, lay, rows);
  if (st) {
    dfree(rows);
    return st;
  }
  *out = rows;
  return HH_OK;
}

// Grows a device array holding `used` elements to `cap` elements, keeping
// its contents.
template <class T>
int grow(T **p, uint64_t used, uint64_t cap) {
  T *q = nullptr;
  int st = dalloc(&q, std::max<uint64_t>(cap, 1));
  if (st) return st;
  if (used && *p) {
    const hipError_t e = hipMemcpy(q, *p, used * sizeof(T), hipMemcpyDeviceToDevice);
    if (e != hipSuccess) {
      dfree(q);
      return fail(HH_DEVICE_ERROR, std::string("grow: ") + hipGetErrorString(e));
    }
  }
  dfree(*p);
  *p = q;
  return HH_OK;
}

// dst[ids[r] * width + j] = src[r * width + j]
__global__ void scatter_rows_kernel(const uint32_t *ids, const uint32_t *src, uint32_t *dst,
                                    uint32_t width, uint64_t count) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count * width;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t r = i / width;
    dst[(uint64_t)ids[r] * width + (i - r * width)] = src[i];
  }
}

int tmax_for(const hh::RowLayout &lay) {
  if (lay.T <= 4) return 4;
  if (lay.T <= 8) return 8;
  if (lay.T <= 32 && lay.dim <= 1024) return 32;
  return 0;
}
// LDS words per wavefront of the search kernel: the visited table, the
// id and distance staging (st entries each: 64, or 128 for beam steps), C
// (128 * sw candidate words) and W (64 * sw result words), candidate words
// being 2 words each.
uint32_t lds_words_for(uint32_t hash_cap, int sw, uint32_t st = 256) {
  return hash_cap + 2 * st + 384u * (uint32_t)sw;
}

int slots_for(uint32_t ef) {
  if (ef <= 64) return 1;
  if (ef <= 128) return 2;
  if (ef <= 256) return 4;
  if (ef <= 512) return 8;
  return 0;
}

// Index of the list-size bucket of ef (ef <= 64, 128, 256, 512), the unit
// the launch shape and order are tuned in; -1 above 512.
int ef_bucket(uint32_t ef) {
  const int sw = slots_for(ef);
  return sw == 1 ? 0 : sw == 2 ? 1 : sw == 4 ? 2 : sw == 8 ? 3 : -1;
}

// Per-index workspace; one search runs in it at a time (ordered by `done`).
struct Workspace {
  std::mutex mu;
  hipEvent_t done = nullptr;
  float *qprep = nullptr;
  int8_t *qcode = nullptr;
  int32_t *status = nullptr;
  uint64_t cap_q = 0;
  uint32_t q_stride = 0, q_cstride = 0;  // row and code strides cap_q was sized for
  uint32_t *counter = nullptr;
  unsigned long long *stats = nullptr;
  // longest-first query order of the two-pass launch
  uint64_t *seed = nullptr;
  uint32_t *order = nullptr;
  uint32_t *spill = nullptr;
  uint64_t *cpool = nullptr;  // per-wave candidate pools (pool_cap entries each)
  uint32_t pool_cap = 0;
  uint32_t *slog = nullptr;   // per-wave spill logs (kSpillLog entries)
  uint32_t *redo = nullptr;   // queries to re-run with the large pools (cap_q)
  float2 *qpd = nullptr;      // per-query (sigma, qconst) of per-dimension codes (cap_q)
  // Pools of `big_cap` >= n entries for the re-run of queries whose pool
  // dropped a candidate: big_slots of them (the re-run's wavefronts).
  uint64_t *bigpool = nullptr;
  uint32_t big_slots = 0, big_cap = 0;
  uint64_t spill_slots = 0;
  uint64_t spill_wcap = 0;   // bitset words per slot the allocation holds
  // brute-force partial lists
  float *part_d = nullptr;
  double *part_d64 = nullptr;
  uint64_t *part_k = nullptr;
  uint32_t *part_n = nullptr;
  uint64_t cap_part = 0, cap_partn = 0;
  // MFMA shortlist buffers
  float *m_dense = nullptr, *m_tau = nullptr, *m_qq = nullptr;
  uint32_t *m_cand = nullptr, *m_cnt = nullptr, *m_redo = nullptr;
  uint64_t cap_mq = 0, cap_dense = 0, cap_cand = 0;

  int ensure_mfma(uint64_t nq, uint64_t dense, uint64_t cand) {
    if (nq <= cap_mq && dense <= cap_dense && cand <= cap_cand) return HH_OK;
    HH_HIP(hipEventSynchronize(done));
    dfree(m_dense);
    dfree(m_tau);
    dfree(m_qq);
    dfree(m_cand);
    dfree(m_cnt);
    dfree(m_redo);
    cap_mq = std::max(nq, cap_mq);
    cap_dense = std::max(dense, cap_dense);
    cap_cand = std::max(cand, cap_cand);
    int st = dalloc(&m_dense, cap_dense);
    if (!st) st = dalloc(&m_tau, cap_mq);
    if (!st) st = dalloc(&m_qq, cap_mq);
    if (!st) st = dalloc(&m_cand, cap_cand);
    if (!st) st = dalloc(&m_cnt, cap_mq);
    if (!st) st = dalloc(&m_redo, cap_mq);
    if (st) cap_mq = cap_dense = cap_cand = 0;
    return st;
  }

  int init() {
    HH_HIP(hipEventCreateWithFlags(&done, hipEventDisableTiming));
    int st = dalloc(&counter, 4);
    if (!st) st = dalloc(&stats, (size_t)hhk::kStatStride * hhk::kStatLines);
    if (st) return st;
    HH_HIP(hipMemset(stats, 0, (size_t)hhk::kStatStride * hhk::kStatLines * sizeof(unsigned long long)));
    HH_HIP(hipMemset(counter, 0, 4 * sizeof(uint32_t)));
    return HH_OK;
  }
  void release() {
    if (done) (void)hipEventSynchronize(done);
    dfree(qprep);
    dfree(qcode);
    dfree(status);
    dfree(counter);
    dfree(stats);
    dfree(seed);
    dfree(order);
    dfree(spill);
    dfree(cpool);
    dfree(slog);
    dfree(redo);
    dfree(qpd);
    dfree(bigpool);
    dfree(part_d);
    dfree(part_d64);
    dfree(part_k);
    dfree(part_n);
    dfree(m_dense);
    dfree(m_tau);
    dfree(m_qq);
    dfree(m_cand);
    dfree(m_cnt);
    dfree(m_redo);
    if (done) (void)hipEventDestroy(done);
    done = nullptr;
  }
  // Per-query buffers for nq queries of `stride` floats and `cstride` code
  // bytes.  The code stride changes when a quantizer is trained on a
  // resident index, so it is part of the capacity.
  int ensure_queries(uint64_t nq, uint32_t stride, uint32_t cstride) {
    if (nq <= cap_q && stride <= q_stride && cstride <= q_cstride) return HH_OK;
    nq = std::max(nq, cap_q);
    stride = std::max(stride, q_stride);
    cstride = std::max(cstride, q_cstride);
    HH_HIP(hipEventSynchronize(done));
    dfree(qprep);
    dfree(qcode);
    dfree(status);
    dfree(seed);
    dfree(order);
    dfree(redo);
    dfree(qpd);
    uint64_t cap = std::max<uint64_t>(nq, 1024);
    int st = dalloc(&qprep, cap * stride);
    if (!st) st = dalloc(&qcode, cap * std::max<uint32_t>(cstride, 1));
    if (!st) st = dalloc(&status, cap);
    if (!st) st = dalloc(&seed, cap);
    if (!st) st = dalloc(&order, cap);
    if (!st) st = dalloc(&redo, cap);
    if (!st) st = dalloc(&qpd, cap);
    if (st) {
      cap_q = q_stride = q_cstride = 0;
      return st;
    }
    cap_q = cap;
    q_stride = stride;
    q_cstride = std::max<uint32_t>(cstride, 1);
    return HH_OK;
  }
  int ensure_spill(uint64_t slots, uint64_t words) {
    if (slots <= spill_slots && words <= spill_wcap) return HH_OK;
    slots = std::max(slots, spill_slots);
    words = std::max(words, spill_wcap);
    HH_HIP(hipEventSynchronize(done));
    dfree(spill);
    dfree(cpool);
    dfree(slog);
    // a query's candidates are distinct nodes, so a pool of n entries (the
    // bitset covers 32 * words >= n nodes) can never overflow
    const uint32_t pcap =
        (uint32_t)std::min<uint64_t>(kCandPool, (std::max<uint64_t>(32 * words, 64) + 63) / 64 * 64);
    int st = dalloc(&spill, slots * words);
    if (!st) st = dalloc(&cpool, slots * pcap);
    if (!st) st = dalloc(&slog, slots * kSpillLog);
    if (st) {
      spill_slots = 0;
      return st;
    }
    HH_HIP(hipMemset(spill, 0, slots * words * sizeof(uint32_t)));
    spill_slots = slots;
    spill_wcap = words;
    pool_cap = pcap;
    return HH_OK;
  }
  // Pools for the re-run pass: n entries each (a query's candidates are
  // distinct nodes, so such a pool never drops one), as many as fit
  // kBigPoolBytes, a multiple of four (the largest waves per block; the
  // re-run's grid is a whole number of blocks) and at least four.
  int ensure_bigpool(uint64_t n) {
    const uint64_t cap = (std::max<uint64_t>(n, 64) + 63) / 64 * 64;
    const uint64_t fit = std::min<uint64_t>(kBigPoolBytes / (cap * 8), kBigPoolSlots);
    const uint64_t want = std::max<uint64_t>(4, fit / 4 * 4);
    if (bigpool && cap <= big_cap && want <= big_slots) return HH_OK;
    HH_HIP(hipEventSynchronize(done));
    dfree(bigpool);
    big_slots = big_cap = 0;
    int st = dalloc(&bigpool, want * cap);
    if (st) return st;
    big_slots = (uint32_t)want;
    big_cap = (uint32_t)cap;
    return HH_OK;
  }
  int ensure_part(uint64_t entries, uint64_t lists, bool f64) {
    if (entries <= cap_part && lists <= cap_partn && (!f64 || part_d64)) return HH_OK;
    HH_HIP(hipEventSynchronize(done));
    dfree(part_d);
    dfree(part_d64);
    dfree(part_k);
    dfree(part_n);
    uint64_t ce = std::max(entries, cap_part), cl = std::max(lists, cap_partn);
    int st = dalloc(&part_d, ce);
    if (!st) st = dalloc(&part_d64, ce);
    if (!st) st = dalloc(&part_k, ce);
    if (!st) st = dalloc(&part_n, cl);
    if (st) {
      cap_part = cap_partn = 0;
      return st;
    }
    cap_part = ce;
    cap_partn = cl;
    return HH_OK;
  }
};

int num_cus(int device) {
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
      cus <= 0)
    cus = 256;
  return cus;
}

// Resident blocks per CU of a kernel at a block size and LDS size, cached
// per device: launch paths ask for every batch, and serving launches often.
int occupancy_blocks(int device, const void *kern, int threads, size_t lds, int *out) {
  static std::mutex mu;
  static std::map<std::tuple<int, const void *, int, size_t>, int> cache;
  const auto key = std::make_tuple(device, kern, threads, lds);
  {
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(key);
    if (it != cache.end()) {
      *out = it->second;
      return HH_OK;
    }
  }
  int b = 0;
  HH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, kern, threads, lds));
  std::lock_guard<std::mutex> lock(mu);
  cache[key] = b;
  *out = b;
  return HH_OK;
}

}  // namespace

struct hh_index {
  int device = 0;
  uint32_t dim = 0, metric = 0, m = 0, m0 = 0;
  uint64_t n = 0, live = 0;
  uint32_t entry = 0;
  int32_t top = -1;
  hh::RowLayout lay;
  uint32_t cstride = 0;
  float *rows = nullptr;
  uint16_t *bf16 = nullptr;
  int8_t *codes = nullptr;
  uint32_t *adj0 = nullptr;
  uint64_t *upper_off = nullptr;
  uint32_t *upper = nullptr;
  uint64_t *keys = nullptr;
  uint32_t *deleted_bits = nullptr;
  float *sq_offset = nullptr;
  float sq_scale = 1.0f;
  std::vector<float> sq_offset_host;  // the resident quantizer's offsets
  // per-dimension min/max codes (hh_index_quantize_perdim)
  int8_t *pd_codes = nullptr;
  float *pd_rnorm = nullptr;          // L2 only
  float *pd_mid = nullptr, *pd_scale = nullptr;
  // Searches hold mu shared while they read the index's arrays into a
  // launch; every change of the resident index (sync, quantizers, bf16
  // rows, tombstones) holds it exclusively, waits for the launches in
  // flight and bumps `generation`, which graph-captured plans check.
  // Changes are refused while a serving loop is open on the index.
  std::shared_mutex mu;
  std::atomic<uint64_t> generation{0};
  std::atomic<int> servers{0};
  // capacities for hh_index_sync (per-node arrays, upper link words)
  uint64_t cap_n = 0, upper_words = 0, cap_words = 0;
  // Workspaces: one search runs in a workspace at a time; calls on
  // different streams take different workspaces, so their kernels overlap.
  std::mutex pool_mu;
  std::condition_variable pool_cv;
  std::vector<std::unique_ptr<Workspace>> pool;
  std::vector<bool> pool_busy;
  Workspace stats_ws;  // holds the stats counters shared by every launch
  // launch shape chosen by hh_index_autotune per ef bucket (index log2 of
  // the list slots: ef <= 64, 128, 256, 512); 0 = rule of thumb
  uint32_t tuned_hash_cap[4] = {0, 0, 0, 0};
  int tuned_wpb[4] = {0, 0, 0, 0};
  hipStream_t stream = nullptr;  // used by the synchronous host-memory API
  // optional timing of the search kernel (hh_index_set_timing)
  bool timing = false;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  // two-pass launches that start the longest queries first
  bool longest_first = true;
  // candidates per layer-0 step of the bf16 and per-dimension int8
  // traversals (hh_index_set_beam_width)
  uint32_t beam_width = 4;
  // per ef bucket, the order hh_index_autotune measured faster: 0 = the
  // rule of thumb, 1 = one pass, 2 = two passes
  int tuned_order[4] = {0, 0, 0, 0};

  ~hh_index() {
    (void)hipSetDevice(device);
    for (auto &w : pool) w->release();
    stats_ws.release();
    dfree(rows);
    dfree(bf16);
    dfree(codes);
    dfree(adj0);
    dfree(upper_off);
    dfree(upper);
    dfree(keys);
    dfree(deleted_bits);
    dfree(sq_offset);
    dfree(pd_codes);
    dfree(pd_rnorm);
    dfree(pd_mid);
    dfree(pd_scale);
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

struct hh_brute {
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t dim = 0, metric = 0;
  uint64_t n = 0;
  hh::RowLayout lay;
  float *rows = nullptr;
  uint64_t *keys = nullptr;
  float *xx = nullptr;       // per-row squared norms (f32), MFMA L2 form
  double max_norm = 0.0;     // largest prepared row norm
  int mfma_mode = 1;          // 0 exact streaming only, 1 automatic, 2 MFMA when supported
  uint64_t last_mfma = 0;    // queries of the last batch that took the MFMA path
  Workspace ws;
  ~hh_brute() {
    (void)hipSetDevice(device);
    ws.release();
    dfree(rows);
    dfree(keys);
    dfree(xx);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

int pack_deleted(const uint8_t *deleted, uint64_t n, std::vector<uint32_t> &bits) {
  bits.assign((n + 31) / 32, 0u);
  if (deleted)
    for (uint64_t i = 0; i < n; i++)
      if (deleted[i]) bits[i >> 5] |= 1u << (i & 31);
  return HH_OK;
}

template <int TMAX, bool EX, bool XA, int SW>
void *search_kernel_ptr(int pc) {
  switch (pc) {
    case hhk::kPrecBf16: return reinterpret_cast<void *>(&hhk::hnsw_search_kernel<TMAX, EX, XA, SW, hhk::kPrecBf16>);
    case hhk::kPrecCodes: return reinterpret_cast<void *>(&hhk::hnsw_search_kernel<TMAX, EX, XA, SW, hhk::kPrecCodes>);
    case hhk::kPrecF32: return reinterpret_cast<void *>(&hhk::hnsw_search_kernel<TMAX, EX, XA, SW, hhk::kPrecF32>);
    case hhk::kPrecI8: return reinterpret_cast<void *>(&hhk::hnsw_search_kernel<TMAX, EX, XA, SW, hhk::kPrecI8>);
    default: return reinterpret_cast<void *>(&hhk::hnsw_search_kernel<TMAX, EX, XA, SW, hhk::kPrecAny>);
  }
}

// The search kernel instance for a row layout, list size and precision class.
void *pick_search_kernel(const hh::RowLayout &lay, int sw, int pc) {
  const int tmax = tmax_for(lay);
  const bool extra = lay.U > 0 || lay.R > 0;
  const bool exact = !extra && (int)lay.T == tmax;
#define HH_PICK(TM, EX, XA)                                        \
  if (tmax == TM && extra == EX && exact == XA) {                  \
    if (sw == 1) return search_kernel_ptr<TM, EX, XA, 1>(pc);      \
    if (sw == 2) return search_kernel_ptr<TM, EX, XA, 2>(pc);      \
    if (sw == 4) return search_kernel_ptr<TM, EX, XA, 4>(pc);      \
    if (sw == 8) return search_kernel_ptr<TM, EX, XA, 8>(pc);      \
  }
  HH_PICK(4, false, false)
  HH_PICK(4, true, false)
  HH_PICK(4, false, true)
  HH_PICK(8, false, false)
  HH_PICK(8, true, false)
  HH_PICK(8, false, true)
  HH_PICK(32, false, false)
  HH_PICK(32, true, false)
  HH_PICK(32, false, true)
#undef HH_PICK
  return nullptr;
}

// Reports the lowest-index failing query of a batch, as an error.
int first_failure(const std::vector<int32_t> &status, uint64_t *first_bad) {
  for (uint64_t i = 0; i < status.size(); i++)
    if (status[i]) {
      if (first_bad) *first_bad = i;
      static const char *msgs[] = {"", "dimension mismatch", "invalid dimension",
                                   "invalid parameter", "vector contains NaN or infinite values",
                                   "zero vector has no direction under cosine distance",
                                   "duplicate key", "capacity exceeded",
                                   "quantized search requested but quantization is not enabled"};
      std::string m = status[i] < 9 ? msgs[status[i]] : "invalid query parameters";
      return fail(status[i], m + " (query " + std::to_string(i) + ")");
    }
  return HH_OK;
}

// Candidates a beam step expands: the index's width, at most 8 and at most
// 256 neighbour slots per step; lists longer than 128 slots keep the
// crate's one-candidate rule.
uint32_t beam_steps(const hh_index *idx) {
  if (idx->beam_width <= 1 || idx->m0 == 0 || idx->m0 > 128) return 1;
  return std::min<uint32_t>(std::min<uint32_t>(idx->beam_width, 8u), 256u / idx->m0);
}

// Launch shape of one search: kernel instance, visited-set capacity,
// waves per block and grid size.
struct SearchConfig {
  void *kern = nullptr;
  uint32_t hash_cap = 0, lds_words = 0, st_cap = 64;
  uint32_t beam = 1;  // candidates per layer-0 step of bf16 / per-dimension queries
  int wpb = 1;
  uint64_t blocks = 0, spill_words = 0;
};

struct SearchCall {
  const float *queries;
  uint64_t nq;
  uint32_t k, ef, precision;
  const uint32_t *allow;
  const hh_query_spec *specs;
  uint32_t k_stride;
  const uint32_t *filters;
  uint32_t nfilters;
  uint64_t *out_keys;
  float *out_dist;
  uint32_t *out_count;
  int32_t *status;
  // construction candidates (hh_index_build_candidates): the queries are
  // the resident rows [rows_first, rows_first + nq), results are node ids,
  // and tombstoned nodes are candidates like any other
  bool build = false;
  uint64_t rows_first = 0;
  // words per filter bitmap in `filters` (0: the index's (n + 31) / 32)
  uint64_t filter_words = 0;
};

// The precision class of a call: a uniform batch runs the instance of its
// precision (no registers or code for the others in its loop; at 1M x 128,
// ef=64, the f32-only instance ran 9.4 M QPS against 8.95 M for the general
// one, the crate-int8 one 11.05 M against 10.6 M).  The bf16 and
// per-dimension instances take beam steps only: at beam width 1 (or lists
// too long for a step) those batches run the general instance, as do
// construction searches and mixed batches.
int precision_class(const hh_index *idx, const SearchCall &c) {
  if (c.specs || c.build) return hhk::kPrecAny;
  const bool beam = beam_steps(idx) > 1;
  if (c.precision == 2) return beam ? hhk::kPrecBf16 : hhk::kPrecAny;
  if (c.precision == 3) return beam ? hhk::kPrecCodes : hhk::kPrecAny;
  if (c.precision == 1) return hhk::kPrecI8;
  return hhk::kPrecF32;
}

// Largest max(ef, k) a launch must serve: the call's own for a uniform
// batch; for a mixed batch the caller's bound (c.ef, 0 = 512).  Mixed-batch
// queries above the bound fail with HH_UNSUPPORTED in the prepare kernel.
uint32_t spec_ef_bound(const SearchCall &c) {
  if (!c.specs) return std::max(c.ef, c.k);
  return c.ef == 0 ? 512u : std::min<uint32_t>(c.ef, 512u);
}

int check_call(hh_index *idx, const SearchCall &c) {
  if (!c.queries || !c.out_count || (c.k_stride > 0 && (!c.out_keys || !c.out_dist)))
    return fail(HH_INVALID_ARGUMENT, "null device pointer");
  if (!tmax_for(idx->lay)) return fail(HH_UNSUPPORTED, "search kernels support dim <= 1024");
  if (!c.specs && c.precision > 3) return fail(HH_INVALID_ARGUMENT, "unknown precision");
  const uint32_t efmax = spec_ef_bound(c);
  if (!slots_for(std::max<uint32_t>(efmax, 1)))
    return fail(HH_UNSUPPORTED, "ef (raised to k) above 512 is not supported");
  if (!c.specs && c.precision == 2 && !idx->bf16)
    return fail(HH_UNSUPPORTED, "bf16 rows not resident: call hh_index_enable_bf16");
  return HH_OK;
}

int search_config(hh_index *idx, const SearchCall &c, SearchConfig &cfg) {
  const uint32_t efmax = std::max<uint32_t>(spec_ef_bound(c), 1);
  const int sw = slots_for(efmax);
  // visited-set capacity: 2560 entries (10 KiB of LDS per wavefront; with
  // the lists and staging at ef <= 64, 12-12.5 KiB, which keeps the
  // register-limited 12 wavefronts per CU resident, and holds the result
  // arrays afterwards); past 3/4 full a query continues on its HBM bitset.
  // Measured at 1M x 128, ef=64: 2048 entries sent 65 of 10,000 f32 queries
  // and 611 per-dimension beam queries to the bitset, 2560 sent 11 and 177,
  // +1% f32 and +9% per-dimension QPS.  hh_index_autotune trades this
  // against occupancy per ef.
  const int bucket = sw == 1 ? 0 : sw == 2 ? 1 : sw == 4 ? 2 : 3;
  cfg.hash_cap = idx->tuned_hash_cap[bucket] ? idx->tuned_hash_cap[bucket] : 2560u;
  cfg.hash_cap = std::max<uint32_t>(cfg.hash_cap, 256u * (uint32_t)sw);
  // beam steps stage every neighbour slot of a step (up to 256)
  const uint32_t bs = beam_steps(idx);
  cfg.beam = bs > 1 && !c.build && (c.specs || c.precision == 2 || c.precision == 3) ? bs : 1u;
  cfg.st_cap = cfg.beam > 1 ? std::max<uint32_t>(64u, (cfg.beam * idx->m0 + 63) / 64 * 64) : 64u;
  cfg.lds_words = lds_words_for(cfg.hash_cap, sw, cfg.st_cap);
  cfg.kern = pick_search_kernel(idx->lay, sw, precision_class(idx, c));
  // waves per block: the tuned value, else whichever of 1, 2, 4 keeps the
  // most waves resident
  int blocks_per_cu = 1, best_waves = 0;
  for (int w : {4, 2, 1}) {
    if (idx->tuned_wpb[bucket] && w != idx->tuned_wpb[bucket]) continue;
    if (cfg.lds_words * 4 * w > 65536) continue;
    int b = 0;
    const int st = occupancy_blocks(idx->device, cfg.kern, w * 64, cfg.lds_words * 4 * w, &b);
    if (st) return st;
    if (b * w > best_waves) {
      best_waves = b * w;
      cfg.wpb = w;
      blocks_per_cu = b;
    }
  }
  blocks_per_cu = std::max(1, blocks_per_cu);
  const uint64_t want = (c.nq + cfg.wpb - 1) / cfg.wpb;
  cfg.blocks = std::min<uint64_t>(want, (uint64_t)blocks_per_cu * num_cus(idx->device));
  cfg.spill_words = std::max<uint64_t>(1, (idx->n + 31) / 32);
  return HH_OK;
}

// The descent pass as descent_kernel: rows of T 4-blocks take the TMAX
// instance that holds them, the exact layouts (D = 32 * TMAX) the unrolled
// one, m <= 16 sixteen lanes per query and m <= 32 thirty-two.  Uniform
// f32 and bf16 batches take thirty-two lanes at any m: two quarters share
// a step's rows, so a step waits for half as many load rounds and twice
// the waves hide the latency (+1% QPS at 1M x 128, m = 16; int8 code rows,
// one per lane, lose 2-3% with idle lanes and keep sixteen).
int launch_descent(const hh_index *idx, uint64_t nq, const hhk::SearchArgs &a,
                   hipStream_t stream) {
  using K = void (*)(const hhk::SearchArgs);
  const int tmax = tmax_for(idx->lay);
  // (cstride is 0 until a quantizer exists; with one it is dim rounded up to
  // 128, which the exact instances take as a compile-time length)
  const bool exact = idx->lay.U == 0 && idx->lay.R == 0 && (int)idx->lay.T == tmax &&
                     (idx->cstride == 0 || idx->cstride == 32u * (uint32_t)tmax);
  const bool row_batch = !a.specs && (a.precision == 0 || a.precision == 2);
  const int wide = idx->m > 16 || row_batch;
  const int ti = tmax == 4 ? 0 : tmax == 8 ? 1 : 2;
  static const K table[3][2][2] = {
      {{hhk::descent_kernel<4, 16, false>, hhk::descent_kernel<4, 16, true>},
       {hhk::descent_kernel<4, 32, false>, hhk::descent_kernel<4, 32, true>}},
      {{hhk::descent_kernel<8, 16, false>, hhk::descent_kernel<8, 16, true>},
       {hhk::descent_kernel<8, 32, false>, hhk::descent_kernel<8, 32, true>}},
      {{hhk::descent_kernel<32, 16, false>, hhk::descent_kernel<32, 16, true>},
       {hhk::descent_kernel<32, 32, false>, hhk::descent_kernel<32, 32, true>}}};
  const int threads = tmax <= 8 ? hhk::descent_threads<8>() : hhk::descent_threads<32>();
  const int stride_cap = tmax == 4 ? hhk::descent_stride<4>()
                         : tmax == 8 ? hhk::descent_stride<8>() : hhk::descent_stride<32>();
  const int code_cap = tmax == 4 ? hhk::descent_code_words<4>()
                       : tmax == 8 ? hhk::descent_code_words<8>() : hhk::descent_code_words<32>();
  if (!tmax || idx->m > 32 || idx->lay.stride > (uint32_t)stride_cap ||
      idx->cstride / 8 > (uint32_t)code_cap)
    return fail(HH_UNSUPPORTED, "descent kernel: layout outside its slot sizes");
  const uint32_t per_block = (uint32_t)threads / (wide ? 32u : 16u);
  const dim3 grid((unsigned)((nq + per_block - 1) / per_block)), block((unsigned)threads);
  hipLaunchKernelGGL(table[ti][wide][exact], grid, block, 0, stream, a);
  HH_HIP(hipGetLastError());
  return HH_OK;
}

// Enqueues the prepare and search kernels of one call on `stream`, using
// workspace `ws` (whose buffers the caller has sized).  No synchronisation.
int enqueue_search(hh_index *idx, Workspace &ws, const SearchCall &c, const SearchConfig &cfg,
                   hipStream_t stream, hipEvent_t before_search = nullptr) {
  hhk::PrepArgs pa{};
  pa.queries = c.queries;
  pa.nq = c.nq;
  pa.dim = idx->dim;
  pa.metric = idx->metric;
  pa.T = (int)idx->lay.T;
  pa.stride = idx->lay.stride;
  pa.cstride = idx->cstride;
  pa.qprep = ws.qprep;
  pa.qcode = idx->codes || idx->pd_codes ? ws.qcode : nullptr;
  pa.sq_offset = idx->sq_offset;
  pa.sq_scale = idx->sq_scale;
  pa.pd_mid = idx->pd_mid;
  pa.pd_scale = idx->pd_scale;
  pa.qpd = ws.qpd;
  pa.has_pd = idx->pd_codes != nullptr;
  pa.precision = c.precision;
  pa.specs = c.specs;
  pa.k_stride = c.k_stride;
  pa.nfilters = c.nfilters;
  pa.ef_cap = c.specs ? spec_ef_bound(c) : 512;
  pa.has_codes = idx->codes != nullptr;
  pa.has_bf16 = idx->bf16 != nullptr;
  pa.status = c.status ? c.status : ws.status;
  pa.counter = ws.counter;
  if (c.build) {
    // stored rows are already prepared and in the row layout
    HH_HIP(hipMemcpyAsync(ws.qprep, idx->rows + c.rows_first * idx->lay.stride,
                          c.nq * idx->lay.stride * sizeof(float), hipMemcpyDeviceToDevice, stream));
    hipLaunchKernelGGL(hhk::fill_u32_kernel, dim3(grid_for(c.nq)), dim3(256), 0, stream,
                       reinterpret_cast<uint32_t *>(pa.status), c.nq, 0u);
    hipLaunchKernelGGL(hhk::fill_u32_kernel, dim3(1), dim3(64), 0, stream, ws.counter,
                       (uint64_t)4, 0u);
  } else {
    hipLaunchKernelGGL(hhk::prepare_queries_kernel, dim3(prep_grid(c.nq)), dim3(256), 0, stream,
                       pa);
  }
  HH_HIP(hipGetLastError());

  hhk::SearchArgs a{};
  a.rows = idx->rows;
  a.bf16 = idx->bf16;
  a.codes = idx->codes;
  a.adj0 = idx->adj0;
  a.upper_off = idx->upper_off;
  a.upper = idx->upper;
  a.keys = idx->keys;
  // the tombstone bitmap is only read when something is deleted
  a.deleted_bits = idx->live < idx->n && !c.build ? idx->deleted_bits : nullptr;
  a.ids_out = c.build ? 1u : 0u;
  a.n = (uint32_t)idx->n;
  a.m = idx->m;
  a.m0 = idx->m0;
  a.metric = idx->metric;
  a.entry = idx->entry;
  a.top = idx->top;
  a.live = idx->live;
  a.T = (int)idx->lay.T;
  a.U = (int)idx->lay.U;
  a.R = (int)idx->lay.R;
  a.stride = idx->lay.stride;
  a.cstride = idx->cstride;
  a.sq_scale = idx->sq_scale;
  a.pd_codes = idx->pd_codes;
  a.pd_rnorm = idx->pd_rnorm;
  a.qprep = ws.qprep;
  a.qcode = ws.qcode;
  a.qpd = ws.qpd;
  a.status = pa.status;
  a.nq = c.nq;
  a.specs = c.specs;
  a.k = c.k;
  a.ef = c.ef;
  a.precision = c.precision;
  a.allow = c.allow;
  a.filters = c.filters;
  a.filter_words = (uint32_t)(c.filter_words ? c.filter_words : (idx->n + 31) / 32);
  a.k_stride = c.k_stride;
  a.out_keys = c.out_keys;
  a.out_dist = c.out_dist;
  a.out_count = c.out_count;
  a.counter = ws.counter;
  a.spill = ws.spill;
  a.spill_words = cfg.spill_words;
  a.cpool = ws.cpool;
  a.cpool_cap = ws.pool_cap;
  a.spill_log = ws.slog;
  a.spill_log_cap = kSpillLog;
  a.stats = idx->stats_ws.stats;
  a.hash_cap = cfg.hash_cap;
  a.hash_limit = cfg.hash_cap / 2 + cfg.hash_cap / 4;
  a.lds_words = cfg.lds_words;
  a.st_cap = cfg.st_cap;
  a.beam = cfg.beam;
  a.seed = ws.seed;
  a.order = ws.order;
  // a pool smaller than the node count can drop a candidate: such queries
  // are listed and re-run below with pools of n entries
  const bool redo = ws.pool_cap < idx->n && ws.bigpool;
  a.redo = redo ? ws.redo : nullptr;
  a.redo_count = ws.counter + 2;
  void *args[] = {&a};
  if (before_search) HH_HIP(hipEventRecord(before_search, stream));
  const dim3 grid((unsigned)cfg.blocks), block(cfg.wpb * 64);
  const size_t lds = cfg.lds_words * 4 * cfg.wpb;
  // With a few queries per resident wave, the kernel's end is set by the
  // long queries that start last.  Two passes start them first: the upper-
  // layer descent of every query, an order by descending descent distance
  // (which tracks the layer-0 work), then layer 0 in that order.  Results
  // do not depend on the order.
  const uint64_t waves = cfg.blocks * (uint64_t)cfg.wpb;
  // The descent pass runs as descent_kernel (one query per 16 or 32 lanes)
  // where it applies: m <= 32, search batches (not construction).
  const bool fast_descent = idx->m <= 32 && !c.build;
  // Rule of thumb, measured at 10k queries: on 1M x 128 clustered L2 rows
  // (m = 16) the order pays from ef = 32 (ef=64 +29%); on 200k x 960
  // cosine (m = 32) and 500k x 256 inner-product (m = 24) rows it is
  // within -3%..+2% of one pass at ef 64 and 256.  hh_index_autotune
  // replaces the rule by a measurement.
  const uint32_t efb = spec_ef_bound(c);
  const int tmax = tmax_for(idx->lay);
  const bool narrow = idx->m <= 16 && idx->lay.U == 0 && idx->lay.R == 0 &&
                      (int)idx->lay.T == tmax && tmax <= 8;
  const int eb = ef_bucket(std::max<uint32_t>(efb, 1));
  const int tuned = eb >= 0 ? idx->tuned_order[eb] : 0;
  const bool two_pass = tuned ? tuned == 2 : ((fast_descent && narrow && efb > 16) || efb > 64);
  // Mixed batches are ordered by their estimated work (ef over the filter's
  // pass rate: most of their cost spread), which pays however many queries
  // each wave takes.
  if (idx->longest_first && idx->top >= 1 && c.nq > waves &&
      (c.nq <= 16 * waves || (c.specs && !c.build)) && two_pass) {
    a.phase = 1;
    if (fast_descent) {
      if (int st = launch_descent(idx, c.nq, a, stream)) return st;
    } else {
      HH_HIP(hipLaunchKernel(cfg.kern, grid, block, args, lds, stream));
    }
    hipLaunchKernelGGL(hhk::order_queries_kernel, dim3(1), dim3(1024), 0, stream, ws.seed,
                       c.specs, c.nq, c.filters, a.filter_words, c.filters ? c.nfilters : 0u,
                       ws.order, ws.counter);
    HH_HIP(hipGetLastError());
    a.phase = 2;
  }
  HH_HIP(hipLaunchKernel(cfg.kern, grid, block, args, lds, stream));
  if (redo) {
    // the re-run: layer 0 of the listed queries from their seeds, on
    // big_slots wavefronts whose pools hold n entries each
    const uint64_t slots = std::min<uint64_t>(ws.big_slots, waves);
    a.phase = 3;
    a.cpool = ws.bigpool;
    a.cpool_cap = ws.big_cap;
    const dim3 rgrid((unsigned)((slots + cfg.wpb - 1) / cfg.wpb));
    HH_HIP(hipLaunchKernel(cfg.kern, rgrid, block, args, lds, stream));
  }
  return HH_OK;
}

// Takes a workspace that no other host thread is enqueuing into (creating
// one, up to a small pool, when all are taken).
Workspace *acquire_ws(hh_index *idx, int &st) {
  std::unique_lock<std::mutex> lock(idx->pool_mu);
  constexpr size_t kMaxPool = 4;
  for (;;) {
    for (size_t i = 0; i < idx->pool.size(); i++)
      if (!idx->pool_busy[i] && hipEventQuery(idx->pool[i]->done) == hipSuccess) {
        idx->pool_busy[i] = true;
        return idx->pool[i].get();
      }
    if (idx->pool.size() < kMaxPool) {
      std::unique_ptr<Workspace> w(new Workspace());
      st = w->init();
      if (st) return nullptr;
      idx->pool.push_back(std::move(w));
      idx->pool_busy.push_back(true);
      return idx->pool.back().get();
    }
    for (size_t i = 0; i < idx->pool.size(); i++)
      if (!idx->pool_busy[i]) {
        idx->pool_busy[i] = true;  // still running on the GPU: the launch waits for its event
        return idx->pool[i].get();
      }
    idx->pool_cv.wait(lock);
  }
}

void release_ws(hh_index *idx, Workspace *w) {
  std::lock_guard<std::mutex> lock(idx->pool_mu);
  for (size_t i = 0; i < idx->pool.size(); i++)
    if (idx->pool[i].get() == w) idx->pool_busy[i] = false;
  idx->pool_cv.notify_one();
}

int launch_call(hh_index *idx, const SearchCall &c, hipStream_t stream);

// Shared launch path of the uniform and the per-query-spec searches.
int launch_search(hh_index *idx, const float *queries, uint64_t nq, uint32_t k, uint32_t ef,
                  uint32_t precision, const uint32_t *allow, const hh_query_spec *specs,
                  uint32_t k_stride, const uint32_t *filters, uint32_t nfilters,
                  uint64_t *out_keys, float *out_dist, uint32_t *out_count, int32_t *status,
                  hipStream_t stream) {
  if (nq == 0) return HH_OK;
  const SearchCall c{queries, nq, k, ef, precision, allow, specs, k_stride, filters, nfilters,
                     out_keys, out_dist, out_count, status};
  return launch_call(idx, c, stream);
}

int launch_call(hh_index *idx, const SearchCall &c, hipStream_t stream) {
  const uint64_t nq = c.nq;
  std::shared_lock<std::shared_mutex> rd(idx->mu);
  int st = check_call(idx, c);
  if (st) return st;
  HH_HIP(hipSetDevice(idx->device));
  SearchConfig cfg;
  st = search_config(idx, c, cfg);
  if (st) return st;
  Workspace *ws = acquire_ws(idx, st);
  if (!ws) return st;
  struct Release {
    hh_index *i;
    Workspace *w;
    ~Release() { release_ws(i, w); }
  } rel{idx, ws};
  st = ws->ensure_queries(nq, idx->lay.stride, idx->cstride);
  if (!st) st = ws->ensure_spill(cfg.blocks * cfg.wpb, cfg.spill_words);
  if (!st && ws->pool_cap < idx->n) st = ws->ensure_bigpool(idx->n);
  if (st) return st;
  HH_HIP(hipStreamWaitEvent(stream, ws->done, 0));
  st = enqueue_search(idx, *ws, c, cfg, stream, idx->timing ? idx->t0 : nullptr);
  if (st) return st;
  if (idx->timing) HH_HIP(hipEventRecord(idx->t1, stream));
  HH_HIP(hipEventRecord(ws->done, stream));
  return HH_OK;
}

}  // namespace

namespace hh {
int index_device(const hh_index *idx) { return idx->device; }
uint32_t index_dim(const hh_index *idx) { return idx->dim; }
uint64_t index_nodes(const hh_index *idx) { return idx->n; }
uint32_t index_resident_waves(hh_index *idx, uint32_t ef) {
  SearchCall c{};
  c.nq = 1u << 30;
  c.k = 1;
  c.ef = std::max<uint32_t>(1, std::min<uint32_t>(ef, 512));
  SearchConfig cfg;
  if (!tmax_for(idx->lay) || search_config(idx, c, cfg)) return 1024;
  return (uint32_t)std::max<uint64_t>(1, cfg.blocks * cfg.wpb);
}

void index_serving(hh_index *idx, int delta) { idx->servers += delta; }

int search_spec_device_strided(hh_index *idx, const float *queries, uint64_t nq,
                               const hh_query_spec *specs, uint32_t k_stride, uint32_t ef_max,
                               const uint32_t *filters, uint32_t nfilters, uint64_t filter_words,
                               uint64_t *out_keys, float *out_dist, uint32_t *out_count,
                               int32_t *status, void *stream) {
  if (!idx || !specs) return fail(HH_INVALID_ARGUMENT, "null argument");
  if (nq == 0) return HH_OK;
  if (filters && filter_words < (idx->n + 31) / 32)
    return fail(HH_INVALID_ARGUMENT, "filter bitmaps shorter than the index");
  SearchCall c{queries, nq, 0u, ef_max, 0u, nullptr, specs, k_stride, filters, nfilters,
               out_keys, out_dist, out_count, status};
  c.filter_words = filters ? filter_words : 0;
  return launch_call(idx, c, (hipStream_t)stream);
}

// Sizes `count` of the index's workspaces for launches of up to nq queries
// at the largest resident wave count, so that no launch of a serving loop
// reallocates (a device free synchronises the whole device).
int index_reserve(hh_index *idx, uint64_t nq, uint32_t count) {
  if (!tmax_for(idx->lay)) return HH_OK;
  HH_HIP(hipSetDevice(idx->device));
  const uint64_t slots = index_resident_waves(idx);
  const uint64_t words = std::max<uint64_t>(1, (idx->n + 31) / 32);
  std::vector<Workspace *> held;
  int st = HH_OK;
  for (uint32_t i = 0; i < count && !st; i++) {
    Workspace *w = acquire_ws(idx, st);
    if (!w) break;
    held.push_back(w);
    st = w->ensure_queries(nq, idx->lay.stride, idx->cstride);
    if (!st) st = w->ensure_spill(slots, words);
    if (!st && w->pool_cap < idx->n) st = w->ensure_bigpool(idx->n);
  }
  for (Workspace *w : held) release_ws(idx, w);
  return st;
}
}  // namespace hh

namespace {

// Changes of the resident index are refused while a serving loop is open
// on it (its launches read the arrays a change would move or free).
int refuse_if_served(hh_index *idx) {
  if (idx->servers.load() > 0)
    return fail(HH_INVALID_ARGUMENT,
                "a serving loop is open on this index: close it before changing the index");
  return HH_OK;
}

// Waits for every launch on the index's workspaces.
int drain(hh_index *idx) {
  std::lock_guard<std::mutex> lock(idx->pool_mu);
  for (auto &w : idx->pool) HH_HIP(hipEventSynchronize(w->done));
  return HH_OK;
}

}  // namespace

extern "C" {

int hh_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

int hh_device_alloc(int device, uint64_t bytes, void **out) {
  if (!out) return fail(HH_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  HH_HIP(hipSetDevice(device));
  if (bytes == 0) return HH_OK;
  if (hipMalloc(out, bytes) != hipSuccess) {
    (void)hipGetLastError();
    *out = nullptr;
    return fail(HH_OUT_OF_MEMORY, "device allocation of " + std::to_string(bytes) + " bytes failed");
  }
  return HH_OK;
}

int hh_device_free(void *p) {
  if (p) HH_HIP(hipFree(p));
  return HH_OK;
}

int hh_memcpy(void *dst, const void *src, uint64_t bytes, int kind, void *stream) {
  if (bytes == 0) return HH_OK;
  if (!dst || !src) return fail(HH_INVALID_ARGUMENT, "null pointer");
  hipMemcpyKind k;
  switch (kind) {
    case 0: k = hipMemcpyHostToDevice; break;
    case 1: k = hipMemcpyDeviceToHost; break;
    case 2: k = hipMemcpyDeviceToDevice; break;
    default: return fail(HH_INVALID_ARGUMENT, "unknown copy kind");
  }
  if (stream) HH_HIP(hipMemcpyAsync(dst, src, bytes, k, (hipStream_t)stream));
  else HH_HIP(hipMemcpy(dst, src, bytes, k));
  return HH_OK;
}

int hh_memset(void *dst, int value, uint64_t bytes) {
  if (bytes == 0) return HH_OK;
  if (!dst) return fail(HH_INVALID_ARGUMENT, "null pointer");
  HH_HIP(hipMemset(dst, value, bytes));
  return HH_OK;
}

int hh_stream_create(int device, void **out) {
  if (!out) return fail(HH_INVALID_ARGUMENT, "null argument");
  HH_HIP(hipSetDevice(device));
  hipStream_t s = nullptr;
  HH_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *out = s;
  return HH_OK;
}

int hh_stream_destroy(void *stream) {
  if (stream) HH_HIP(hipStreamDestroy((hipStream_t)stream));
  return HH_OK;
}

int hh_stream_synchronize(void *stream) {
  HH_HIP(hipStreamSynchronize((hipStream_t)stream));
  return HH_OK;
}

int hh_device_synchronize(int device) {
  HH_HIP(hipSetDevice(device));
  HH_HIP(hipDeviceSynchronize());
  return HH_OK;
}

int hh_index_upload(const hh_graph_desc *d, int device, hh_index **out) {
  if (!d || !out) return fail(HH_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (d->dim == 0 || d->dim > 65536) return fail(HH_INVALID_DIMENSION, "invalid dimension");
  if (d->metric > 2) return fail(HH_INVALID_ARGUMENT, "unknown metric");
  if (d->n > 0xFFFFFFFEull) return fail(HH_CAPACITY_EXCEEDED, "too many nodes");
  if (d->m < 1 || d->m0 < 1) return fail(HH_INVALID_PARAMETER, "m and m0 must be positive");
  if (d->n > 0 && (!d->vectors || !d->keys || !d->levels || !d->layer0 || !d->upper_off))
    return fail(HH_INVALID_ARGUMENT, "null graph array");
  if (d->top_level >= 0 && d->entry_id >= d->n)
    return fail(HH_INVALID_ARGUMENT, "entry point out of range");
  HH_HIP(hipSetDevice(device));
  std::unique_ptr<hh_index> idx(new hh_index());
  idx->device = device;
  idx->dim = d->dim;
  idx->metric = d->metric;
  idx->m = d->m;
  idx->m0 = d->m0;
  idx->n = d->n;
  idx->live = d->live;
  idx->entry = d->entry_id;
  idx->top = d->top_level;
  idx->lay = hh::RowLayout(d->dim);
  int st = idx->stats_ws.init();
  if (st) return st;
  HH_HIP(hipStreamCreateWithFlags(&idx->stream, hipStreamNonBlocking));
  if (d->n == 0) {
    // an empty index may still hold a quantizer (searches then succeed
    // with no results instead of reporting quantization disabled)
    if (d->codes) {
      if (!d->sq_offset) return fail(HH_INVALID_ARGUMENT, "codes without quantizer offsets");
      idx->cstride = (d->dim + 127) / 128 * 128;
      st = dalloc(&idx->codes, idx->cstride);
      if (!st) st = dalloc(&idx->sq_offset, d->dim);
      if (st) return st;
      HH_HIP(hipMemset(idx->codes, 0, idx->cstride));
      HH_HIP(hipMemcpy(idx->sq_offset, d->sq_offset, d->dim * 4, hipMemcpyHostToDevice));
      idx->sq_scale = d->sq_scale;
      idx->sq_offset_host.assign(d->sq_offset, d->sq_offset + d->dim);
    }
    *out = idx.release();
    return HH_OK;
  }
  const uint64_t n = d->n;
  st = upload_rows(d->vectors, n, d->dim, idx->lay, &idx->rows);
  if (st) return st;
  // layer 0: crate blocks (count, ids) -> n x m0 ids with unused slots
  // 0xFFFFFFFF (the count is implied by the first unused slot).
  {
    std::vector<uint32_t> ell((size_t)n * d->m0, 0xFFFFFFFFu);
    for (uint64_t i = 0; i < n; i++) {
      const uint32_t *blk = d->layer0 + i * (d->m0 + 1);
      uint32_t c = std::min(blk[0], d->m0);
      for (uint32_t s = 0; s < c; s++) {
        if (blk[1 + s] >= n) return fail(HH_INVALID_ARGUMENT, "layer-0 neighbour out of range");
        ell[i * d->m0 + s] = blk[1 + s];
      }
    }
    st = dalloc(&idx->adj0, ell.size());
    if (st) return st;
    HH_HIP(hipMemcpy(idx->adj0, ell.data(), ell.size() * 4, hipMemcpyHostToDevice));
  }
  {
    uint64_t words = 0;
    for (uint64_t i = 0; i < n; i++)
      words = std::max<uint64_t>(words, d->upper_off[i] + (uint64_t)d->levels[i] * (d->m + 1));
    st = dalloc(&idx->upper_off, n);
    if (!st) st = dalloc(&idx->upper, std::max<uint64_t>(words, 1));
    if (st) return st;
    idx->upper_words = words;
    idx->cap_words = std::max<uint64_t>(words, 1);
    HH_HIP(hipMemcpy(idx->upper_off, d->upper_off, n * 8, hipMemcpyHostToDevice));
    if (words) HH_HIP(hipMemcpy(idx->upper, d->upper, words * 4, hipMemcpyHostToDevice));
  }
  st = dalloc(&idx->keys, n);
  if (st) return st;
  HH_HIP(hipMemcpy(idx->keys, d->keys, n * 8, hipMemcpyHostToDevice));
  {
    std::vector<uint32_t> bits;
    pack_deleted(d->deleted, n, bits);
    st = dalloc(&idx->deleted_bits, bits.size());
    if (st) return st;
    HH_HIP(hipMemcpy(idx->deleted_bits, bits.data(), bits.size() * 4, hipMemcpyHostToDevice));
  }
  if (d->codes) {
    if (!d->sq_offset) return fail(HH_INVALID_ARGUMENT, "codes without quantizer offsets");
    idx->cstride = (d->dim + 127) / 128 * 128;
    std::vector<int8_t> packed((size_t)n * idx->cstride, 0);
    for (uint64_t i = 0; i < n; i++)
      std::memcpy(packed.data() + i * idx->cstride, d->codes + i * d->dim, d->dim);
    st = dalloc(&idx->codes, packed.size());
    if (!st) st = dalloc(&idx->sq_offset, d->dim);
    if (st) return st;
    HH_HIP(hipMemcpy(idx->codes, packed.data(), packed.size(), hipMemcpyHostToDevice));
    HH_HIP(hipMemcpy(idx->sq_offset, d->sq_offset, d->dim * 4, hipMemcpyHostToDevice));
    idx->sq_scale = d->sq_scale;
    idx->sq_offset_host.assign(d->sq_offset, d->sq_offset + d->dim);
  }
  idx->cap_n = n;
  *out = idx.release();
  return HH_OK;
}

void hh_index_free(hh_index *idx) { delete idx; }

int hh_index_update_deleted(hh_index *idx, const uint8_t *deleted, uint64_t live) {
  if (!idx) return fail(HH_INVALID_ARGUMENT, "null index");
  std::unique_lock<std::shared_mutex> wr(idx->mu);
  if (int st = refuse_if_served(idx)) return st;
  idx->generation++;
  if (idx->n == 0) return HH_OK;
  HH_HIP(hipSetDevice(idx->device));
  std::vector<uint32_t> bits;
  pack_deleted(deleted, idx->n, bits);
  int st = drain(idx);
  if (st) return st;
  HH_HIP(hipMemcpy(idx->deleted_bits, bits.data(), bits.size() * 4, hipMemcpyHostToDevice));
  idx->live = live;
  return HH_OK;
}

int hh_index_enable_bf16(hh_index *idx) {
  if (!idx) return fail(HH_INVALID_ARGUMENT, "null index");
  std::unique_lock<std::shared_mutex> wr(idx->mu);
  if (idx->bf16 || idx->n == 0) return HH_OK;
  if (int st = refuse_if_served(idx)) return st;
  idx->generation++;
  HH_HIP(hipSetDevice(idx->device));
  const uint64_t count = idx->n * idx->lay.stride;
  int st = dalloc(&idx->bf16, std::max(idx->cap_n, idx->n) * idx->lay.stride);
  if (st) return st;
  to_bf16_kernel<<<grid_for(count), 256>>>(idx->rows, idx->bf16, count);
  HH_HIP(hipGetLastError());
  HH_HIP(hipDeviceSynchronize());
  return HH_OK;
}

int hh_index_quantize(hh_index *idx, float *offset, float *scale) {
  if (!idx) return fail(HH_INVALID_ARGUMENT, "null index");
  std::unique_lock<std::shared_mutex> wr(idx->mu);
  if (int st = refuse_if_served(idx)) return st;
  idx->generation++;
  HH_HIP(hipSetDevice(idx->device));
  int st = drain(idx);
  if (st) return st;
  const uint32_t dim = idx->dim, stride = idx->lay.stride;
  const uint32_t cstride = (dim + 127) / 128 * 128;
  const uint64_t n = idx->n;
  hipStream_t s = idx->stream;
  // new buffers first: the current quantizer stays whole if one fails
  int8_t *codes = nullptr;
  float *off = nullptr, *off_pos = nullptr;
  uint32_t *mx = nullptr;
  const uint64_t code_rows = std::max<uint64_t>(idx->cap_n, 1);
  st = dalloc(&codes, code_rows * cstride);
  if (!st) st = dalloc(&off, dim);
  if (!st) st = dalloc(&off_pos, stride);
  if (!st) st = dalloc(&mx, 1);
  auto release = [&] {
    dfree(codes);
    dfree(off);
    dfree(off_pos);
    dfree(mx);
  };
  if (st) {
    release();
    return st;
  }
  hhk::SqArgs a{idx->rows, n, stride, dim, idx->lay.T, cstride, off, off_pos, mx, codes};
  const uint64_t code_words = code_rows * cstride / 4;
  hipLaunchKernelGGL(hhk::fill_u32_kernel, dim3(grid_for(code_words)), dim3(256), 0, s,
                     reinterpret_cast<uint32_t *>(codes), code_words, 0u);
  hipLaunchKernelGGL(hhk::fill_u32_kernel, dim3(1), dim3(256), 0, s, mx, (uint64_t)1, 0u);
  if (idx->metric == HH_L2 && n > 0) {
    hipLaunchKernelGGL(hhk::sq_mean_kernel, dim3(stride / 16), dim3(256), 0, s, a);
  } else {
    hipLaunchKernelGGL(hhk::fill_u32_kernel, dim3(grid_for(dim)), dim3(256), 0, s,
                       reinterpret_cast<uint32_t *>(off), (uint64_t)dim, 0u);
    hipLaunchKernelGGL(hhk::fill_u32_kernel, dim3(grid_for(stride)), dim3(256), 0, s,
                       reinterpret_cast<uint32_t *>(off_pos), (uint64_t)stride, 0u);
  }
  if (n > 0) {
    const int grid = (int)std::min<uint64_t>((n + 1023) / 1024, 4096);
    hipLaunchKernelGGL(hhk::sq_max_kernel, dim3(grid), dim3(256), 0, s, a);
    hipLaunchKernelGGL(hhk::sq_encode_kernel, dim3(grid), dim3(256), 0, s, a);
  }
  uint32_t bits = 0;
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(&bits, mx, 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  std::vector<float> off_host(dim);
  if (e == hipSuccess && dim)
    e = hipMemcpy(off_host.data(), off, (size_t)dim * 4, hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    release();
    return fail(HH_DEVICE_ERROR, std::string("device quantizer: ") + hipGetErrorString(e));
  }
  // the scale rule of the training kernels, evaluated the same way here
  float max_abs;
  std::memcpy(&max_abs, &bits, 4);
  const float sc = (max_abs > 0.0f && std::isfinite(max_abs)) ? max_abs / 127.0f : 1.0f;
  dfree(idx->codes);
  dfree(idx->sq_offset);
  idx->codes = codes;
  idx->sq_offset = off;
  idx->cstride = cstride;
  idx->sq_scale = sc;
  idx->sq_offset_host = off_host;
  if (offset && dim) std::memcpy(offset, off_host.data(), (size_t)dim * 4);
  codes = nullptr;
  off = nullptr;
  release();
  if (scale) *scale = sc;
  return HH_OK;
}

int hh_index_sync(hh_index *idx, const hh_graph_desc *d, const uint32_t *dirty, uint64_t ndirty) {
  if (!idx || !d || (ndirty && !dirty)) return fail(HH_INVALID_ARGUMENT, "null argument");
  std::unique_lock<std::shared_mutex> wr(idx->mu);
  if (int st = refuse_if_served(idx)) return st;
  idx->generation++;
  if (d->dim != idx->dim || d->metric != idx->metric || d->m != idx->m || d->m0 != idx->m0)
    return fail(HH_INVALID_ARGUMENT, "graph does not match the resident index: upload it again");
  if (d->n < idx->n) return fail(HH_INVALID_ARGUMENT, "graph has fewer nodes than the resident index");
  if (d->n > 0xFFFFFFFEull) return fail(HH_CAPACITY_EXCEEDED, "too many nodes");
  if (d->n > 0 && (!d->vectors || !d->keys || !d->levels || !d->layer0 || !d->upper_off))
    return fail(HH_INVALID_ARGUMENT, "null graph array");
  if (d->top_level >= 0 && d->entry_id >= d->n)
    return fail(HH_INVALID_ARGUMENT, "entry point out of range");
  // the quantizer must be the resident one (new rows are encoded with it)
  const bool codes = idx->codes != nullptr;
  if (codes != (d->codes != nullptr))
    return fail(HH_INVALID_ARGUMENT, "quantizer changed: upload the index again");
  if (codes && (!d->sq_offset || d->sq_scale != idx->sq_scale ||
                std::memcmp(d->sq_offset, idx->sq_offset_host.data(), (size_t)d->dim * 4) != 0))
    return fail(HH_INVALID_ARGUMENT, "quantizer changed: upload the index again");
  const uint64_t n0 = idx->n, n1 = d->n, add = n1 - n0;
  const uint32_t m0 = d->m0, stride = idx->lay.stride, dim = idx->dim;
  // links of the new nodes and of the changed old ones, as ELL rows
  std::vector<uint32_t> ell_new(add * m0, 0xFFFFFFFFu), ids_old, ell_old;
  bool upper_changed = false;
  uint64_t words = idx->upper_words;
  auto ell_row = [&](uint64_t i, uint32_t *row) -> int {
    const uint32_t *blk = d->layer0 + i * (m0 + 1);
    const uint32_t c = std::min(blk[0], m0);
    for (uint32_t s = 0; s < c; s++) {
      if (blk[1 + s] >= n1) return fail(HH_INVALID_ARGUMENT, "layer-0 neighbour out of range");
      row[s] = blk[1 + s];
    }
    return HH_OK;
  };
  for (uint64_t i = n0; i < n1; i++) {
    int st = ell_row(i, ell_new.data() + (i - n0) * m0);
    if (st) return st;
    const uint64_t end = d->upper_off[i] + (uint64_t)d->levels[i] * (d->m + 1);
    if (d->levels[i]) upper_changed = true;
    words = std::max(words, end);
  }
  for (uint64_t j = 0; j < ndirty; j++) {
    const uint32_t id = dirty[j];
    if (id >= n1) return fail(HH_INVALID_ARGUMENT, "dirty node id out of range");
    if (id >= n0) continue;  // new nodes go over whole
    ids_old.push_back(id);
    ell_old.resize(ids_old.size() * m0, 0xFFFFFFFFu);
    int st = ell_row(id, ell_old.data() + (ids_old.size() - 1) * m0);
    if (st) return st;
    if (d->levels[id]) upper_changed = true;
  }
  HH_HIP(hipSetDevice(idx->device));
  int st = drain(idx);
  if (st) return st;
  // capacity for the new nodes: half as much again, so a run of small
  // inserts moves the arrays rarely
  if (n1 > idx->cap_n) {
    const uint64_t cap = std::max<uint64_t>({n1, idx->cap_n + idx->cap_n / 2, 1024});
    st = grow(&idx->rows, n0 * stride, cap * stride);
    if (!st && idx->bf16) st = grow(&idx->bf16, n0 * stride, cap * stride);
    if (!st && codes) st = grow(&idx->codes, n0 * idx->cstride, cap * idx->cstride);
    if (!st && idx->pd_codes) st = grow(&idx->pd_codes, n0 * idx->cstride, cap * idx->cstride);
    if (!st && idx->pd_rnorm) st = grow(&idx->pd_rnorm, n0, cap);
    if (!st) st = grow(&idx->adj0, n0 * m0, cap * m0);
    if (!st) st = grow(&idx->upper_off, n0, cap);
    if (!st) st = grow(&idx->keys, n0, cap);
    if (!st) st = grow(&idx->deleted_bits, (n0 + 31) / 32, (cap + 31) / 32);
    if (st) return st;
    idx->cap_n = cap;
  }
  if (words > idx->cap_words) {
    const uint64_t cap = std::max<uint64_t>(words, idx->cap_words + idx->cap_words / 2);
    st = grow(&idx->upper, idx->upper_words, cap);
    if (st) return st;
    idx->cap_words = cap;
  }
  if (add) {
    st = upload_rows_into(d->vectors + n0 * dim, add, dim, idx->lay, idx->rows + n0 * stride);
    if (st) return st;
    if (idx->bf16) {
      to_bf16_kernel<<<grid_for(add * stride), 256>>>(idx->rows + n0 * stride,
                                                      idx->bf16 + n0 * stride, add * stride);
      HH_HIP(hipGetLastError());
    }
    if (codes) {
      std::vector<int8_t> packed(add * idx->cstride, 0);
      for (uint64_t i = 0; i < add; i++)
        std::memcpy(packed.data() + i * idx->cstride, d->codes + (n0 + i) * dim, dim);
      HH_HIP(hipMemcpy(idx->codes + n0 * idx->cstride, packed.data(), packed.size(),
                       hipMemcpyHostToDevice));
    }
    if (idx->pd_codes) {
      // new rows take the resident per-dimension quantizer's codes
      hhk::PdArgs pa{idx->rows, n0, n1, stride, dim, idx->lay.T, idx->cstride, idx->metric,
                     nullptr, nullptr, idx->pd_mid, idx->pd_scale, idx->pd_codes, idx->pd_rnorm};
      hipLaunchKernelGGL(hhk::pd_encode_kernel, dim3(grid_for(add * 64)), dim3(256), 0, 0, pa);
      HH_HIP(hipGetLastError());
    }
    HH_HIP(hipMemcpy(idx->adj0 + n0 * m0, ell_new.data(), ell_new.size() * 4,
                     hipMemcpyHostToDevice));
    HH_HIP(hipMemcpy(idx->upper_off + n0, d->upper_off + n0, add * 8, hipMemcpyHostToDevice));
    HH_HIP(hipMemcpy(idx->keys + n0, d->keys + n0, add * 8, hipMemcpyHostToDevice));
  }
  if (!ids_old.empty()) {
    uint32_t *dids = nullptr, *drows = nullptr;
    st = dalloc(&dids, ids_old.size());
    if (!st) st = dalloc(&drows, ell_old.size());
    hipError_t e = hipSuccess;
    if (!st) e = hipMemcpy(dids, ids_old.data(), ids_old.size() * 4, hipMemcpyHostToDevice);
    if (!st && e == hipSuccess)
      e = hipMemcpy(drows, ell_old.data(), ell_old.size() * 4, hipMemcpyHostToDevice);
    if (!st && e == hipSuccess) {
      scatter_rows_kernel<<<grid_for(ell_old.size()), 256>>>(dids, drows, idx->adj0, m0,
                                                             ids_old.size());
      e = hipGetLastError();
    }
    if (!st && e == hipSuccess) e = hipDeviceSynchronize();
    dfree(dids);
    dfree(drows);
    if (st) return st;
    if (e != hipSuccess) return fail(HH_DEVICE_ERROR, std::string("sync: ") + hipGetErrorString(e));
  }
  // upper layers: small (about n/15 blocks of m+1 words); sent whole when
  // any upper block changed
  if (upper_changed && words)
    HH_HIP(hipMemcpy(idx->upper, d->upper, words * 4, hipMemcpyHostToDevice));
  idx->upper_words = words;
  {
    std::vector<uint32_t> bits;
    pack_deleted(d->deleted, n1, bits);
    if (!bits.empty())
      HH_HIP(hipMemcpy(idx->deleted_bits, bits.data(), bits.size() * 4, hipMemcpyHostToDevice));
  }
  HH_HIP(hipDeviceSynchronize());
  idx->n = n1;
  idx->live = d->live;
  idx->entry = d->entry_id;
  idx->top = d->top_level;
  return HH_OK;
}

int hh_index_build_candidates(hh_index *idx, uint64_t first, uint64_t count, uint32_t ef,
                              uint32_t *ids, float *dists, uint32_t *counts) {
  if (!idx || (count && (!ids || !dists || !counts))) return fail(HH_INVALID_ARGUMENT, "null argument");
  if (first > idx->n || count > idx->n - first)
    return fail(HH_INVALID_ARGUMENT, "rows outside the resident index");
  if (ef == 0 || ef > 512) return fail(HH_UNSUPPORTED, "ef must be in 1..512");
  if (count == 0) return HH_OK;
  HH_HIP(hipSetDevice(idx->device));
  uint64_t *dk = nullptr;
  float *dd = nullptr;
  uint32_t *dc = nullptr;
  int32_t *ds = nullptr;
  int st = dalloc(&dk, count * ef);
  if (!st) st = dalloc(&dd, count * ef);
  if (!st) st = dalloc(&dc, count);
  if (!st) st = dalloc(&ds, count);
  auto cleanup = [&] {
    dfree(dk);
    dfree(dd);
    dfree(dc);
    dfree(ds);
  };
  if (st) {
    cleanup();
    return st;
  }
  SearchCall c{idx->rows, count, ef, ef, 0, nullptr, nullptr, ef, nullptr, 0, dk, dd, dc, ds};
  c.build = true;
  c.rows_first = first;
  hipStream_t s = idx->stream;
  st = launch_call(idx, c, s);
  std::vector<uint64_t> hk(count * ef);
  hipError_t e = hipSuccess;
  if (!st) e = hipMemcpyAsync(hk.data(), dk, count * ef * 8, hipMemcpyDeviceToHost, s);
  if (!st && e == hipSuccess) e = hipMemcpyAsync(dists, dd, count * ef * 4, hipMemcpyDeviceToHost, s);
  if (!st && e == hipSuccess) e = hipMemcpyAsync(counts, dc, count * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  cleanup();
  if (st) return st;
  if (e != hipSuccess) return fail(HH_DEVICE_ERROR, std::string("candidates: ") + hipGetErrorString(e));
  for (uint64_t i = 0; i < count * ef; i++) ids[i] = (uint32_t)hk[i];
  return HH_OK;
}

namespace {
// Copies n code rows (cstride bytes apart on the device) into n*dim bytes.
int download_code_rows(hh_index *idx, const int8_t *src, int8_t *codes) {
  // rows are cstride bytes apart on the device: pack chunks of them densely
  // there (one kernel), then copy each chunk in one piece
  const uint64_t chunk = std::min<uint64_t>(idx->n, std::max<uint64_t>(1, (64ull << 20) / idx->dim));
  int8_t *stage = nullptr;
  HH_HIP(hipMalloc((void **)&stage, chunk * idx->dim));
  hipError_t e = hipSuccess;
  for (uint64_t r0 = 0; r0 < idx->n && e == hipSuccess; r0 += chunk) {
    const uint64_t rows = std::min(chunk, idx->n - r0);
    hipLaunchKernelGGL(hhk::pack_codes_kernel, dim3((unsigned)std::min<uint64_t>(rows, 65536)),
                       dim3(64), 0, idx->stream, src + r0 * idx->cstride, idx->cstride,
                       idx->dim, rows, stage);
    e = hipGetLastError();
    if (e == hipSuccess)
      e = hipMemcpyAsync(codes + r0 * idx->dim, stage, rows * idx->dim, hipMemcpyDeviceToHost,
                         idx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(idx->stream);
  }
  dfree(stage);
  if (e != hipSuccess)
    return fail(HH_DEVICE_ERROR, std::string("code download: ") + hipGetErrorString(e));
  return HH_OK;
}
}  // namespace

int hh_index_download_codes(hh_index *idx, int8_t *codes) {
  if (!idx || !codes) return fail(HH_INVALID_ARGUMENT, "null argument");
  std::shared_lock<std::shared_mutex> rd(idx->mu);
  if (!idx->codes) return fail(HH_UNSUPPORTED, "no quantizer codes resident");
  if (idx->n == 0) return HH_OK;
  HH_HIP(hipSetDevice(idx->device));
  if (int st = drain(idx)) return st;
  return download_code_rows(idx, idx->codes, codes);
}

int hh_index_quantize_perdim(hh_index *idx, float *mid, float *scale) {
  if (!idx) return fail(HH_INVALID_ARGUMENT, "null index");
  std::unique_lock<std::shared_mutex> wr(idx->mu);
  if (int st = refuse_if_served(idx)) return st;
  idx->generation++;
  HH_HIP(hipSetDevice(idx->device));
  int st = drain(idx);
  if (st) return st;
  const uint32_t dim = idx->dim, stride = idx->lay.stride;
  const uint32_t cstride = (dim + 127) / 128 * 128;
  const uint64_t n = idx->n, rows_cap = std::max<uint64_t>({idx->cap_n, n, 1});
  hipStream_t s = idx->stream;
  int8_t *codes = nullptr;
  float *rnorm = nullptr, *md = nullptr, *sc = nullptr;
  int32_t *lo = nullptr, *hi = nullptr;
  auto release = [&] {
    dfree(codes);
    dfree(rnorm);
    dfree(md);
    dfree(sc);
    dfree(lo);
    dfree(hi);
  };
  st = dalloc(&codes, rows_cap * cstride);
  if (!st && idx->metric == HH_L2) st = dalloc(&rnorm, rows_cap);
  if (!st) st = dalloc(&md, std::max<uint32_t>(dim, 1));
  if (!st) st = dalloc(&sc, std::max<uint32_t>(dim, 1));
  if (!st) st = dalloc(&lo, stride);
  if (!st) st = dalloc(&hi, stride);
  if (st) {
    release();
    return st;
  }
  hhk::PdArgs a{idx->rows, 0, n, stride, dim, idx->lay.T, cstride, idx->metric, lo, hi, md, sc,
                codes, rnorm};
  std::vector<float> mid_h(dim, 0.0f), scale_h(dim, 1.0f);
  hipError_t e = hipSuccess;
  if (n > 0) {
    hipLaunchKernelGGL(hhk::fill_u32_kernel, dim3(grid_for(stride)), dim3(256), 0, s,
                       reinterpret_cast<uint32_t *>(lo), (uint64_t)stride, 0x7FFFFFFFu);
    hipLaunchKernelGGL(hhk::fill_u32_kernel, dim3(grid_for(stride)), dim3(256), 0, s,
                       reinterpret_cast<uint32_t *>(hi), (uint64_t)stride, 0x80000000u);
    hipLaunchKernelGGL(hhk::pd_minmax_kernel, dim3((unsigned)std::min<uint64_t>((n + 1023) / 1024, 4096)),
                       dim3(256), 0, s, a);
    hipLaunchKernelGGL(hhk::pd_params_kernel, dim3((dim + 255) / 256), dim3(256), 0, s, a);
    hipLaunchKernelGGL(hhk::pd_encode_kernel, dim3(grid_for(n * 64)), dim3(256), 0, s, a);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(mid_h.data(), md, dim * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(scale_h.data(), sc, dim * 4, hipMemcpyDeviceToHost, s);
  } else {
    // no rows: the rule's value for an empty set (mid 0, scale 1)
    e = hipMemcpyAsync(md, mid_h.data(), dim * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(sc, scale_h.data(), dim * 4, hipMemcpyHostToDevice, s);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    release();
    return fail(HH_DEVICE_ERROR, std::string("per-dimension quantizer: ") + hipGetErrorString(e));
  }
  dfree(idx->pd_codes);
  dfree(idx->pd_rnorm);
  dfree(idx->pd_mid);
  dfree(idx->pd_scale);
  idx->pd_codes = codes;
  idx->pd_rnorm = rnorm;
  idx->pd_mid = md;
  idx->pd_scale = sc;
  idx->cstride = cstride;
  codes = nullptr;
  rnorm = md = sc = nullptr;
  release();
  if (mid && dim) std::memcpy(mid, mid_h.data(), (size_t)dim * 4);
  if (scale && dim) std::memcpy(scale, scale_h.data(), (size_t)dim * 4);
  return HH_OK;
}

int hh_index_download_perdim(hh_index *idx, int8_t *codes, float *rnorm) {
  if (!idx) return fail(HH_INVALID_ARGUMENT, "null index");
  std::shared_lock<std::shared_mutex> rd(idx->mu);
  if (!idx->pd_codes) return fail(HH_UNSUPPORTED, "no per-dimension codes resident");
  if (idx->n == 0) return HH_OK;
  HH_HIP(hipSetDevice(idx->device));
  if (int st = drain(idx)) return st;
  if (codes)
    if (int st = download_code_rows(idx, idx->pd_codes, codes)) return st;
  if (rnorm) {
    if (!idx->pd_rnorm) return fail(HH_UNSUPPORTED, "row terms are kept for L2 only");
    HH_HIP(hipMemcpy(rnorm, idx->pd_rnorm, idx->n * 4, hipMemcpyDeviceToHost));
  }
  return HH_OK;
}

int hh_index_download_bf16(hh_index *idx, uint16_t *out) {
  if (!idx || !out) return fail(HH_INVALID_ARGUMENT, "null argument");
  std::shared_lock<std::shared_mutex> rd(idx->mu);
  if (!idx->bf16) return fail(HH_UNSUPPORTED, "bf16 rows not resident");
  HH_HIP(hipSetDevice(idx->device));
  if (int st = drain(idx)) return st;
  const uint32_t dim = idx->dim, stride = idx->lay.stride;
  const uint64_t chunk = std::max<uint64_t>(1, (32ull << 20) / stride);
  std::vector<uint16_t> stage;
  for (uint64_t r0 = 0; r0 < idx->n; r0 += chunk) {
    const uint64_t rows = std::min(chunk, idx->n - r0);
    stage.resize(rows * stride);
    HH_HIP(hipMemcpy(stage.data(), idx->bf16 + r0 * stride, rows * stride * 2,
                     hipMemcpyDeviceToHost));
    for (uint64_t r = 0; r < rows; r++)
      for (uint32_t d = 0; d < dim; d++) out[(r0 + r) * dim + d] = stage[r * stride + idx->lay.pos(d)];
  }
  return HH_OK;
}

int hh_index_stats(hh_index *idx, hh_search_stats *out) {
  if (!idx || !out) return fail(HH_INVALID_ARGUMENT, "null argument");
  HH_HIP(hipSetDevice(idx->device));
  int st = drain(idx);
  if (st) return st;
  std::vector<unsigned long long> lines((size_t)hhk::kStatStride * hhk::kStatLines);
  HH_HIP(hipMemcpy(lines.data(), idx->stats_ws.stats, lines.size() * sizeof(unsigned long long),
                   hipMemcpyDeviceToHost));
  unsigned long long v[kStatCount] = {0};
  for (uint32_t l = 0; l < hhk::kStatLines; l++)
    for (int c = 0; c < kStatCount; c++) v[c] += lines[(size_t)l * hhk::kStatStride + c];
  out->queries = v[0];
  out->n_dist = v[1];
  out->n_upper = v[2];
  out->n_expand = v[3];
  out->n_rerank = v[4];
  out->n_overflow = v[5];
  out->n_visited_spill = v[6];
  out->n_cand_spill = v[7];
  out->n_redo = v[8];
  return HH_OK;
}

int hh_index_set_query_order(hh_index *idx, int longest_first) {
  if (!idx) return fail(HH_INVALID_ARGUMENT, "null index");
  int st = drain(idx);
  if (st) return st;
  idx->longest_first = longest_first != 0;
  return HH_OK;
}

int hh_index_set_beam_width(hh_index *idx, uint32_t width) {
  if (!idx) return fail(HH_INVALID_ARGUMENT, "null index");
  if (width < 1 || width > 8) return fail(HH_INVALID_ARGUMENT, "beam width must be 1..8");
  std::unique_lock<std::shared_mutex> wr(idx->mu);
  if (int st = refuse_if_served(idx)) return st;
  int st = drain(idx);
  if (st) return st;
  idx->beam_width = width;
  idx->generation++;  // captured plans hold the old step rule
  return HH_OK;
}

int hh_index_beam_steps(const hh_index *idx, uint32_t *steps) {
  if (!idx || !steps) return fail(HH_INVALID_ARGUMENT, "null argument");
  *steps = beam_steps(idx);
  return HH_OK;
}

int hh_index_set_timing(hh_index *idx, int on) {
  if (!idx) return fail(HH_INVALID_ARGUMENT, "null index");
  HH_HIP(hipSetDevice(idx->device));
  if (on && !idx->t0) {
    HH_HIP(hipEventCreate(&idx->t0));
    HH_HIP(hipEventCreate(&idx->t1));
  }
  idx->timing = on != 0;
  return HH_OK;
}

int hh_index_last_kernel_ms(hh_index *idx, float *ms) {
  if (!idx || !ms) return fail(HH_INVALID_ARGUMENT, "null argument");
  if (!idx->timing) return fail(HH_INVALID_ARGUMENT, "timing is off (hh_index_set_timing)");
  HH_HIP(hipSetDevice(idx->device));
  HH_HIP(hipEventSynchronize(idx->t1));
  HH_HIP(hipEventElapsedTime(ms, idx->t0, idx->t1));
  return HH_OK;
}

int hh_index_gather_probe(hh_index *idx, uint32_t rounds, double *gbps, float *ms) {
  if (!idx || !gbps || rounds == 0) return fail(HH_INVALID_ARGUMENT, "null argument");
  if (idx->n == 0) return fail(HH_INVALID_ARGUMENT, "empty index");
  if (idx->lay.stride > 256) return fail(HH_UNSUPPORTED, "the probe covers rows of up to 256 floats");
  HH_HIP(hipSetDevice(idx->device));
  if (int st = drain(idx)) return st;
  // 16 waves per CU, the search kernel's residency at ef <= 64
  const uint32_t blocks = 4u * (uint32_t)num_cus(idx->device);
  float *out = nullptr;
  HH_HIP(hipMalloc((void **)&out, (size_t)blocks * 256 * sizeof(float)));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  float best = 0.0f;
  for (int rep = 0; rep < 4 && e == hipSuccess; rep++) {
    e = hipEventRecord(e0, idx->stream);
    if (e == hipSuccess)
      hipLaunchKernelGGL(gather_probe_kernel, dim3(blocks), dim3(256), 0, idx->stream, idx->rows,
                         idx->n, idx->lay.stride, rounds, (uint64_t)rep * 0x9E3779B97F4A7C15ull, out);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(e1, idx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (e == hipSuccess && rep > 0 && (best == 0.0f || t < best)) best = t;  // first run warms up
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(out);
  if (e != hipSuccess) return fail(HH_DEVICE_ERROR, std::string("gather probe: ") + hipGetErrorString(e));
  const double bytes = (double)blocks * 4 * rounds * 16 * idx->dim * 4.0;
  *gbps = bytes / (best * 1e-3) / 1e9;
  if (ms) *ms = best;
  return HH_OK;
}

int hh_index_autotune(hh_index *idx, const float *queries, uint64_t nq, uint32_t k, uint32_t ef,
                      uint32_t precision, uint64_t *out_keys, float *out_dist,
                      uint32_t *out_count, int32_t *status, void *stream,
                      uint32_t *chosen_hash_cap, int *chosen_wpb) {
  if (!idx || !queries || nq == 0) return fail(HH_INVALID_ARGUMENT, "null or empty argument");
  const uint32_t efmax = std::max<uint32_t>(std::max(ef, k), 1);
  const int sw = slots_for(efmax);
  if (!sw) return fail(HH_UNSUPPORTED, "ef (raised to k) above 512 is not supported");
  const int bucket = sw == 1 ? 0 : sw == 2 ? 1 : sw == 4 ? 2 : 3;
  HH_HIP(hipSetDevice(idx->device));
  hipStream_t s = (hipStream_t)stream;
  hipEvent_t e0, e1;
  HH_HIP(hipEventCreate(&e0));
  HH_HIP(hipEventCreate(&e1));
  float best_ms = 1e30f;
  uint32_t best_cap = 0;
  int best_wpb = 0;
  const uint32_t caps[] = {1024, 1536, 1792, 2048, 2560, 3072, 4096, 6144, 8192, 12288, 16384};
  int st = HH_OK;
  // best of three timed launches after a warm-up, in the current shape and order
  auto time_launch = [&](float &ms_best) {
    ms_best = 1e30f;
    for (int rep = 0; rep < 4 && !st; rep++) {
      if (hipEventRecord(e0, s) != hipSuccess) st = HH_DEVICE_ERROR;
      if (!st)
        st = launch_search(idx, queries, nq, k, ef, precision, nullptr, nullptr, k, nullptr, 0,
                           out_keys, out_dist, out_count, status, s);
      if (!st && hipEventRecord(e1, s) != hipSuccess) st = HH_DEVICE_ERROR;
      float ms = 0;
      if (!st && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        st = HH_DEVICE_ERROR;
      if (rep > 0) ms_best = std::min(ms_best, ms);  // first run warms up
    }
  };
  const int saved_order = idx->tuned_order[bucket];
  idx->tuned_order[bucket] = 0;
  for (uint32_t cap : caps) {
    if (cap < 256u * (uint32_t)sw || cap * 4 < efmax * 16) continue;
    for (int w : {1, 2, 4}) {
      if ((uint64_t)lds_words_for(cap, sw) * 4 * (uint32_t)w > 65536) continue;
      idx->tuned_hash_cap[bucket] = cap;
      idx->tuned_wpb[bucket] = w;
      float ms_best;
      time_launch(ms_best);
      if (st) break;
      if (ms_best < best_ms) {
        best_ms = ms_best;
        best_cap = cap;
        best_wpb = w;
      }
    }
    if (st) break;
  }
  // then one pass against two in the chosen shape (when the order is on)
  int best_order = 0;
  if (!st && idx->longest_first) {
    idx->tuned_hash_cap[bucket] = best_cap;
    idx->tuned_wpb[bucket] = best_wpb;
    float ms1 = 0, ms2 = 0;
    idx->tuned_order[bucket] = 1;
    time_launch(ms1);
    idx->tuned_order[bucket] = 2;
    if (!st) time_launch(ms2);
    best_order = ms2 < ms1 ? 2 : 1;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  idx->tuned_hash_cap[bucket] = st ? 0 : best_cap;
  idx->tuned_wpb[bucket] = st ? 0 : best_wpb;
  idx->tuned_order[bucket] = st ? saved_order : best_order;
  if (st) return st == HH_DEVICE_ERROR ? fail(st, "autotune launch failed") : st;
  if (chosen_hash_cap) *chosen_hash_cap = best_cap;
  if (chosen_wpb) *chosen_wpb = best_wpb;
  return HH_OK;
}

int hh_index_set_launch_shape(hh_index *idx, uint32_t ef, uint32_t hash_cap, int wpb) {
  if (!idx) return fail(HH_INVALID_ARGUMENT, "null index");
  const int sw = slots_for(std::max<uint32_t>(ef, 1));
  if (!sw) return fail(HH_UNSUPPORTED, "ef above 512 is not supported");
  const int bucket = sw == 1 ? 0 : sw == 2 ? 1 : sw == 4 ? 2 : 3;
  if (hash_cap == 0) {
    idx->tuned_hash_cap[bucket] = 0;
    idx->tuned_wpb[bucket] = 0;
    idx->tuned_order[bucket] = 0;
    return HH_OK;
  }
  if (wpb != 1 && wpb != 2 && wpb != 4)
    return fail(HH_INVALID_ARGUMENT, "waves per block must be 1, 2 or 4");
  if (hash_cap < 256u * (uint32_t)sw || (uint64_t)lds_words_for(hash_cap, sw) * 4 * (uint64_t)wpb > 65536)
    return fail(HH_INVALID_ARGUMENT, "visited-set capacity out of range for this ef and block");
  idx->tuned_hash_cap[bucket] = hash_cap;
  idx->tuned_wpb[bucket] = wpb;
  return HH_OK;
}

int hh_index_reset_stats(hh_index *idx) {
  if (!idx) return fail(HH_INVALID_ARGUMENT, "null index");
  HH_HIP(hipSetDevice(idx->device));
  int st = drain(idx);
  if (st) return st;
  HH_HIP(hipMemset(idx->stats_ws.stats, 0,
                   (size_t)hhk::kStatStride * hhk::kStatLines * sizeof(unsigned long long)));
  return HH_OK;
}

int hh_search_batch_device(hh_index *idx, const float *queries, uint64_t nq, uint32_t k,
                           uint32_t ef, uint32_t precision, const uint32_t *allow_bits,
                           uint64_t *out_keys, float *out_dist, uint32_t *out_count,
                           int32_t *status, void *stream) {
  if (!idx) return fail(HH_INVALID_ARGUMENT, "null index");
  return launch_search(idx, queries, nq, k, ef, precision, allow_bits, nullptr, k, nullptr, 0,
                       out_keys, out_dist, out_count, status, (hipStream_t)stream);
}

int hh_search_batch_spec_device(hh_index *idx, const float *queries, uint64_t nq,
                                const hh_query_spec *specs, uint32_t k_stride, uint32_t ef_max,
                                const uint32_t *filters, uint32_t nfilters, uint64_t *out_keys,
                                float *out_dist, uint32_t *out_count, int32_t *status,
                                void *stream) {
  if (!idx) return fail(HH_INVALID_ARGUMENT, "null index");
  if (!specs) return fail(HH_INVALID_ARGUMENT, "null specs");
  return launch_search(idx, queries, nq, 0, ef_max, 0, nullptr, specs, k_stride, filters,
                       nfilters, out_keys, out_dist, out_count, status, (hipStream_t)stream);
}

int hh_search_batch(hh_index *idx, const float *queries, uint64_t nq, uint32_t k, uint32_t ef,
                    uint32_t precision, const uint32_t *allow_bits, uint64_t *out_keys,
                    float *out_dist, uint32_t *out_count, uint64_t *first_bad) {
  if (!idx) return fail(HH_INVALID_ARGUMENT, "null index");
  if (nq == 0) return HH_OK;
  if (!queries || !out_count || (k && (!out_keys || !out_dist)))
    return fail(HH_INVALID_ARGUMENT, "null argument");
  HH_HIP(hipSetDevice(idx->device));
  float *dq = nullptr, *dd = nullptr;
  uint64_t *dk = nullptr;
  uint32_t *dc = nullptr, *da = nullptr;
  int32_t *ds = nullptr;
  auto cleanup = [&]() {
    dfree(dq);
    dfree(dd);
    dfree(dk);
    dfree(dc);
    dfree(da);
    dfree(ds);
  };
  const uint64_t kk = std::max<uint32_t>(k, 1);
  int st = dalloc(&dq, nq * idx->dim);
  if (!st) st = dalloc(&dd, nq * kk);
  if (!st) st = dalloc(&dk, nq * kk);
  if (!st) st = dalloc(&dc, nq);
  if (!st) st = dalloc(&ds, nq);
  if (!st && allow_bits) st = dalloc(&da, std::max<uint64_t>(1, (idx->n + 31) / 32));
  if (st) {
    cleanup();
    return st;
  }
  hipStream_t s = idx->stream;
  hipError_t e = hipMemcpyAsync(dq, queries, nq * idx->dim * 4, hipMemcpyHostToDevice, s);
  if (e == hipSuccess && allow_bits)
    e = hipMemcpyAsync(da, allow_bits, ((idx->n + 31) / 32) * 4, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) {
    cleanup();
    return fail(HH_DEVICE_ERROR, hipGetErrorString(e));
  }
  st = launch_search(idx, dq, nq, k, ef, precision, da, nullptr, k, nullptr, 0, dk, dd, dc, ds, s);
  if (st) {
    (void)hipStreamSynchronize(s);
    cleanup();
    return st;
  }
  std::vector<int32_t> status(nq);
  e = hipMemcpyAsync(status.data(), ds, nq * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(out_count, dc, nq * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess && k) e = hipMemcpyAsync(out_keys, dk, nq * k * 8, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess && k) e = hipMemcpyAsync(out_dist, dd, nq * k * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  cleanup();
  if (e != hipSuccess) return fail(HH_DEVICE_ERROR, hipGetErrorString(e));
  return first_failure(status, first_bad);
}

int hh_search_batch_spec(hh_index *idx, const float *queries, uint64_t nq,
                         const hh_query_spec *specs, uint32_t k_stride, const uint32_t *filters,
                         uint32_t nfilters, uint64_t *out_keys, float *out_dist,
                         uint32_t *out_count, int32_t *status) {
  if (!idx) return fail(HH_INVALID_ARGUMENT, "null index");
  if (nq == 0) return HH_OK;
  if (!queries || !specs || !out_count || !status || (k_stride && (!out_keys || !out_dist)))
    return fail(HH_INVALID_ARGUMENT, "null argument");
  HH_HIP(hipSetDevice(idx->device));
  float *dq = nullptr, *dd = nullptr;
  uint64_t *dk = nullptr;
  uint32_t *dc = nullptr, *df = nullptr;
  int32_t *ds = nullptr;
  hh_query_spec *dsp = nullptr;
  auto cleanup = [&]() {
    dfree(dq);
    dfree(dd);
    dfree(dk);
    dfree(dc);
    dfree(df);
    dfree(ds);
    dfree(dsp);
  };
  const uint64_t kk = std::max<uint32_t>(k_stride, 1);
  const uint64_t words = (idx->n + 31) / 32;
  int st = dalloc(&dq, nq * idx->dim);
  if (!st) st = dalloc(&dd, nq * kk);
  if (!st) st = dalloc(&dk, nq * kk);
  if (!st) st = dalloc(&dc, nq);
  if (!st) st = dalloc(&ds, nq);
  if (!st) st = dalloc(&dsp, nq);
  if (!st && filters && nfilters) st = dalloc(&df, std::max<uint64_t>(1, words * nfilters));
  if (st) {
    cleanup();
    return st;
  }
  hipStream_t s = idx->stream;
  hipError_t e = hipMemcpyAsync(dq, queries, nq * idx->dim * 4, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(dsp, specs, nq * sizeof(hh_query_spec), hipMemcpyHostToDevice, s);
  if (e == hipSuccess && df && words)
    e = hipMemcpyAsync(df, filters, words * nfilters * 4, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) {
    cleanup();
    return fail(HH_DEVICE_ERROR, hipGetErrorString(e));
  }
  uint32_t ef_max = 1;
  for (uint64_t i = 0; i < nq; i++)
    ef_max = std::max(ef_max, std::max(specs[i].ef, specs[i].k));
  ef_max = std::min<uint32_t>(ef_max, 512);  // larger ones fail per query
  st = launch_search(idx, dq, nq, 0, ef_max, 0, nullptr, dsp, k_stride, df, df ? nfilters : 0, dk, dd,
                     dc, ds, s);
  if (st) {
    (void)hipStreamSynchronize(s);
    cleanup();
    return st;
  }
  e = hipMemcpyAsync(status, ds, nq * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(out_count, dc, nq * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess && k_stride)
    e = hipMemcpyAsync(out_keys, dk, nq * k_stride * 8, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess && k_stride)
    e = hipMemcpyAsync(out_dist, dd, nq * k_stride * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  cleanup();
  if (e != hipSuccess) return fail(HH_DEVICE_ERROR, hipGetErrorString(e));
  return HH_OK;
}

}  // extern "C"

struct hh_search_plan {
  hh_index *idx = nullptr;
  uint64_t generation = 0;  // the index's generation the graph was captured at
  Workspace ws;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t cap = nullptr;
  ~hh_search_plan() {
    if (idx) (void)hipSetDevice(idx->device);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (cap) (void)hipStreamDestroy(cap);
    ws.release();
  }
};

extern "C" {

int hh_search_plan_create(hh_index *idx, const float *queries, uint64_t nq, uint32_t k,
                          uint32_t ef, uint32_t precision, const uint32_t *allow_bits,
                          const hh_query_spec *specs, uint32_t k_stride, const uint32_t *filters,
                          uint32_t nfilters, uint64_t *out_keys, float *out_dist,
                          uint32_t *out_count, int32_t *status, hh_search_plan **out) {
  if (!idx || !out) return fail(HH_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (nq == 0) return fail(HH_INVALID_ARGUMENT, "empty batch");
  const SearchCall c{queries, nq, specs ? 0u : k, ef, specs ? 0u : precision,
                     specs ? nullptr : allow_bits, specs, specs ? k_stride : k,
                     specs ? filters : nullptr, specs ? nfilters : 0u, out_keys, out_dist,
                     out_count, status};
  std::shared_lock<std::shared_mutex> rd(idx->mu);
  int st = check_call(idx, c);
  if (st) return st;
  HH_HIP(hipSetDevice(idx->device));
  std::unique_ptr<hh_search_plan> p(new hh_search_plan());
  p->idx = idx;
  p->generation = idx->generation.load();
  SearchConfig cfg;
  st = search_config(idx, c, cfg);
  if (!st) st = p->ws.init();
  if (!st) st = p->ws.ensure_queries(nq, idx->lay.stride, idx->cstride);
  if (!st) st = p->ws.ensure_spill(cfg.blocks * cfg.wpb, cfg.spill_words);
  if (!st && p->ws.pool_cap < idx->n) st = p->ws.ensure_bigpool(idx->n);
  if (st) return st;
  HH_HIP(hipStreamCreateWithFlags(&p->cap, hipStreamNonBlocking));
  HH_HIP(hipStreamBeginCapture(p->cap, hipStreamCaptureModeThreadLocal));
  st = enqueue_search(idx, p->ws, c, cfg, p->cap);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(p->cap, &g);
  p->graph = g;
  if (st) return st;
  if (e != hipSuccess) return fail(HH_DEVICE_ERROR, std::string("graph capture: ") + hipGetErrorString(e));
  HH_HIP(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
  *out = p.release();
  return HH_OK;
}

int hh_search_plan_launch(hh_search_plan *p, void *stream) {
  if (!p) return fail(HH_INVALID_ARGUMENT, "null plan");
  std::shared_lock<std::shared_mutex> rd(p->idx->mu);
  if (p->generation != p->idx->generation.load())
    return fail(HH_INVALID_ARGUMENT,
                "the index changed since the plan was created: create the plan again");
  HH_HIP(hipSetDevice(p->idx->device));
  hipStream_t s = (hipStream_t)stream;
  HH_HIP(hipStreamWaitEvent(s, p->ws.done, 0));
  HH_HIP(hipGraphLaunch(p->exec, s));
  HH_HIP(hipEventRecord(p->ws.done, s));
  return HH_OK;
}

void hh_search_plan_free(hh_search_plan *p) { delete p; }

// ---------------------------------------------------------------- brute
int hh_brute_upload(uint32_t dim, uint32_t metric, const uint64_t *keys, const float *vectors,
                    uint64_t n, int device, hh_brute **out, uint64_t *bad_key) {
  if (!out) return fail(HH_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (dim == 0 || dim > 65536)
    return fail(HH_INVALID_DIMENSION, "invalid dimension " + std::to_string(dim) +
                                          ": must be between 1 and 65536");
  if (metric > 2) return fail(HH_INVALID_ARGUMENT, "unknown metric");
  if (n && (!keys || !vectors)) return fail(HH_INVALID_ARGUMENT, "null argument");
  // prepare_batch then the duplicate check (src/brute.rs:102-114)
  std::vector<float> prepared((size_t)n * dim);
  for (uint64_t i = 0; i < n; i++) {
    int st = hh::prepare_vector(metric, dim, vectors + i * dim, prepared.data() + i * dim);
    if (st) return st;
  }
  std::unordered_set<uint64_t> seen;
  seen.reserve(n);
  for (uint64_t i = 0; i < n; i++)
    if (!seen.insert(keys[i]).second) {
      if (bad_key) *bad_key = keys[i];
      return fail(HH_DUPLICATE_KEY, "key " + std::to_string(keys[i]) + " is already present");
    }
  HH_HIP(hipSetDevice(device));
  std::unique_ptr<hh_brute> b(new hh_brute());
  b->device = device;
  b->dim = dim;
  b->metric = metric;
  b->n = n;
  b->lay = hh::RowLayout(dim);
  int st = b->ws.init();
  if (st) return st;
  HH_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  st = upload_rows(prepared.data(), n, dim, b->lay, &b->rows);
  if (!st) st = dalloc(&b->keys, n);
  if (!st) st = dalloc(&b->xx, n);
  if (st) return st;
  if (n) {
    HH_HIP(hipMemcpy(b->keys, keys, n * 8, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(hhk::row_norms_kernel, dim3(grid_for(n)), dim3(256), 0, nullptr, b->rows, n,
                       b->lay.stride, b->xx);
    HH_HIP(hipGetLastError());
    HH_HIP(hipDeviceSynchronize());
    double mx = 0.0;
    for (uint64_t i = 0; i < n; i++) {
      double ss = 0.0;
      for (uint32_t e = 0; e < dim; e++) {
        const double v = prepared[i * dim + e];
        ss += v * v;
      }
      mx = std::max(mx, ss);
    }
    b->max_norm = std::sqrt(mx);
  }
  *out = b.release();
  return HH_OK;
}

void hh_brute_free(hh_brute *b) { delete b; }

}  // extern "C"

namespace {

template <int TM, bool EX, int QB>
void *brute_partial_ptr(int s) {
  if (s == 1) return reinterpret_cast<void *>(&hhk::brute_partial_kernel<TM, EX, QB, 1>);
  if (s == 2) return reinterpret_cast<void *>(&hhk::brute_partial_kernel<TM, EX, QB, 2>);
  if (s == 4) return reinterpret_cast<void *>(&hhk::brute_partial_kernel<TM, EX, QB, 4>);
  return reinterpret_cast<void *>(&hhk::brute_partial_kernel<TM, EX, QB, 8>);
}

void *mfma_ptr(uint32_t ks) {
  switch (ks) {
    case 8: return reinterpret_cast<void *>(&hhk::brute_mfma_kernel<8>);
    case 16: return reinterpret_cast<void *>(&hhk::brute_mfma_kernel<16>);
    case 24: return reinterpret_cast<void *>(&hhk::brute_mfma_kernel<24>);
    case 32: return reinterpret_cast<void *>(&hhk::brute_mfma_kernel<32>);
    case 40: return reinterpret_cast<void *>(&hhk::brute_mfma_kernel<40>);
    case 48: return reinterpret_cast<void *>(&hhk::brute_mfma_kernel<48>);
    case 56: return reinterpret_cast<void *>(&hhk::brute_mfma_kernel<56>);
    case 64: return reinterpret_cast<void *>(&hhk::brute_mfma_kernel<64>);
    default: return nullptr;
  }
}

// Whether a batch takes the MFMA shortlist path: f32, unfiltered, D <= 128,
// k <= 64, and (in automatic mode) enough rows for the GEMM to pay.
bool mfma_eligible(const hh_brute *b, uint64_t nq, uint32_t kc, const uint32_t *allow, bool f64) {
  if (b->mfma_mode == 0 || f64 || allow || kc == 0 || kc > 64 || b->n == 0) return false;
  if (b->lay.stride > 128 || !mfma_ptr(b->lay.stride / 2)) return false;
  if (b->mfma_mode == 2) return b->n >= 64;
  return b->n >= 32768 && nq >= 32;
}

// MFMA shortlist for the queries [c0, c0 + cn) of a batch (queries already
// prepared into ws.qprep).  Writes results for queries proven exact and
// sets ws.m_redo for the rest.
int mfma_chunk(hh_brute *b, Workspace &ws, const hhk::BruteArgs &ba, uint64_t c0, uint64_t cn,
               uint32_t kc, uint64_t row_step, uint32_t cap, hipStream_t stream) {
  const uint32_t stride = b->lay.stride;
  hhk::MfmaArgs m{};
  m.rows = b->rows;
  m.n = b->n;
  m.stride = stride;
  m.dim = b->dim;
  m.metric = b->metric;
  m.qprep = ws.qprep + c0 * stride;
  m.status = ba.status + c0;
  m.nq = cn;
  m.xx = b->xx;
  m.qq = ws.m_qq;
  m.dense = ws.m_dense;
  m.tau = ws.m_tau;
  m.cand = ws.m_cand;
  m.cand_count = ws.m_cnt;
  m.cand_cap = cap;
  m.max_norm = b->max_norm;
  m.redo = ws.m_redo + c0;
  void *mk = mfma_ptr(stride / 2);
  const unsigned gx = (unsigned)((cn + 63) / 64);
  hipLaunchKernelGGL(hhk::query_norms_kernel, dim3(grid_for(cn)), dim3(256), 0, stream, m.qprep,
                     cn, stride, ws.m_qq);
  HH_HIP(hipGetLastError());
  // 1. sample pass (dense) and per-query thresholds
  m.mode = 0;
  m.row_step = row_step;
  m.nrows = (b->n + row_step - 1) / row_step;
  {
    const uint64_t tiles = (m.nrows + 31) / 32;
    const uint64_t gy = std::max<uint64_t>(
        1, std::min<uint64_t>((tiles + 3) / 4, (2048 + gx - 1) / gx));
    void *args[] = {&m};
    HH_HIP(hipLaunchKernel(mk, dim3(gx, (unsigned)gy), dim3(256), args, 0, stream));
  }
  hipLaunchKernelGGL(hhk::brute_tau_kernel, dim3((unsigned)((cn + 3) / 4)), dim3(256), 0, stream,
                     m);
  HH_HIP(hipGetLastError());
  hipLaunchKernelGGL(hhk::fill_u32_kernel, dim3(grid_for(cn)), dim3(256), 0, stream, ws.m_cnt, cn,
                     0u);
  HH_HIP(hipGetLastError());
  // 2. candidate filter over every row
  m.mode = 1;
  m.row_step = 1;
  m.nrows = b->n;
  {
    const uint64_t tiles = (b->n + 31) / 32;
    const uint64_t gy = std::max<uint64_t>(
        1, std::min<uint64_t>((tiles + 3) / 4, (2048 + gx - 1) / gx));
    void *args[] = {&m};
    HH_HIP(hipLaunchKernel(mk, dim3(gx, (unsigned)gy), dim3(256), args, 0, stream));
  }
  // 3. exact re-rank of the candidates
  hhk::BruteArgs rb = ba;
  rb.qprep = ws.qprep + c0 * stride;
  rb.out_keys = ba.out_keys + c0 * ba.k_stride;
  rb.out_dist = ba.out_dist + c0 * ba.k_stride;
  rb.out_count = ba.out_count + c0;
  const bool ex = b->lay.U > 0 || b->lay.R > 0;
  const int s = slots_for(kc);
  void *rk;
  if (ex)
    rk = s == 1 ? (void *)&hhk::brute_rerank_kernel<4, true, 1>
                : (void *)&hhk::brute_rerank_kernel<4, true, 2>;
  else
    rk = s == 1 ? (void *)&hhk::brute_rerank_kernel<4, false, 1>
                : (void *)&hhk::brute_rerank_kernel<4, false, 2>;
  void *args[] = {&m, &rb};
  HH_HIP(hipLaunchKernel(rk, dim3((unsigned)((cn + 3) / 4)), dim3(256), args, 0, stream));
  return HH_OK;
}

int brute_launch(hh_brute *b, const float *queries, uint64_t nq, uint32_t k,
                 const uint32_t *allow, uint64_t *out_keys, float *out_dist, double *out_dist64,
                 uint32_t *out_count, int32_t *status, hipStream_t stream, bool f64) {
  if (!b) return fail(HH_INVALID_ARGUMENT, "null index");
  if (nq == 0) return HH_OK;
  if (!queries || !out_count) return fail(HH_INVALID_ARGUMENT, "null device pointer");
  const int tmax = tmax_for(b->lay);
  if (!tmax) return fail(HH_UNSUPPORTED, "brute-force kernels support dim <= 1024");
  const uint32_t kc = (uint32_t)std::min<uint64_t>(k, b->n);  // src/brute.rs:141-142
  if (kc > 512) return fail(HH_UNSUPPORTED, "k above 512 (after clamping to len) is not supported");
  if (kc && (!out_keys || (!out_dist && !out_dist64)))
    return fail(HH_INVALID_ARGUMENT, "null device pointer");
  const int s = slots_for(std::max<uint32_t>(kc, 1));
  HH_HIP(hipSetDevice(b->device));
  Workspace &ws = b->ws;
  std::lock_guard<std::mutex> lock(ws.mu);
  b->last_mfma = 0;
  int st = ws.ensure_queries(nq, b->lay.stride, 0);
  if (st) return st;
  const int qb = f64 ? 1 : (tmax == 4 ? 4 : (tmax == 8 ? 2 : 1));
  const uint64_t groups = (nq + qb - 1) / qb;
  const uint64_t target_waves = 8192;
  uint64_t nslices = std::max<uint64_t>(1, (target_waves + groups - 1) / groups);
  nslices = std::min<uint64_t>(nslices, std::max<uint64_t>(1, (b->n + 1) / 2));
  nslices = std::min<uint64_t>(nslices, 1024);
  uint64_t rps = (b->n + nslices - 1) / nslices;
  rps = (rps + 3) / 4 * 4;
  if (rps == 0) rps = 4;
  nslices = std::max<uint64_t>(1, (b->n + rps - 1) / rps);
  st = ws.ensure_part(nq * nslices * std::max<uint32_t>(kc, 1), nq * nslices, f64);
  if (st) return st;
  HH_HIP(hipStreamWaitEvent(stream, ws.done, 0));
  hhk::PrepArgs pa{};
  pa.queries = queries;
  pa.nq = nq;
  pa.dim = b->dim;
  pa.metric = b->metric;
  pa.T = (int)b->lay.T;
  pa.stride = b->lay.stride;
  pa.qprep = ws.qprep;
  pa.sq_scale = 1.0f;
  pa.status = status ? status : ws.status;
  pa.counter = nullptr;
  pa.ef_cap = 512;
  hipLaunchKernelGGL(hhk::prepare_queries_kernel, dim3(prep_grid(nq)), dim3(256), 0, stream, pa);
  HH_HIP(hipGetLastError());
  hhk::BruteArgs a{};
  a.rows = b->rows;
  a.keys = b->keys;
  a.allow = allow;
  a.n = b->n;
  a.dim = b->dim;
  a.metric = b->metric;
  a.T = (int)b->lay.T;
  a.U = (int)b->lay.U;
  a.R = (int)b->lay.R;
  a.stride = b->lay.stride;
  a.qprep = ws.qprep;
  a.status = pa.status;
  a.nq = nq;
  a.k = kc;
  a.k_stride = k;
  a.nslices = (uint32_t)nslices;
  a.rows_per_slice = rps;
  a.part_d = ws.part_d;
  a.part_d64 = ws.part_d64;
  a.part_k = ws.part_k;
  a.part_n = ws.part_n;
  a.out_keys = out_keys;
  a.out_dist = out_dist;
  a.out_dist64 = out_dist64;
  a.out_count = out_count;
  void *args[] = {&a};
  if (kc == 0 || b->n == 0) {
    // nothing to rank: every valid query gets zero results
    a.nslices = 0;
    void *mk = reinterpret_cast<void *>(&hhk::brute_merge_kernel<1>);
    HH_HIP(hipLaunchKernel(mk, dim3((unsigned)((nq + 3) / 4)), dim3(256), args, 0, stream));
    HH_HIP(hipEventRecord(ws.done, stream));
    return HH_OK;
  }
  if (mfma_eligible(b, nq, kc, allow, f64)) {
    // sample stride: about 8 * row_step candidates per query, >= 4k
    const uint64_t row_step = std::min<uint64_t>(256, std::max<uint64_t>(16, 4 * (uint64_t)kc));
    const uint32_t cap = (uint32_t)(32 * row_step);
    const uint64_t nsample = (b->n + row_step - 1) / row_step;
    // query chunk: dense sample block <= 1 GiB
    uint64_t chunk = std::max<uint64_t>(64, ((1ull << 28) / std::max<uint64_t>(1, nsample)) / 64 * 64);
    chunk = std::min<uint64_t>(chunk, nq);
    st = ws.ensure_mfma(nq, chunk * nsample, chunk * (uint64_t)cap);
    if (st) return st;
    for (uint64_t c0 = 0; c0 < nq; c0 += chunk) {
      st = mfma_chunk(b, ws, a, c0, std::min<uint64_t>(chunk, nq - c0), kc, row_step, cap, stream);
      if (st) return st;
    }
    // exact streaming pass for the flagged queries only
    a.only = ws.m_redo;
    b->last_mfma = nq;
  }
  const uint64_t work = f64 ? nq * nslices : groups * nslices;
  void *pk;
  void *mk;
  if (f64) {
    pk = s == 1 ? (void *)&hhk::brute_f64_kernel<1>
                : s == 2 ? (void *)&hhk::brute_f64_kernel<2>
                         : s == 4 ? (void *)&hhk::brute_f64_kernel<4> : (void *)&hhk::brute_f64_kernel<8>;
    mk = s == 1 ? (void *)&hhk::brute_merge_f64_kernel<1>
                : s == 2 ? (void *)&hhk::brute_merge_f64_kernel<2>
                         : s == 4 ? (void *)&hhk::brute_merge_f64_kernel<4>
                                  : (void *)&hhk::brute_merge_f64_kernel<8>;
  } else {
    const bool ex = b->lay.U > 0 || b->lay.R > 0;
    if (tmax == 4) pk = ex ? brute_partial_ptr<4, true, 4>(s) : brute_partial_ptr<4, false, 4>(s);
    else if (tmax == 8) pk = ex ? brute_partial_ptr<8, true, 2>(s) : brute_partial_ptr<8, false, 2>(s);
    else pk = ex ? brute_partial_ptr<32, true, 1>(s) : brute_partial_ptr<32, false, 1>(s);
    mk = s == 1 ? (void *)&hhk::brute_merge_kernel<1>
                : s == 2 ? (void *)&hhk::brute_merge_kernel<2>
                         : s == 4 ? (void *)&hhk::brute_merge_kernel<4> : (void *)&hhk::brute_merge_kernel<8>;
  }
  HH_HIP(hipLaunchKernel(pk, dim3((unsigned)((work + 3) / 4)), dim3(256), args, 0, stream));
  HH_HIP(hipLaunchKernel(mk, dim3((unsigned)((nq + 3) / 4)), dim3(256), args, 0, stream));
  HH_HIP(hipEventRecord(ws.done, stream));
  return HH_OK;
}

}  // namespace

extern "C" {

int hh_brute_search_batch_device(hh_brute *b, const float *queries, uint64_t nq, uint32_t k,
                                 const uint32_t *allow_bits, uint64_t *out_keys, float *out_dist,
                                 uint32_t *out_count, int32_t *status, void *stream) {
  return brute_launch(b, queries, nq, k, allow_bits, out_keys, out_dist, nullptr, out_count,
                      status, (hipStream_t)stream, false);
}

int hh_brute_search_batch_f64_device(hh_brute *b, const float *queries, uint64_t nq, uint32_t k,
                                     uint64_t *out_keys, double *out_dist, uint32_t *out_count,
                                     int32_t *status, void *stream) {
  return brute_launch(b, queries, nq, k, nullptr, out_keys, nullptr, out_dist, out_count, status,
                      (hipStream_t)stream, true);
}

}  // extern "C"

extern "C" {

uint64_t hh_brute_len(const hh_brute *b) { return b ? b->n : 0; }

int hh_brute_set_mfma(hh_brute *b, int mode) {
  if (!b) return fail(HH_INVALID_ARGUMENT, "null index");
  if (mode < 0 || mode > 2) return fail(HH_INVALID_ARGUMENT, "mode must be 0, 1 or 2");
  std::lock_guard<std::mutex> lock(b->ws.mu);
  b->mfma_mode = mode;
  return HH_OK;
}

int hh_brute_mfma_stats(hh_brute *b, uint64_t *mfma_queries, uint64_t *redo_queries) {
  if (!b) return fail(HH_INVALID_ARGUMENT, "null index");
  HH_HIP(hipSetDevice(b->device));
  std::lock_guard<std::mutex> lock(b->ws.mu);
  HH_HIP(hipEventSynchronize(b->ws.done));
  uint64_t redo = 0;
  if (b->last_mfma) {
    std::vector<uint32_t> f(b->last_mfma);
    HH_HIP(hipMemcpy(f.data(), b->ws.m_redo, f.size() * 4, hipMemcpyDeviceToHost));
    for (uint32_t v : f) redo += v != 0;
  }
  if (mfma_queries) *mfma_queries = b->last_mfma;
  if (redo_queries) *redo_queries = redo;
  return HH_OK;
}

static int brute_host(hh_brute *b, const float *queries, uint64_t nq, uint32_t k,
                      const uint32_t *allow_bits, uint64_t *out_keys, float *out_dist,
                      double *out_dist64, uint32_t *out_count, uint64_t *first_bad, bool f64) {
  if (!b) return fail(HH_INVALID_ARGUMENT, "null index");
  if (nq == 0) return HH_OK;
  if (!queries || !out_count || (k && (!out_keys || (!out_dist && !out_dist64))))
    return fail(HH_INVALID_ARGUMENT, "null argument");
  HH_HIP(hipSetDevice(b->device));
  const uint64_t kk = std::max<uint32_t>(k, 1);
  float *dq = nullptr, *dd = nullptr;
  double *dd64 = nullptr;
  uint64_t *dk = nullptr;
  uint32_t *dc = nullptr, *da = nullptr;
  int32_t *ds = nullptr;
  auto cleanup = [&]() {
    dfree(dq);
    dfree(dd);
    dfree(dd64);
    dfree(dk);
    dfree(dc);
    dfree(da);
    dfree(ds);
  };
  const uint64_t words = std::max<uint64_t>(1, (b->n + 31) / 32);
  int st = dalloc(&dq, nq * b->dim);
  if (!st) st = f64 ? dalloc(&dd64, nq * kk) : dalloc(&dd, nq * kk);
  if (!st) st = dalloc(&dk, nq * kk);
  if (!st) st = dalloc(&dc, nq);
  if (!st) st = dalloc(&ds, nq);
  if (!st && allow_bits) st = dalloc(&da, words);
  if (st) {
    cleanup();
    return st;
  }
  hipStream_t s = b->stream;
  hipError_t e = hipMemcpyAsync(dq, queries, nq * b->dim * 4, hipMemcpyHostToDevice, s);
  if (e == hipSuccess && allow_bits)
    e = hipMemcpyAsync(da, allow_bits, words * 4, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) {
    cleanup();
    return fail(HH_DEVICE_ERROR, hipGetErrorString(e));
  }
  st = brute_launch(b, dq, nq, k, da, dk, dd, dd64, dc, ds, s, f64);
  if (st) {
    (void)hipStreamSynchronize(s);
    cleanup();
    return st;
  }
  std::vector<int32_t> status(nq);
  e = hipMemcpyAsync(status.data(), ds, nq * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(out_count, dc, nq * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess && k) e = hipMemcpyAsync(out_keys, dk, nq * k * 8, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess && k && !f64)
    e = hipMemcpyAsync(out_dist, dd, nq * k * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess && k && f64)
    e = hipMemcpyAsync(out_dist64, dd64, nq * k * 8, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  cleanup();
  if (e != hipSuccess) return fail(HH_DEVICE_ERROR, hipGetErrorString(e));
  return first_failure(status, first_bad);
}

int hh_brute_search_batch(hh_brute *b, const float *queries, uint64_t nq, uint32_t k,
                          const uint32_t *allow_bits, uint64_t *out_keys, float *out_dist,
                          uint32_t *out_count, uint64_t *first_bad) {
  return brute_host(b, queries, nq, k, allow_bits, out_keys, out_dist, nullptr, out_count,
                    first_bad, false);
}

int hh_brute_search_batch_f64(hh_brute *b, const float *queries, uint64_t nq, uint32_t k,
                              uint64_t *out_keys, double *out_dist, uint32_t *out_count,
                              uint64_t *first_bad) {
  return brute_host(b, queries, nq, k, nullptr, out_keys, nullptr, out_dist, out_count, first_bad,
                    true);
}

}  // extern "C"

// ------------------------------------------------ no-allocation search
// hh_search_batch_device_noalloc: the search of hh_search_batch_device with
// its whole workspace (prepared queries and codes, statuses, seeds, order,
// re-run list, visited bitsets, candidate pools, spill logs, re-run pools)
// carved from one static device array instead of allocated, for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops).  Each device holds its own copy of the array for the
// life of the process; every call on a device shares it, so calls that
// overlap on one device race.
namespace {
constexpr uint64_t kScratchBytes = 64ull << 20;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
unsigned char *scratch_addr[kMaxDevices] = {};

// The device address of `scratch` on `device` (current), looked up on the
// first call there.  The array's name is a host-side stand-in, never a
// device address: kernels and copies get this pointer.
int scratch_base(int device, unsigned char **out) {
  if (device < 0 || device >= kMaxDevices) return fail(HH_INVALID_ARGUMENT, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_addr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(HH_DEVICE_ERROR, std::string("scratch address: ") + hipGetErrorString(e));
    scratch_addr[device] = static_cast<unsigned char *>(p);
  }
  *out = scratch_addr[device];
  return HH_OK;
}

// Bump allocator over the scratch array (256-byte aligned pieces); the
// caller checks `used` against the array's size before using any piece.
struct Carve {
  unsigned char *base;
  uint64_t used;
  template <class T>
  T *take(uint64_t count) {
    used = (used + 255) / 256 * 256;
    T *p = reinterpret_cast<T *>(base + used);
    used += count * sizeof(T);
    return p;
  }
};

uint64_t round256(uint64_t b) { return (b + 255) / 256 * 256; }

}  // namespace

extern "C" {

int hh_search_batch_device_noalloc(hh_index *idx, const float *queries, uint64_t nq, uint32_t k,
                                   uint32_t ef, uint32_t precision, const uint32_t *allow_bits,
                                   uint64_t *out_keys, float *out_dist, uint32_t *out_count,
                                   int32_t *status, void *stream) {
  if (!idx) return fail(HH_INVALID_ARGUMENT, "null index");
  if (nq == 0) return HH_OK;
  if (nq > kScratchBytes)
    return fail(HH_UNSUPPORTED, "batch too large for the 64 MiB scratch of the no-allocation search");
  const SearchCall c{queries, nq, k, ef, precision, allow_bits, nullptr, k, nullptr, 0,
                     out_keys, out_dist, out_count, status};
  std::shared_lock<std::shared_mutex> rd(idx->mu);
  int st = check_call(idx, c);
  if (st) return st;
  HH_HIP(hipSetDevice(idx->device));
  unsigned char *base = nullptr;
  st = scratch_base(idx->device, &base);
  if (st) return st;
  SearchConfig cfg;
  st = search_config(idx, c, cfg);
  if (st) return st;
  const uint64_t n = idx->n, words = cfg.spill_words;
  const uint32_t stride = idx->lay.stride, cstride = std::max<uint32_t>(idx->cstride, 1);
  // pools as the allocating path sizes them; when they can drop a
  // candidate, four n-entry pools for the re-run pass
  const uint32_t pcap =
      (uint32_t)std::min<uint64_t>(kCandPool, (std::max<uint64_t>(32 * words, 64) + 63) / 64 * 64);
  const uint64_t big_cap = (std::max<uint64_t>(n, 64) + 63) / 64 * 64;
  const bool need_big = pcap < n;
  Carve cv{base, 0};
  Workspace ws;  // views into the scratch array only: nothing to release
  ws.counter = cv.take<uint32_t>(4);
  ws.qprep = cv.take<float>(nq * stride);
  ws.qcode = cv.take<int8_t>(nq * cstride);
  ws.status = status ? nullptr : cv.take<int32_t>(nq);
  ws.seed = cv.take<uint64_t>(nq);
  ws.order = cv.take<uint32_t>(nq);
  ws.redo = cv.take<uint32_t>(nq);
  ws.qpd = cv.take<float2>(nq);
  if (need_big) {
    ws.bigpool = cv.take<uint64_t>(4 * big_cap);
    ws.big_slots = 4;
    ws.big_cap = (uint32_t)big_cap;
  }
  const uint64_t fixed = round256(cv.used);
  const uint64_t per_wave = round256(words * 4) + round256((uint64_t)pcap * 8) + round256(kSpillLog * 4);
  // waves: the launch shape's, no more than one per query (whole blocks),
  // and no more than the rest of the array holds
  const uint64_t wpb = (uint64_t)cfg.wpb;
  uint64_t waves = std::min<uint64_t>(cfg.blocks * wpb, (nq + wpb - 1) / wpb * wpb);
  waves = fixed < kScratchBytes ? std::min<uint64_t>(waves, (kScratchBytes - fixed) / per_wave / wpb * wpb) : 0;
  if (waves < wpb)
    return fail(HH_UNSUPPORTED,
                "batch and index too large for the 64 MiB scratch of the no-allocation search "
                "(it holds the queries, " + std::to_string(wpb) +
                    " candidate pools and visited bitsets, and the re-run pools); "
                    "use hh_search_batch_device or a smaller batch");
  ws.spill = cv.take<uint32_t>(waves * words);
  ws.cpool = cv.take<uint64_t>(waves * pcap);
  ws.slog = cv.take<uint32_t>(waves * kSpillLog);
  if (cv.used > kScratchBytes) return fail(HH_UNSUPPORTED, "scratch layout exceeds the array");
  ws.pool_cap = pcap;
  ws.spill_slots = waves;
  ws.spill_wcap = words;
  ws.cap_q = nq;
  cfg.blocks = waves / wpb;
  hipStream_t s = (hipStream_t)stream;
  // the array keeps the last call's contents: the visited bitsets must
  // start clear (a fill kernel, so the call can be captured in a graph)
  hipLaunchKernelGGL(hhk::fill_u32_kernel, dim3(grid_for(waves * words)), dim3(256), 0, s,
                     ws.spill, waves * words, 0u);
  HH_HIP(hipGetLastError());
  return enqueue_search(idx, ws, c, cfg, s);
}

}  // extern "C"
