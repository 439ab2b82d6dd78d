// search.hip — the query preparation, the descent pass and the query
// order of the batched HNSW search on gfx950 (the layer-0 search kernel is
// in search_impl.h, instantiated by search_f32.hip, search_bf16.hip,
// search_codes.hip and search_any.hip).
#include <hip/hip_runtime.h>

#include "search_impl.h"

using namespace hhd;

namespace hhk {

// ------------------------------------------------------ descent pass
// The first pass of a two-pass launch when m <= 32: one query per 16 lanes
// (m <= 16, four per wavefront) or 32 lanes (two per wavefront).  f32 and
// bf16 rows are read and reduced by 16-lane quarters exactly as the search
// kernel's gather does (the AVX2 kernel's accumulators and lane
// combination, x86.rs:21-27, 47-77, 96-129), so every distance has the
// search kernel's bits; int8 code rows are summed one per lane (exact
// integers).  Without a per-query copy of the query in LDS, more blocks
// stay resident.  The greedy steps are those of search_graph / greedy_closest
// (src/hnsw/search.rs:49-68,146-161): the first list position holding the
// smallest distance, taken when strictly below the best.

// Integer codes: the exact integer sums over the whole code row (zero
// padded to cstride on both sides) in any order, rescaled as the search
// kernel does: prec 1 the crate's quantized distance
// (src/kernels/int8.rs:13-48, src/quantization.rs:85-92), prec 3 the
// per-dimension codes.
// WORDS: the code row's uint2 count when known at compile time (exact
// layouts: cstride = D), else 0 and cstride / 8 is read at run time.
template <int WORDS>
__device__ __forceinline__ float lane_distance_i8(const SearchArgs &P, uint32_t prec,
                                                  const QCodeInfo &qs, const uint2 *qc,
                                                  uint32_t id) {
  const int words = WORDS ? WORDS : (int)(P.cstride >> 3);
  const bool pd = prec == 3;
  gptr<uint2> row = (gptr<uint2>)((pd ? P.pd_codes : P.codes) + (size_t)id * P.cstride);
  const float rn = pd && P.metric == 0 ? G(P.pd_rnorm)[id] : 0.0f;
  int acc = 0;
  if (!pd && P.metric == 0) {
#pragma unroll 8
    for (int c = 0; c < words; c++) {
      const uint2 a = qc[c], b = row[c];
      acc += __builtin_amdgcn_sdot4((int)a.x, (int)a.x, 0, false) -
             2 * __builtin_amdgcn_sdot4((int)a.x, (int)b.x, 0, false) +
             __builtin_amdgcn_sdot4((int)b.x, (int)b.x, 0, false);
      acc += __builtin_amdgcn_sdot4((int)a.y, (int)a.y, 0, false) -
             2 * __builtin_amdgcn_sdot4((int)a.y, (int)b.y, 0, false) +
             __builtin_amdgcn_sdot4((int)b.y, (int)b.y, 0, false);
    }
  } else {
#pragma unroll 8
    for (int c = 0; c < words; c++) {
      const uint2 a = qc[c], b = row[c];
      acc = __builtin_amdgcn_sdot4((int)a.x, (int)b.x, acc, false);
      acc = __builtin_amdgcn_sdot4((int)a.y, (int)b.y, acc, false);
    }
  }
  if (pd) {
    const float t = __fmul_rn(qs.sigma, __int2float_rn(acc));
    if (P.metric == 0) return __fsub_rn(__fadd_rn(qs.qconst, rn), __fadd_rn(t, t));
    if (P.metric == 1) return -__fadd_rn(qs.qconst, t);
    return __fsub_rn(1.0f, __fadd_rn(qs.qconst, t));
  }
  const float s2 = __fmul_rn(P.sq_scale, P.sq_scale);
  if (P.metric == 0) return __fmul_rn(__uint2float_rn((uint32_t)acc), s2);
  if (P.metric == 1) return __fmul_rn(-__int2float_rn(acc), s2);
  return __fsub_rn(1.0f, __fmul_rn(__int2float_rn(acc), s2));
}

// f32 or bf16 distances from one query to the rows ids[0..cnt) (this
// query's LDS slots), written to out[0..cnt): the query's QPQ quarters take
// rows qq, qq + QPQ, ..., B at a time with all of a round's loads issued
// first, each row read by a quarter's 16 lanes in the search kernel's lane
// layout and reduced as it reduces (the same bits as its gather_round).
// Sixteen lanes reading one row keep a load instruction on 4 rows per wave
// instead of 64.
template <int TMAX, bool EXTRA, bool EXACT, class E, int QPQ>
__device__ __forceinline__ void descent_rows(const SearchArgs &P, const QLanes<TMAX, EXTRA> &qv,
                                             const E *rows, const uint32_t *ids, int cnt, int qq,
                                             float *out) {
  constexpr int B = rows_per_quarter<TMAX>();
  const int lane = lane_id(), i = lane & 15, quarter = lane >> 4;
  const Shape s{P.T, P.U, P.R, P.stride};
  for (int base = qq; base < cnt; base += B * QPQ) {
    QLanes<TMAX, EXTRA> x[B];
#pragma unroll
    for (int b = 0; b < B; b++) {
      const int row = base + b * QPQ;
      const uint32_t rid = ids[row < cnt ? row : base];
      if constexpr (!EXACT) {
#pragma unroll
        for (int t = 0; t < TMAX; t++) x[b].lo[t] = x[b].hi[t] = 0.0f;
        x[b].e[0] = x[b].r = 0.0f;
        if constexpr (EXTRA) x[b].e[1] = x[b].e[2] = 0.0f;
      }
      load_q<TMAX, EXTRA, E, EXACT>(rows + (size_t)rid * P.stride, s, i, x[b]);
    }
#pragma unroll
    for (int b = 0; b < B; b++) {
      const int row = base + b * QPQ;
      if (row >= cnt) break;
      float lo, hi, tail;
      if (P.metric == 0) quarter_acc<TMAX, EXTRA, true, EXACT>(qv, x[b], s, i, lo, hi, tail);
      else quarter_acc<TMAX, EXTRA, false, EXACT>(qv, x[b], s, i, lo, hi, tail);
      const float h = finish_metric(quarter_reduce<EXTRA>(lo, hi, tail, s.R, quarter), P.metric);
      if (i == 0) out[row] = h;
    }
  }
}

template <int TMAX, int LPQ, bool EXACT>
__global__ void __launch_bounds__(descent_threads<TMAX>()) descent_kernel(const SearchArgs P) {
  static_assert(!EXACT || TMAX % 4 == 0, "exact layout");
  constexpr int kCodeWords = descent_code_words<TMAX>();
  constexpr int kQueries = descent_threads<TMAX>() / LPQ;
  constexpr bool kExtra = !EXACT;
  constexpr int kQpq = LPQ / 16;  // quarters per query
  __shared__ __align__(16) uint2 qcs[kQueries][kCodeWords];
  __shared__ uint32_t sid[kQueries][LPQ];  // a step's neighbour ids
  __shared__ float sdist[kQueries][LPQ];   // and their distances
  const int il = lane_id() & (LPQ - 1);
  const int qslot = (int)(threadIdx.x / LPQ);
  const uint64_t qi = (uint64_t)blockIdx.x * kQueries + qslot;
  const Shape shp{P.T, P.U, P.R, P.stride};
  unsigned long long up = 0;
  if (qi < P.nq) {
    uint32_t k = P.k, prec = P.precision;
    if (P.specs) {
      const hh_query_spec sp = G(P.specs)[qi];
      k = sp.k;
      prec = sp.precision;
    }
    const int status = G(P.status)[qi];
    if (status != 0 || k == 0 || P.live == 0 || P.top < 0) {
      GM(P.seed)[qi] = ((uint64_t)__float_as_uint(-INFINITY) << 32) | kInvalid;
    } else {
      // the query (prepared, row layout) in the quarter-lane layout, and
      // its codes into this slot
      QLanes<TMAX, kExtra> qv;
      if constexpr (!EXACT) {
#pragma unroll
        for (int t = 0; t < TMAX; t++) qv.lo[t] = qv.hi[t] = 0.0f;
        qv.e[0] = qv.r = 0.0f;
        if constexpr (kExtra) qv.e[1] = qv.e[2] = 0.0f;
      }
      load_q<TMAX, kExtra, float, EXACT>(P.qprep + qi * P.stride, shp, il & 15, qv);
      const bool codes = prec == 1 || prec == 3;
      QCodeInfo qk{0, 1.0f, 0.0f};
      if (prec == 1 || prec == 3) {
        gptr<uint2> csrc = (gptr<uint2>)(P.qcode + qi * P.cstride);
        for (int c = il; c < (int)(P.cstride / 8); c += LPQ) qcs[qslot][c] = csrc[c];
        if (prec == 3) {
          const float2 v = G(P.qpd)[qi];
          qk.sigma = v.x;
          qk.qconst = v.y;
        }
      }
      wave_sync();
      const uint2 *qc = qcs[qslot];
      uint32_t *ids = sid[qslot];
      float *dists = sdist[qslot];
      // f32 / bf16 distances of rows ids[0..cnt) into dists (quarter lanes)
      auto rows_dist = [&](int cnt) {
        if (prec == 2)
          descent_rows<TMAX, kExtra, EXACT, uint16_t, kQpq>(P, qv, P.bf16, ids, cnt, il >> 4, dists);
        else
          descent_rows<TMAX, kExtra, EXACT, float, kQpq>(P, qv, P.rows, ids, cnt, il >> 4, dists);
        wave_sync();
      };
      uint32_t cur = P.entry;
      uint64_t cur_off = G(P.upper_off)[cur];
      float best;
      if (codes) {
        best = lane_distance_i8<EXACT ? 4 * TMAX : 0>(P, prec, qk, qc, cur);
      } else {
        ids[il] = cur;
        wave_sync();
        rows_dist(1);
        best = dists[0];
      }
      up++;
      for (int layer = P.top; layer >= 1; layer--) {
        for (;;) {
          gptr<uint32_t> blk = G(P.upper) + cur_off + (size_t)(layer - 1) * (P.m + 1);
          uint32_t cnt = blk[0];
          if (cnt > P.m) cnt = P.m;
          const bool valid = (uint32_t)il < cnt;
          const uint32_t nid = valid ? blk[1 + il] : kInvalid;
          const uint64_t noff = valid ? (uint64_t)G(P.upper_off)[nid] : 0;
          float md;
          if (codes) {
            md = valid ? lane_distance_i8<EXACT ? 4 * TMAX : 0>(P, prec, qk, qc, nid) : 0.0f;
          } else {
            ids[il] = nid;
            wave_sync();
            rows_dist((int)cnt);
            md = valid ? dists[il] : 0.0f;
          }
          int mi = valid ? il : LPQ;
#pragma unroll
          for (int m = 1; m < LPQ; m <<= 1) {
            const float od = __shfl_xor(md, m, LPQ);
            const int oi = __shfl_xor(mi, m, LPQ);
            const bool take = oi < LPQ && (mi == LPQ || od < md || (!(md < od) && oi < mi));
            if (take) {
              md = od;
              mi = oi;
            }
          }
          up += cnt;
          const int src_lane = mi & (LPQ - 1);
          const uint32_t bn = (uint32_t)__shfl((int)nid, src_lane, LPQ);
          const uint32_t bo_lo = (uint32_t)__shfl((int)(uint32_t)noff, src_lane, LPQ);
          const uint32_t bo_hi = (uint32_t)__shfl((int)(uint32_t)(noff >> 32), src_lane, LPQ);
          if (mi < LPQ && md < best) {
            best = md;
            cur = bn;
            cur_off = ((uint64_t)bo_hi << 32) | bo_lo;
          } else {
            break;
          }
        }
      }
      GM(P.seed)[qi] = ((uint64_t)__float_as_uint(best) << 32) | cur;
    }
  }
  // one count per query (its first lane holds the query's total), summed
  // over the wave and added by lane 0 into the wave's line of the counters
  uint32_t wsum = il == 0 ? (uint32_t)up : 0u;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) wsum += (uint32_t)__shfl_xor((int)wsum, off, 64);
  if ((threadIdx.x & 63) == 0 && wsum && P.stats) {
    const uint32_t wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    atomicAdd(&P.stats[(size_t)(wave % kStatLines) * kStatStride + 2], (unsigned long long)wsum);
  }
}

// Longest-first query order for the two-pass launch: a histogram of the
// seeds' order bins in LDS, bins taken from the largest descent distance
// down (queries within a bin in any order); then resets the search kernel's
// query counter for the second pass.
// Mixed batches: by an estimate of the query's layer-0 work, ef (raised
// to k) over its filter's pass rate (a query keeps expanding until ef
// admitted nodes are found), on a log scale; plain batches: by the
// descent distance.
__device__ __forceinline__ uint32_t query_order_bin(const uint64_t *seed, const hh_query_spec *specs,
                                                    const float *rate, uint32_t nrate, uint64_t i) {
  const uint32_t b = order_bin(__uint_as_float((uint32_t)(G(seed)[i] >> 32)));
  if (!specs) return b;
  const hh_query_spec sp = G(specs)[i];
  uint32_t ef = sp.ef > sp.k ? sp.ef : sp.k;
  ef = ef < 1 ? 1 : (ef > 512 ? 512 : ef);
  float r = sp.filter < nrate ? rate[sp.filter] : 1.0f;
  r = r < (1.0f / 256.0f) ? (1.0f / 256.0f) : r;
  const float cost = __log2f((float)ef / r);  // 0 .. 17
  const int c = (int)(cost * 224.0f);
  return (uint32_t)(c < 0 ? 0 : (c > 4095 ? 4095 : c));
}

__global__ void __launch_bounds__(1024) order_queries_kernel(const uint64_t *seed,
                                                             const hh_query_spec *specs, uint64_t nq,
                                                             const uint32_t *filters, uint32_t filter_words,
                                                             uint32_t nfilters, uint32_t *order,
                                                             uint32_t *counter) {
  constexpr uint32_t kRated = 256;     // filters with a pass-rate estimate
  constexpr uint32_t kSampled = 4096;  // bitmap words sampled per filter
  __shared__ float rate[kRated];
  __shared__ uint32_t ones[kRated];
  const uint32_t nrate = specs && filters ? (nfilters < kRated ? nfilters : kRated) : 0u;
  {
    const uint32_t t0 = threadIdx.x;
    if (t0 < kRated) ones[t0] = 0u;
    __syncthreads();
    const uint32_t sw = filter_words < kSampled ? filter_words : kSampled;
    if (nrate && sw) {
      // threads t, t + nrate, ... of filter t % nrate sum strided words
      const uint32_t per = 1024 / nrate;
      const uint32_t f = t0 % nrate, part = t0 / nrate;
      if (part < per) {
        uint32_t sum = 0;
        for (uint32_t w = part; w < sw; w += per)
          sum += __popc(G(filters)[(size_t)f * filter_words + w]);
        atomicAdd(&ones[f], sum);
      }
    }
    __syncthreads();
    if (t0 < nrate) rate[t0] = sw ? (float)ones[t0] / (32.0f * (float)sw) : 1.0f;
  }
  static_assert(kOrderBins == 4 * 1024, "four bins per thread");
  __shared__ uint32_t off[kOrderBins];
  __shared__ uint32_t part[1024];
  const uint32_t t = threadIdx.x;
#pragma unroll
  for (int i = 0; i < 4; i++) off[4 * t + i] = 0u;
  __syncthreads();
  for (uint64_t i = t; i < nq; i += 1024)
    atomicAdd(&off[query_order_bin(seed, specs, rate, nrate, i)], 1u);
  __syncthreads();
  // thread t owns the descending-order bins 4t..4t+3 (bin kOrderBins-1-r)
  uint32_t c[4], sum = 0;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    c[i] = off[kOrderBins - 1 - (4 * t + i)];
    sum += c[i];
  }
  part[t] = sum;
  __syncthreads();
  for (uint32_t o = 1; o < 1024; o <<= 1) {
    const uint32_t v = t >= o ? part[t - o] : 0u;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  uint32_t run = part[t] - sum;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    off[kOrderBins - 1 - (4 * t + i)] = run;
    run += c[i];
  }
  if (t == 0) GM(counter)[0] = 0u;
  __syncthreads();
  for (uint64_t i = t; i < nq; i += 1024) {
    const uint32_t pos = atomicAdd(&off[query_order_bin(seed, specs, rate, nrate, i)], 1u);
    if (pos < nq) GM(order)[pos] = (uint32_t)i;
  }
}

// Validation and preparation of every query (src/vector.rs:22-45,90-104),
// spec checks (src/hnsw/mod.rs:468-473), int8 encoding of the query
// (src/quantization.rs:73-81), then the row layout.  One wavefront per
// query: lanes load the query coalesced and square it in f64 (exact for f32
// inputs); the f64 norm is then summed left to right, as the crate does, by
// a uniform chain over the lanes' squares.
__device__ __forceinline__ void prepare_one(const PrepArgs &A, uint64_t qi, int lane) {
  gptr<float> v = G(A.queries) + qi * A.dim;
  int status = 0;
  double norm_sq = 0.0;
  for (uint32_t base = 0; base < A.dim; base += 64) {
    const uint32_t e = base + (uint32_t)lane;
    const float x = e < A.dim ? v[e] : 0.0f;
    if (__ballot(!isfinite(x))) {
      status = 4;
      break;
    }
    const double sq = __dmul_rn((double)x, (double)x);
    const int cnt = A.dim - base < 64 ? (int)(A.dim - base) : 64;
    const int lo = __double2loint(sq), hi = __double2hiint(sq);
    for (int l = 0; l < cnt; l++)
      norm_sq = __dadd_rn(norm_sq, __hiloint2double(__builtin_amdgcn_readlane(hi, l),
                                                    __builtin_amdgcn_readlane(lo, l)));
  }
  if (status == 0 && norm_sq > 3.4028234663852886e38 / 8.0) status = 4;
  if (status == 0 && A.metric == 2 && norm_sq == 0.0) status = 5;
  uint32_t prec = A.precision;
  if (A.specs) {
    const hh_query_spec sp = G(A.specs)[qi];
    prec = sp.precision;
    if (status == 0) {
      if (sp.precision > 3) status = 103;
      else if (sp.k > A.k_stride) status = 103;
      else if (sp.filter != 0xFFFFFFFFu && sp.filter >= A.nfilters) status = 103;
      else if ((sp.ef > sp.k ? sp.ef : sp.k) > A.ef_cap) status = 102;
    }
  }
  if (status == 0 && prec == 1 && !A.has_codes) status = 8;
  if (status == 0 && prec == 2 && !A.has_bf16) status = 102;
  if (status == 0 && prec == 3 && !A.has_pd) status = 8;
  GM(A.status)[qi] = status;
  if (status != 0) return;
  const double inv = A.metric == 2 ? 1.0 / sqrt(norm_sq) : 1.0;
  gmut<float> out = GM(A.qprep) + qi * A.stride;
  gmut<int8_t> code = A.qcode ? GM(A.qcode) + qi * A.cstride : nullptr;
  const float sinv = 1.0f / A.sq_scale;
  const bool encode = code && prec == 1;
  for (uint32_t e = (uint32_t)lane; e < A.stride; e += 64) {
    float x = 0.0f;
    uint32_t pos = e;
    if (e < A.dim) {
      x = v[e];
      if (A.metric == 2) x = (float)__dmul_rn((double)x, inv);
      if (e < 32u * A.T) {
        const uint32_t t = e >> 5, jj = e & 31, b = t >> 2;
        const uint32_t tw = A.T - 4 * b < 4 ? A.T - 4 * b : 4;
        pos = b * 128 + jj * tw + (t & 3);
      }
    }
    out[pos] = x;
  }
  if (code && prec == 3) {
    // per-dimension codes: weights w_d = (x_d - mid_d) * scale_d (L2) or
    // x_d * scale_d, sigma = max |w_d| / 127 (1 when zero), code_d =
    // clamp(round_half_away(w_d * (1 / sigma)), -127, 127); qconst =
    // sum_d (x_d - mid_d)^2 (L2) or sum_d x_d * mid_d in f64, summed in
    // dimension order by a uniform chain, rounded to f32 once.
    float wmax = 0.0f;
    double qconst = 0.0;
    for (uint32_t base = 0; base < A.dim; base += 64) {
      const uint32_t e = base + (uint32_t)lane;
      float w = 0.0f;
      double term = 0.0;
      if (e < A.dim) {
        float x = v[e];
        if (A.metric == 2) x = (float)__dmul_rn((double)x, inv);
        const float md = G(A.pd_mid)[e], sc = G(A.pd_scale)[e];
        if (A.metric == 0) {
          w = __fmul_rn(__fsub_rn(x, md), sc);
          const double t = __dsub_rn((double)x, (double)md);
          term = __dmul_rn(t, t);
        } else {
          w = __fmul_rn(x, sc);
          term = __dmul_rn((double)x, (double)md);
        }
      }
      wmax = fmaxf(wmax, fabsf(w));
      const int cnt = A.dim - base < 64 ? (int)(A.dim - base) : 64;
      const int lo = __double2loint(term), hi = __double2hiint(term);
      for (int l = 0; l < cnt; l++)
        qconst = __dadd_rn(qconst, __hiloint2double(__builtin_amdgcn_readlane(hi, l),
                                                    __builtin_amdgcn_readlane(lo, l)));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) wmax = fmaxf(wmax, __shfl_xor(wmax, o, 64));
    float sigma = __fdiv_rn(wmax, 127.0f);
    if (!(sigma > 0.0f)) sigma = 1.0f;
    const float winv = __fdiv_rn(1.0f, sigma);
    for (uint32_t e = (uint32_t)lane; e < A.cstride; e += 64) {
      int8_t cv = 0;
      if (e < A.dim) {
        float x = v[e];
        if (A.metric == 2) x = (float)__dmul_rn((double)x, inv);
        const float md = G(A.pd_mid)[e], sc = G(A.pd_scale)[e];
        const float w = A.metric == 0 ? __fmul_rn(__fsub_rn(x, md), sc) : __fmul_rn(x, sc);
        float c = roundf(__fmul_rn(w, winv));
        c = fminf(127.0f, fmaxf(-127.0f, c));
        cv = (int8_t)c;
      }
      code[e] = cv;
    }
    if (lane == 0) GM(A.qpd)[qi] = make_float2(sigma, (float)qconst);
  }
  if (encode) {
    for (uint32_t e = (uint32_t)lane; e < A.cstride; e += 64) {
      int8_t cv = 0;
      if (e < A.dim) {
        float x = v[e];
        if (A.metric == 2) x = (float)__dmul_rn((double)x, inv);
        float c = roundf(__fmul_rn(__fsub_rn(x, A.sq_offset[e]), sinv));
        c = fminf(127.0f, fmaxf(-127.0f, c));
        cv = (int8_t)c;
      }
      code[e] = cv;
    }
  }
}

// grid: any number of 256-thread blocks (four queries per block at a time)
__global__ void __launch_bounds__(256) prepare_queries_kernel(const PrepArgs A) {
  const int lane = (int)(threadIdx.x & 63);
  if (blockIdx.x == 0 && threadIdx.x < 4 && A.counter) GM(A.counter)[threadIdx.x] = 0u;
  for (uint64_t qi = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6); qi < A.nq;
       qi += (uint64_t)gridDim.x * 4)
    prepare_one(A, qi, lane);
}

#define HH_DESCENT_INST(TM, XA)                                             \
  template __global__ void descent_kernel<TM, 16, XA>(const SearchArgs);    \
  template __global__ void descent_kernel<TM, 32, XA>(const SearchArgs);
HH_DESCENT_INST(4, true)
HH_DESCENT_INST(8, true)
HH_DESCENT_INST(32, true)
HH_DESCENT_INST(4, false)
HH_DESCENT_INST(8, false)
HH_DESCENT_INST(32, false)

}  // namespace hhk
