// search.hip — batched HNSW search on gfx950.
//
// One wavefront runs one query end to end (search_impl,
// src/hnsw/mod.rs:460-511): greedy descent through the upper layers
// (greedy_closest, src/hnsw/search.rs:49-68), then the ef-bounded layer-0
// search (search_layer, src/hnsw/search.rs:81-142), optional exact re-rank,
// and the final (distance, key) ordering.  Waves are persistent and pull
// query indices from a counter, so uneven per-query work (ef, filters)
// balances itself.
//
// Per wave: the candidate list C and the result list W are sorted arrays of
// candidate words in LDS, each neighbour batch merged in one pass; C spills
// its largest entries to a per-wave HBM pool; the visited set is an LDS hash
// table that moves to a per-wave HBM bitset if a query visits more nodes
// than the table holds; neighbour rows are gathered in rounds of 16, one row
// per 16-lane quarter at a time, 512 contiguous bytes per row at D=128.  The
// adjacency row of the next candidate is fetched while the current one is
// expanded.
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "search_kernels.h"

using namespace hhd;

namespace hhk {

struct Stats {
  unsigned long long q, dist, up, exp, rr, ovf, spill, cspill, redo;
};

// A candidate as one 64-bit word ordered like (distance total order, id).
__device__ __forceinline__ uint64_t ckey(float d, uint32_t id) {
  return ((uint64_t)((uint32_t)tkey(d) ^ 0x80000000u) << 32) | id;
}
__device__ __forceinline__ float ckey_d(uint64_t k) {
  const int32_t t = (int32_t)((uint32_t)(k >> 32) ^ 0x80000000u);
  return __int_as_float(t ^ (int32_t)(((uint32_t)(t >> 31)) >> 1));
}

// ---------------------------------------------------------------- lists
// C (candidates) and W (results) are sorted arrays of candidate words
// (`ckey`) in the wave's LDS.  A neighbour batch is merged in one pass: the
// list is read into registers (element i in lane i % 64, chunk i / 64), every
// element's merged position is its rank, counted with ballots against the
// batch's few new elements (a scalar loop over them, independent steps),
// then all lanes write.  The caller issues the merges with all 64 lanes
// (wave-uniform arguments).

// Rank of each batch element (lanes in `amask`, word `akey`) among the
// list elements e[0..CPL) (valid where chunk*64 + lane < size), plus the
// number of batch elements ordering before it; and, for every list
// element, its index plus the number of batch elements ordering before it.
// N: the chunks that can hold list elements (chunks past the list rank
// nothing), a compile-time count so the per-element loop stays straight.
template <int CPL, int N>
__device__ __forceinline__ uint32_t merge_ranks_n(const uint64_t (&e)[CPL], uint32_t (&ep)[CPL],
                                                  uint32_t size, uint64_t amask, uint64_t akey) {
  const int lane = lane_id();
  uint32_t ap = 0;
#pragma unroll
  for (int c = 0; c < CPL; c++) ep[c] = c * 64 + lane;
  for (uint64_t m = amask; m; m &= m - 1) {
    const int j = __ffsll((unsigned long long)m) - 1;
    const uint64_t kj = rl64(akey, j);
    uint32_t r = 0;
#pragma unroll
    for (int c = 0; c < N; c++) {
      const bool valid = c * 64 + lane < (int)size;
      r += (uint32_t)popc64(__ballot(valid && e[c] < kj));
      ep[c] += kj < e[c] ? 1u : 0u;
    }
    ap += (kj < akey ? 1u : 0u) + (lane == j ? r : 0u);
  }
  return ap;
}
template <int CPL>
__device__ __forceinline__ uint32_t merge_ranks(const uint64_t (&e)[CPL], uint32_t (&ep)[CPL],
                                                uint32_t size, uint64_t amask, uint64_t akey) {
  if (size <= 64) return merge_ranks_n<CPL, 1>(e, ep, size, amask, akey);
  if constexpr (CPL > 2) {
    if (size <= 128) return merge_ranks_n<CPL, 2>(e, ep, size, amask, akey);
  }
  if constexpr (CPL > 4) {
    if (size <= 256) return merge_ranks_n<CPL, 4>(e, ep, size, amask, akey);
  }
  if constexpr (CPL > 8) {
    if (size <= 512) return merge_ranks_n<CPL, 8>(e, ep, size, amask, akey);
  }
  return merge_ranks_n<CPL, CPL>(e, ep, size, amask, akey);
}

// The crate's sequential acceptance (search.rs:118-132) counted for a
// neighbour batch: for each surviving neighbour (lanes in sm, distance d),
// the results at or below d plus the admitted survivors before it at or
// below d.  N: the chunks of W that can hold results.
template <int SW, int N>
__device__ __forceinline__ uint32_t accept_counts(const uint64_t (&we)[SW], uint32_t wsize,
                                                  uint64_t sm, uint64_t am, float myd) {
  const int lane = lane_id();
  uint32_t below = 0;
  for (uint64_t m = sm; m; m &= m - 1) {
    const int j = __ffsll((unsigned long long)m) - 1;
    const float dj = rlf(myd, j);
    uint32_t cw = 0;  // results at or below d_j
#pragma unroll
    for (int c = 0; c < N; c++)
      cw += (uint32_t)popc64(__ballot(c * 64 + lane < (int)wsize && !(ckey_d(we[c]) > dj)));
    below += (lane == j ? cw : 0u) + (((am >> j) & 1) && j < lane && !(dj > myd) ? 1u : 0u);
  }
  return below;
}

// W <- the `cap` smallest of W[0..size) (held in e[]) and the batch
// elements in `amask`; returns the new worst (largest) word.
template <int WPL>
__device__ __forceinline__ uint64_t w_merge(uint64_t *W, const uint64_t (&e)[WPL], uint32_t &size,
                                            uint32_t cap, uint64_t amask, uint64_t akey) {
  const int lane = lane_id();
  const bool mine = (amask >> lane) & 1;
  uint32_t ep[WPL];
  const uint32_t ap = merge_ranks<WPL>(e, ep, size, amask, akey);
  const uint32_t total = size + (uint32_t)popc64(amask);
  const uint32_t nsize = total < cap ? total : cap;
  uint64_t last = ~0ull;
  const uint64_t la = __ballot(mine && ap == nsize - 1);
  if (la) last = rl64(akey, __ffsll((unsigned long long)la) - 1);
#pragma unroll
  for (int c = 0; c < WPL; c++) {
    if (c > 0 && c * 64 >= (int)size) break;  // chunks past W hold nothing
    const bool valid = c * 64 + lane < (int)size;
    const uint64_t lb = __ballot(valid && ep[c] == nsize - 1);
    if (lb) last = rl64(e[c], __ffsll((unsigned long long)lb) - 1);
    if (valid && ep[c] < nsize) W[ep[c]] = e[c];
  }
  if (mine && ap < nsize) W[ap] = akey;
  size = nsize;
  return last;
}

// C[head..head+size) (held in e[]) merged with the batch elements in
// `amask` into C[0..): the merged order's first `keep` entries stay, keep =
// min(ccap, entries below pmin); the rest go to the pool unless dead
// (results full and distance above the worst), and pmin stays a lower
// bound of the pool.  With results full, dead entries are also trimmed from
// C's end.
template <int CPL>
__device__ __forceinline__ void c_merge(uint64_t *C, const uint64_t (&e)[CPL], uint32_t &head,
                                        uint32_t &size, uint32_t ccap, uint64_t amask,
                                        uint64_t akey, bool wfull, float worst,
                                        gmut<uint64_t> pool, bool has_pool, uint32_t pool_cap,
                                        uint32_t &pn, uint64_t &pmin, unsigned long long &ovf,
                                        unsigned long long &moved) {
  const int lane = lane_id();
  const bool mine = (amask >> lane) & 1;
  uint32_t ep[CPL];
  const uint32_t ap = merge_ranks<CPL>(e, ep, size, amask, akey);
  // every entry of C is below pmin; batch entries at or above it go out
  const uint32_t below = size + (uint32_t)popc64(__ballot(mine && akey < pmin));
  const uint32_t keep = below < ccap ? below : ccap;
  wave_sync();  // every lane has read C before any writes
  // the smallest entry leaving C bounds the pool from below
  const uint64_t at_keep_a = __ballot(mine && ap == keep);
  uint64_t first_out = at_keep_a ? rl64(akey, __ffsll((unsigned long long)at_keep_a) - 1) : ~0ull;
  uint32_t live = 0;  // kept entries that are not dead
#pragma unroll
  for (int c = 0; c < CPL; c++) {
    if (c > 0 && c * 64 >= (int)size) break;  // chunks past C hold nothing
    const bool valid = c * 64 + lane < (int)size;
    const uint64_t b = __ballot(valid && ep[c] == keep);
    if (b) first_out = rl64(e[c], __ffsll((unsigned long long)b) - 1);
    const bool dead = wfull && ckey_d(e[c]) > worst;
    if (valid && ep[c] < keep) C[ep[c]] = e[c];
    live += (uint32_t)popc64(__ballot(valid && ep[c] < keep && !dead));
    const bool out = valid && ep[c] >= keep && !dead;
    const uint64_t om = __ballot(out);
    if (om) {
      const uint32_t room = has_pool ? pool_cap - pn : 0u;
      const uint32_t slot = (uint32_t)mbcnt(om);
      if (out && slot < room) pool[pn + slot] = e[c];
      const uint32_t no = (uint32_t)popc64(om), put = no < room ? no : room;
      pn += put;
      moved += put;
      ovf += no - put;
    }
  }
  {
    const bool dead = wfull && ckey_d(akey) > worst;
    if (mine && ap < keep) C[ap] = akey;
    live += (uint32_t)popc64(__ballot(mine && ap < keep && !dead));
    const bool out = mine && ap >= keep && !dead;
    const uint64_t om = __ballot(out);
    if (om) {
      const uint32_t room = has_pool ? pool_cap - pn : 0u;
      const uint32_t slot = (uint32_t)mbcnt(om);
      if (out && slot < room) pool[pn + slot] = akey;
      const uint32_t no = (uint32_t)popc64(om), put = no < room ? no : room;
      pn += put;
      moved += put;
      ovf += no - put;
    }
  }
  if (first_out < pmin && !(wfull && ckey_d(first_out) > worst)) pmin = first_out;
  wave_sync();
  head = 0;
  size = wfull ? live : keep;  // dead entries form C's tail: trimmed
}

// Loads list entries [base, base + size) into registers (element i in lane
// i % 64, chunk i / 64; ~0 past the end).
template <int CPL>
__device__ __forceinline__ void list_load(const uint64_t *L, uint32_t size, uint64_t (&e)[CPL]) {
  const int lane = lane_id();
#pragma unroll
  for (int c = 0; c < CPL; c++) {
    const uint32_t i = c * 64 + lane;
    e[c] = i < size ? L[i] : ~0ull;
  }
}

// f32 or bf16 rows: B rows per quarter per round, all of a round's loads
// issued before any is used (rows past cnt repeat the last id, so every
// lane loads unconditionally and no branch splits the loads).
template <int TMAX>
constexpr int rows_per_quarter() { return TMAX <= 4 ? 4 : (TMAX <= 8 ? 2 : 1); }

template <int TMAX, bool EXTRA, bool EXACT, class E>
__device__ __forceinline__ void gather_round(const SearchArgs &P, const QLanes<TMAX, EXTRA> &q,
                                             const E *rows, const uint32_t *ids, int cnt,
                                             int base, float *out) {
  const int lane = lane_id(), i = lane & 15, quarter = lane >> 4;
  const Shape s{P.T, P.U, P.R, P.stride};
  constexpr int B = rows_per_quarter<TMAX>();
  // groups of four rows (one per quarter) this round needs: a group wholly
  // past cnt is skipped (wave-uniform), so at most three slots are wasted
  const int groups = (cnt - base + 3) >> 2;
  {
    uint32_t rid[B];
#pragma unroll
    for (int b = 0; b < B; b++) {
      const int row = base + 4 * b + quarter;
      rid[b] = ids[row < cnt ? row : cnt - 1];
    }
    QLanes<TMAX, EXTRA> x[B];
#pragma unroll
    for (int b = 0; b < B; b++) {
      if constexpr (!EXACT) {
#pragma unroll
        for (int t = 0; t < TMAX; t++) x[b].lo[t] = x[b].hi[t] = 0.0f;
        x[b].e[0] = x[b].r = 0.0f;
        if constexpr (EXTRA) x[b].e[1] = x[b].e[2] = 0.0f;
      }
      if (b == 0 || b < groups) load_q<TMAX, EXTRA, E, EXACT>(rows + (size_t)rid[b] * P.stride, s, i, x[b]);
    }
#pragma unroll
    for (int b = 0; b < B; b++) {
      if (b > 0 && b >= groups) break;
      const int row = base + 4 * b + quarter;
      float lo, hi, tail;
      if (P.metric == 0) quarter_acc<TMAX, EXTRA, true, EXACT>(q, x[b], s, i, lo, hi, tail);
      else quarter_acc<TMAX, EXTRA, false, EXACT>(q, x[b], s, i, lo, hi, tail);
      const float h = finish_metric(quarter_reduce<EXTRA>(lo, hi, tail, s.R, quarter), P.metric);
      if (i == 0 && row < cnt) out[row] = h;
    }
  }
}

// The first round is straight-line code, so its loads issue right after
// the caller's (no loop-header wait for earlier loads in flight).
template <int TMAX, bool EXTRA, bool EXACT, class E>
__device__ __forceinline__ void gather_rows(const SearchArgs &P, const QLanes<TMAX, EXTRA> &q,
                                            const E *rows, const uint32_t *ids, int cnt,
                                            float *out) {
  constexpr int R = 4 * rows_per_quarter<TMAX>();
  gather_round<TMAX, EXTRA, EXACT, E>(P, q, rows, ids, cnt, 0, out);
  for (int base = R; base < cnt; base += R)
    gather_round<TMAX, EXTRA, EXACT, E>(P, q, rows, ids, cnt, base, out);
}

// 128-byte code chunks a lane group holds: cstride / 128 for the exact
// layouts, else the most any dim of that TMAX needs (dim <= 32 * TMAX + 31,
// and dim <= 1024).
template <int TMAX, bool EXACT>
constexpr int code_chunks() {
  return EXACT ? TMAX / 4 : ((TMAX * 32 + 31 + 127) / 128 < 8 ? (TMAX * 32 + 31 + 127) / 128 : 8);
}

// Per-query constants of the integer traversals: the crate's L2 (sum of
// the query code's squares, so a row costs two exact dot products per word:
// sum (a-b)^2 = sum a^2 + sum (b^2 - 2ab)) and the per-dimension codes
// (sigma and qconst of the query's weights).
struct QCodeInfo {
  int qaa;
  float sigma, qconst;
};

// Code rows a quarter gathers per round: 8 (32 per round) while the codes
// take at most 2 x 128 bytes per row, 4 beyond.
template <int CC>
constexpr int code_rows_per_quarter() { return CC <= 2 ? 8 : 4; }

// Integer-code distances of rows ids[0..cnt): prec 1 the crate's codes
// (exact sums, src/kernels/int8.rs:13-48, rescaled as
// src/quantization.rs:85-92), prec 3 the per-dimension codes (exact
// sum_d qcode_d * code_d, then t = sigma * acc and L2 (qconst + rnorm) -
// (t + t), inner product -(qconst + t), cosine 1 - (qconst + t)).  A
// quarter takes several rows per round, every load issued first.
template <int CC>
__device__ __forceinline__ void gather_codes(const SearchArgs &P, const uint2 (&qc)[CC],
                                             const QCodeInfo &qs, int prec, const uint32_t *ids,
                                             int cnt, float *out, bool exact_layout) {
  const int lane = lane_id(), i = lane & 15, quarter = lane >> 4;
  constexpr int B = code_rows_per_quarter<CC>();
  const bool pd = prec == 3;
  const int8_t *codes = pd ? P.pd_codes : P.codes;
  const int chunks = (int)(P.cstride >> 7);
  const bool l2 = P.metric == 0;
  const float s2 = __fmul_rn(P.sq_scale, P.sq_scale);
  for (int base = 0; base < cnt; base += 4 * B) {
    const int groups = (cnt - base + 3) >> 2;  // four-row groups needed (wave-uniform)
    uint2 cb[B][CC];
    float rn[B];
#pragma unroll
    for (int b = 0; b < B; b++) {
      const int row = base + 4 * b + quarter;
      const uint32_t id = ids[row < cnt ? row : cnt - 1];
      gptr<uint2> cr = (gptr<uint2>)(codes + (size_t)id * P.cstride);
      const bool need = b == 0 || b < groups;
#pragma unroll
      for (int c = 0; c < CC; c++)
        cb[b][c] = need && (exact_layout || c < chunks) ? (uint2)cr[c * 16 + i] : make_uint2(0u, 0u);
      rn[b] = need && pd && l2 ? G(P.pd_rnorm)[id] : 0.0f;
    }
#pragma unroll
    for (int b = 0; b < B; b++) {
      if (b > 0 && b >= groups) break;
      const int row = base + 4 * b + quarter;
      int acc = 0;
      if (!pd && l2) {
        int ab = 0, bb = 0;
#pragma unroll
        for (int c = 0; c < CC; c++) {
          ab = __builtin_amdgcn_sdot4((int)qc[c].x, (int)cb[b][c].x, ab, false);
          ab = __builtin_amdgcn_sdot4((int)qc[c].y, (int)cb[b][c].y, ab, false);
          bb = __builtin_amdgcn_sdot4((int)cb[b][c].x, (int)cb[b][c].x, bb, false);
          bb = __builtin_amdgcn_sdot4((int)cb[b][c].y, (int)cb[b][c].y, bb, false);
        }
        acc = bb - 2 * ab;
      } else {
#pragma unroll
        for (int c = 0; c < CC; c++) {
          acc = __builtin_amdgcn_sdot4((int)qc[c].x, (int)cb[b][c].x, acc, false);
          acc = __builtin_amdgcn_sdot4((int)qc[c].y, (int)cb[b][c].y, acc, false);
        }
      }
      acc = quarter_sum_i(acc);
      if (i == 0 && row < cnt) {
        float v;
        if (pd) {
          const float t = __fmul_rn(qs.sigma, __int2float_rn(acc));
          if (l2) v = __fsub_rn(__fadd_rn(qs.qconst, rn[b]), __fadd_rn(t, t));
          else if (P.metric == 1) v = -__fadd_rn(qs.qconst, t);
          else v = __fsub_rn(1.0f, __fadd_rn(qs.qconst, t));
        } else if (l2) {
          v = __fmul_rn(__uint2float_rn((uint32_t)(acc + qs.qaa)), s2);
        } else if (P.metric == 1) {
          v = __fmul_rn(-__int2float_rn(acc), s2);
        } else {
          v = __fsub_rn(1.0f, __fmul_rn(__int2float_rn(acc), s2));
        }
        out[row] = v;
      }
    }
  }
}

// Distances from the query to rows ids[0..cnt) (ids in LDS), written to
// out[0..cnt) (LDS), with traversal precision `prec` (0 exact f32, 1 the
// crate's int8 codes, 2 bf16 rows, 3 per-dimension int8 codes).  All 64
// lanes must call it.
template <int TMAX, bool EXTRA, bool EXACT>
__device__ __forceinline__ void gather_dists(const SearchArgs &P, const QLanes<TMAX, EXTRA> &q,
                                             const uint2 (&qc)[code_chunks<TMAX, EXACT>()],
                                             const QCodeInfo &qs, int prec, const uint32_t *ids,
                                             int cnt, float *out) {
  if (prec == 1 || prec == 3) {
    gather_codes<code_chunks<TMAX, EXACT>()>(P, qc, qs, prec, ids, cnt, out, EXACT);
    return;
  }
  if (prec == 2) gather_rows<TMAX, EXTRA, EXACT, uint16_t>(P, q, P.bf16, ids, cnt, out);
  else gather_rows<TMAX, EXTRA, EXACT, float>(P, q, P.rows, ids, cnt, out);
}

__device__ __forceinline__ bool admit(const SearchArgs &P, const uint32_t *allow, uint32_t id) {
  if (P.deleted_bits && bit_test(P.deleted_bits, id)) return false;
  if (allow && !bit_test(allow, id)) return false;
  return true;
}

// Phase-1 output of one query (all lanes store the same word).
__device__ __forceinline__ void seed_query(const SearchArgs &P, uint32_t qi, uint32_t entry,
                                           float best) {
  GM(P.seed)[qi] = ((uint64_t)__float_as_uint(best) << 32) | entry;
}

// EXACT: D == 32 * TMAX (branch-free row loads, see load_q).
template <int TMAX, bool EXTRA, bool EXACT, int SW>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TMAX <= 8 ? 4 : 2)))
hnsw_search_kernel(const SearchArgs P) {
  extern __shared__ uint32_t lds[];
  const int lane = lane_id(), i16 = lane & 15;
  const int wave = threadIdx.x >> 6;
  const uint32_t slot = blockIdx.x * (blockDim.x >> 6) + wave;
  uint32_t *table = lds + (size_t)wave * P.lds_words;
  uint32_t *st_id = table + P.hash_cap;
  float *st_d = reinterpret_cast<float *>(st_id + 64);
  uint64_t *Ck = reinterpret_cast<uint64_t *>(st_d + 64);  // 128 * SW candidate words
  uint64_t *Wk = Ck + 128 * SW;                              // 64 * SW result words
  uint32_t *spill = P.spill ? P.spill + (size_t)slot * P.spill_words : nullptr;
  gmut<uint32_t> slog = GM(P.spill_log) + (size_t)slot * P.spill_log_cap;
  const uint32_t hcap = P.hash_cap;
  const Shape shp{P.T, P.U, P.R, P.stride};
  gptr<uint32_t> adj0 = G(P.adj0);
  Stats st = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  // Candidates beyond the register list go to a per-wave pool in HBM, so
  // the candidate set is as unbounded as the crate's heap up to the pool's
  // capacity.  Invariant: every pooled entry orders after every entry of C,
  // so the front of C is always the global minimum.
  gmut<uint64_t> pool = GM(P.cpool) + (size_t)slot * P.cpool_cap;
  const bool has_pool = P.cpool != nullptr;

  // Phase 1 (descent only, nearly uniform work per query) takes queries
  // statically, without the shared counter.
  uint32_t static_next = slot;
  const uint32_t total_waves = gridDim.x * (blockDim.x >> 6);
  for (;;) {
    // Side effects in this loop run on every lane with wave-uniform values
    // (never under `lane == 0` branches): correlated lane-0 branches around
    // the back edge let the compiler run lanes a loop iteration apart, which
    // breaks the readlane broadcasts below.  Only lane 0 adds 1 here.
    uint32_t qi;
    if (P.phase == 1) {
      qi = static_next;
      static_next += total_waves;
    } else if (P.phase == 3) {
      qi = rl(atomicAdd(P.counter + 1, lane == 0 ? 1u : 0u), 0);
    } else {
      qi = rl(atomicAdd(P.counter, lane == 0 ? 1u : 0u), 0);
    }
    if (P.phase == 3) {
      const uint32_t nredo = rl(G(P.redo_count)[0], 0);
      if (qi >= (nredo < P.nq ? nredo : (uint32_t)P.nq)) break;
      qi = rl(G(P.redo)[qi], 0);
    } else if (qi >= P.nq) {
      break;
    }
    if (P.phase == 2) qi = rl(G(P.order)[qi], 0);
    uint32_t k = P.k, ef = P.ef, prec = P.precision, filt = 0xFFFFFFFFu;
    if (P.specs) {
      const hh_query_spec sp = G(P.specs)[qi];
      k = sp.k;
      ef = sp.ef;
      prec = sp.precision;
      filt = sp.filter;
    }
    k = rl(k, 0); ef = rl(ef, 0); prec = rl(prec, 0); filt = rl(filt, 0);
    const uint64_t obase = (uint64_t)qi * P.k_stride;
    const int status = (int)rl((uint32_t)G(P.status)[qi], 0);
    if (status != 0 || k == 0 || P.live == 0 || P.top < 0) {
      GM(P.out_count)[qi] = 0;
      if (P.phase == 1) seed_query(P, qi, kInvalid, -INFINITY);
      continue;
    }
    if (ef < k) ef = k;
    // Candidate-list capacity depends on the query's own ef only (twice
    // its 64-multiple bucket), never on the kernel instance, so a query
    // returns the same results alone, in a mixed batch or when served.
    const uint32_t ccap = ef <= 64 ? 128 : ef <= 128 ? 256 : ef <= 256 ? 512 : 1024;
    const uint32_t *allow = P.allow;
    if (P.specs) allow = filt == 0xFFFFFFFFu ? nullptr : P.filters + (size_t)filt * P.filter_words;
    if (P.phase == 0 || P.phase == 2) st.q++;

    // query, in the row layout, plus its int8 code words
    QLanes<TMAX, EXTRA> q;
#pragma unroll
    for (int t = 0; t < TMAX; t++) q.lo[t] = q.hi[t] = 0.0f;
    q.e[0] = q.r = 0.0f;
    if constexpr (EXTRA) q.e[1] = q.e[2] = 0.0f;
    load_q<TMAX, EXTRA, float>(P.qprep + (size_t)qi * P.stride, shp, i16, q);
    uint2 qc[code_chunks<TMAX, EXACT>()];
#pragma unroll
    for (int c = 0; c < code_chunks<TMAX, EXACT>(); c++) qc[c] = make_uint2(0, 0);
    QCodeInfo qs{0, 1.0f, 0.0f};
    if (prec == 1 || prec == 3) {
      gptr<uint2> cq = (gptr<uint2>)(P.qcode + (size_t)qi * P.cstride);
      const int chunks = (int)(P.cstride >> 7);
#pragma unroll
      for (int c = 0; c < code_chunks<TMAX, EXACT>(); c++)
        if (c < chunks) qc[c] = cq[c * 16 + i16];
      if (prec == 1 && P.metric == 0) {
        int aa = 0;
#pragma unroll
        for (int c = 0; c < code_chunks<TMAX, EXACT>(); c++) {
          aa = __builtin_amdgcn_sdot4((int)qc[c].x, (int)qc[c].x, aa, false);
          aa = __builtin_amdgcn_sdot4((int)qc[c].y, (int)qc[c].y, aa, false);
        }
        qs.qaa = quarter_sum_i(aa);
      } else if (prec == 3) {
        const float2 v = G(P.qpd)[qi];
        qs.sigma = __builtin_bit_cast(float, rl(__float_as_uint(v.x), 0));
        qs.qconst = __builtin_bit_cast(float, rl(__float_as_uint(v.y), 0));
      }
    }

    int mode = 0;
    uint32_t vcount = 0;
    uint32_t slog_n = 0;  // nodes logged since the visited set left LDS

    // ---- entry point and greedy descent (search_graph, search.rs:146-161)
    uint32_t cur;
    float best;
    if (P.phase >= 2) {
      const uint64_t sd = G(P.seed)[qi];
      cur = rl((uint32_t)sd, 0);
      best = __uint_as_float(rl((uint32_t)(sd >> 32), 0));
    } else {
      cur = P.entry;
      // The upper-layer block offset of the current node is always known
      // before its block is needed: the entry's is loaded with the entry's
      // row, a neighbour's with the neighbours' rows.  So each greedy step
      // costs two dependent round trips (block, rows).
      uint64_t cur_off = P.top >= 1 ? (uint64_t)G(P.upper_off)[cur] : 0;
      st_id[0] = cur;
      wave_sync();
      gather_dists<TMAX, EXTRA, EXACT>(P, q, qc, qs, (int)prec, st_id, 1, st_d);
      wave_sync();
      best = rlf(st_d[0], 0);
      st.up++;
      for (int layer = P.top; layer >= 1; layer--) {
        for (;;) {
          const uint32_t from = cur;
          gptr<uint32_t> blk = G(P.upper) + cur_off + (size_t)(layer - 1) * (P.m + 1);
          // the count and the first 64 slots in one round trip (slots past
          // the count are inside the block, m+1 words)
          const uint32_t cnt_word = blk[0];
          const uint32_t slot0 = lane < (int)P.m ? blk[1 + lane] : kInvalid;
          uint32_t cnt = rl(cnt_word, 0);
          if (cnt > P.m) cnt = P.m;
          for (uint32_t base = 0; base < cnt; base += 64) {
            const int c = (int)(cnt - base < 64 ? cnt - base : 64);
            const uint32_t nid = lane < c ? (base == 0 ? slot0 : blk[1 + base + lane]) : kInvalid;
            const uint64_t noff = lane < c ? (uint64_t)G(P.upper_off)[nid] : 0;
            if (lane < c) st_id[lane] = nid;
            wave_sync();
            gather_dists<TMAX, EXTRA, EXACT>(P, q, qc, qs, (int)prec, st_id, c, st_d);
            wave_sync();
            const float dv = lane < c ? st_d[lane] : 0.0f;
            // first list position holding the IEEE-smallest distance
            float md = dv;
            int mi = lane < c ? lane : 64;
#pragma unroll
            for (int m = 1; m < 64; m <<= 1) {
              const float od = __shfl_xor(md, m, 64);
              const int oi = __shfl_xor(mi, m, 64);
              const bool take = oi < 64 && (mi == 64 || od < md || (!(md < od) && oi < mi));
              if (take) {
                md = od;
                mi = oi;
              }
            }
            md = rlf(md, 0);
            mi = (int)rl((uint32_t)mi, 0);
            st.up += c;
            if (md < best) {
              best = md;
              cur = rl(nid, mi);
              cur_off = ((uint64_t)rl((uint32_t)(noff >> 32), mi) << 32) | rl((uint32_t)noff, mi);
            }
            wave_sync();
          }
          if (cur == from) break;
        }
      }
      if (P.phase == 1) {
        seed_query(P, qi, cur, best);
        continue;
      }
    }

    // ---- layer 0 (search_layer, search.rs:81-142)
    // counters as they stand before layer 0: restored if this attempt is
    // dropped for a re-run
    const Stats st0 = st;
    const uint32_t ep_id = cur;
    const float ep_d = best;
    for (uint32_t t = lane; t < hcap; t += 64) table[t] = kInvalid;
    wave_sync();
    // C and W: sorted candidate words in LDS (C from `head`, W from 0).
    uint32_t head = 0, csize = 0, wsize = 0;
    float worst = 0.0f;
    uint32_t pn = 0;            // pooled candidates
    uint64_t pmin = ~0ull;      // lower bound of the pooled candidates
    lds_visit(table, hcap, cur);
    wave_sync();
    vcount = 1;
    const uint64_t ep_key = ckey(best, cur);
    Ck[0] = ep_key;
    csize = 1;
    if (admit(P, allow, cur)) {
      Wk[0] = ep_key;
      wsize = 1;
      worst = best;
    }
    wave_sync();
    const bool pf_ok = P.m0 <= 64;
    uint32_t pf_id = kInvalid, pf_adj = kInvalid, pf2_id = kInvalid, pf2_adj = kInvalid;

    for (;;) {
      if (csize == 0) {
        if (pn == 0) break;
        // refill C from the pool in chunks of 64, each merged like a
        // neighbour batch; what does not fit goes back to the pool, which
        // is compacted in place (a chunk writes at most as many entries as
        // it read)
        const uint32_t total = pn;
        const bool wfull = wsize >= ef;
        pn = 0;
        pmin = ~0ull;
        head = 0;
        for (uint32_t base = 0; base < total; base += 64) {
          const uint32_t c = total - base < 64 ? total - base : 64;
          const uint64_t x = lane < (int)c ? (uint64_t)pool[base + lane] : ~0ull;
          const uint64_t live = __ballot(lane < (int)c && !(wfull && ckey_d(x) > worst));
          uint64_t ce[2 * SW];
          list_load<2 * SW>(Ck + head, csize, ce);
          c_merge<2 * SW>(Ck, ce, head, csize, ccap, live, x, wfull, worst, pool, has_pool,
                          P.cpool_cap, pn, pmin, st.ovf, st.cspill);
        }
        if (csize == 0) break;
      }
      const uint64_t top = Ck[head];
      const float cd = ckey_d(top);
      const uint32_t cid = (uint32_t)top;
      head++;
      csize--;
      if (wsize >= ef && cd > worst) break;
      st.exp++;
      const size_t arow = (size_t)cid * P.m0;
      for (uint32_t base = 0; base < P.m0; base += 64) {
        const int c = (int)(P.m0 - base < 64 ? P.m0 - base : 64);
        const bool last = base + 64 >= P.m0;
        uint32_t nid;
        if (base == 0 && cid == pf_id) nid = pf_adj;
        else if (base == 0 && cid == pf2_id) nid = pf2_adj;
        else nid = lane < c ? adj0[arow + base + lane] : kInvalid;
        bool isnew = false;
        if (nid != kInvalid) isnew = mode == 0 ? lds_visit(table, hcap, nid) : bit_visit(spill, nid);
        const uint64_t nm = __ballot(isnew);
        const int nn = popc64(nm);
        if (mode == 1 && nn > 0) {
          // log nodes marked in the bitset so the query's end clears just
          // their words (a full clear past the log's capacity)
          if (slog_n + (uint32_t)nn <= P.spill_log_cap && isnew) slog[slog_n + mbcnt(nm)] = nid;
          slog_n += (uint32_t)nn;
        }
        // The adjacency row of the next candidate is fetched while this
        // expansion's rows are gathered and merged: first C's front (the
        // next pop unless a new neighbour beats it), below the best new one.
        if (last) {
          pf_id = kInvalid;
          pf2_id = kInvalid;
          if (pf_ok && csize > 0) {
            pf_id = (uint32_t)Ck[head];
            pf_adj = lane < (int)P.m0 ? adj0[(size_t)pf_id * P.m0 + lane] : kInvalid;
          }
        }
        if (nn == 0) continue;
        if (isnew) st_id[mbcnt(nm)] = nid;
        wave_sync();
        const uint32_t myid = lane < nn ? st_id[lane] : kInvalid;
        const bool adm = lane < nn && admit(P, allow, myid);
        gather_dists<TMAX, EXTRA, EXACT>(P, q, qc, qs, (int)prec, st_id, nn, st_d);
        wave_sync();
        const float myd = lane < nn ? st_d[lane] : 0.0f;
        const uint64_t am = __ballot(adm);
        vcount += nn;
        st.dist += nn;
        // Which new neighbours the crate's sequential loop accepts
        // (search.rs:118-132): neighbour t passes iff |W_t| < ef or
        // d_t < worst_t, and worst_t is the ef-th smallest distance of W and
        // the admitted neighbours before t.  So t passes iff fewer than ef
        // of those are at or below d_t.  (Neighbours at or above the worst
        // at the batch start can never pass.)
        const uint64_t sm = __ballot(lane < nn && (wsize < ef || myd < worst));
        const bool surv = (sm >> lane) & 1;
        if (last && pf_ok && sm) {
          // the smallest surviving neighbour (accepted below unless ties
          // fill W first), if it beats C's front, is the next candidate:
          // fetch its adjacency row during the acceptance and the merges
          // (a hint only: the smallest distance by row DPP and the four
          // row minima, its lowest lane on ties)
          float dm = surv ? myd : INFINITY;
          dm = fminf(dm, dpp<kRor8>(dm));
          dm = fminf(dm, dpp<kRor4>(dm));
          dm = fminf(dm, dpp<kXor1>(dm));
          dm = fminf(dm, dpp<kXor2>(dm));
          dm = fminf(fminf(rlf(dm, 0), rlf(dm, 16)), fminf(rlf(dm, 32), rlf(dm, 48)));
          const uint64_t at = __ballot(surv && myd == dm);
          const int jl = __ffsll((unsigned long long)at) - 1;
          const uint64_t key = at ? ckey(dm, rl(myid, jl)) : ~0ull;
          if (at && (csize == 0 || key < Ck[head])) {
            pf2_id = (uint32_t)key;
            pf2_adj = lane < (int)P.m0 ? adj0[(size_t)pf2_id * P.m0 + lane] : kInvalid;
          }
        }
        uint64_t we[SW];
        list_load<SW>(Wk, wsize, we);
        // Shortcut: every survivor passes when even the largest survivor
        // distance has fewer than ef entries at or below it counting W's and
        // every admitted survivor (then no neighbour's count can reach ef).
        const uint32_t nadm = (uint32_t)popc64(sm & am);
        bool all_pass = wsize + nadm < ef;
        if (!all_pass && sm) {
          float dx = surv ? myd : -INFINITY;
          dx = fmaxf(dx, dpp<kRor8>(dx));
          dx = fmaxf(dx, dpp<kRor4>(dx));
          dx = fmaxf(dx, dpp<kXor1>(dx));
          dx = fmaxf(dx, dpp<kXor2>(dx));
          dx = fmaxf(fmaxf(rlf(dx, 0), rlf(dx, 16)), fmaxf(rlf(dx, 32), rlf(dx, 48)));
          uint32_t cw = 0;
#pragma unroll
          for (int c = 0; c < SW; c++)
            cw += (uint32_t)popc64(__ballot(c * 64 + lane < (int)wsize && !(ckey_d(we[c]) > dx)));
          all_pass = cw + nadm < ef;
        }
        uint32_t below = 0;
        if (!all_pass) {
          if (wsize <= 64) below = accept_counts<SW, 1>(we, wsize, sm, am, myd);
          else if (SW > 2 && wsize <= 128) below = accept_counts<SW, (SW > 2 ? 2 : SW)>(we, wsize, sm, am, myd);
          else if (SW > 4 && wsize <= 256) below = accept_counts<SW, (SW > 4 ? 4 : SW)>(we, wsize, sm, am, myd);
          else below = accept_counts<SW, SW>(we, wsize, sm, am, myd);
        }
        const bool acc = surv && (all_pass || below < ef);
        const uint64_t amask = __ballot(acc);
        const uint64_t akey = acc ? ckey(myd, myid) : ~0ull;
        if (amask) {
          uint64_t ce[2 * SW];
          list_load<2 * SW>(Ck + head, csize, ce);
          if (amask & am) worst = ckey_d(w_merge<SW>(Wk, we, wsize, ef, amask & am, akey));
          c_merge<2 * SW>(Ck, ce, head, csize, ccap, amask, akey, wsize >= ef, worst, pool,
                          has_pool, P.cpool_cap, pn, pmin, st.ovf, st.cspill);
        }
        if (mode == 0 && vcount > P.hash_limit && spill) {
          for (uint32_t t = lane; t < hcap; t += 64) {
            const uint32_t v = table[t];
            if (v != kInvalid)
              __hip_atomic_fetch_or(spill + (v >> 5), 1u << (v & 31), __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT);
          }
          __threadfence();
          mode = 1;
          st.spill++;
        }
      }
    }

    if (mode == 1) {
      // clear this query's bits: the LDS table's nodes and the logged ones
      __threadfence();
      for (uint32_t t = lane; t < hcap; t += 64) {
        const uint32_t v = table[t];
        if (v != kInvalid) GM(spill)[v >> 5] = 0u;
      }
      if (slog_n <= P.spill_log_cap) {
        for (uint32_t j = lane; j < slog_n; j += 64) GM(spill)[slog[j] >> 5] = 0u;
      } else {
        for (uint32_t w = lane; w < P.spill_words; w += 64) GM(spill)[w] = 0u;
      }
      __threadfence();
    }

    if (P.redo && P.phase != 3 && st.ovf != st0.ovf) {
      // the pool dropped a candidate: list the query for the phase-3 launch
      // (which starts from the same layer-0 entry) and drop this attempt
      st = st0;
      st.redo++;
      seed_query(P, qi, ep_id, ep_d);
      const uint32_t at = rl(atomicAdd(P.redo_count, lane == 0 ? 1u : 0u), 0);
      if (at < P.nq) GM(P.redo)[at] = qi;
      wave_sync();
      continue;
    }

    // ---- results: optional exact re-rank, keys, (distance, key) order
    float *res_d = reinterpret_cast<float *>(table);
    uint32_t *res_id = table + 64 * SW;
    uint32_t *res_key = table + 128 * SW;  // 2 words per entry
    for (uint32_t idx = lane; idx < wsize; idx += 64) {
      const uint64_t w = Wk[idx];
      res_d[idx] = ckey_d(w);
      res_id[idx] = (uint32_t)w;
    }
    wave_sync();
    if (prec != 0) {
      for (int base = 0; base < wsize; base += 64) {
        const int c = wsize - base < 64 ? wsize - base : 64;
        gather_dists<TMAX, EXTRA, EXACT>(P, q, qc, qs, 0, res_id + base, c, res_d + base);
        wave_sync();
      }
      st.rr += wsize;
    }
    for (int idx = lane; idx < wsize; idx += 64) {
      const uint64_t key = P.ids_out ? (uint64_t)res_id[idx] : G(P.keys)[res_id[idx]];
      res_key[2 * idx] = (uint32_t)key;
      res_key[2 * idx + 1] = (uint32_t)(key >> 32);
    }
    wave_sync();
    // W is sorted by (distance, id); the crate orders by (distance, key).
    // Without equal distances the orders agree; otherwise rank by counting.
    bool ties = false;
    for (int idx = lane; idx + 1 < wsize; idx += 64) ties |= tkey(res_d[idx]) == tkey(res_d[idx + 1]);
    const bool direct = prec == 0 && __ballot(ties) == 0;
    for (int idx = lane; idx < wsize; idx += 64) {
      const float d = res_d[idx];
      const uint64_t key = ((uint64_t)res_key[2 * idx + 1] << 32) | res_key[2 * idx];
      int rank = idx;
      if (!direct) {
        rank = 0;
        for (int o = 0; o < wsize; o++) {
          const uint64_t ko = ((uint64_t)res_key[2 * o + 1] << 32) | res_key[2 * o];
          rank += nless(res_d[o], ko, d, key);
        }
      }
      if (rank < (int)k) {
        GM(P.out_keys)[obase + rank] = key;
        GM(P.out_dist)[obase + rank] = d;
      }
    }
    GM(P.out_count)[qi] = (uint32_t)(wsize < (int)k ? wsize : (int)k);
    wave_sync();
  }
  if (lane == 0 && P.stats) {
    atomicAdd(&P.stats[0], st.q);
    atomicAdd(&P.stats[1], st.dist);
    atomicAdd(&P.stats[2], st.up);
    atomicAdd(&P.stats[3], st.exp);
    atomicAdd(&P.stats[4], st.rr);
    atomicAdd(&P.stats[5], st.ovf);
    atomicAdd(&P.stats[6], st.spill);
    atomicAdd(&P.stats[7], st.cspill);
    atomicAdd(&P.stats[8], st.redo);
  }
}

// ------------------------------------------------------ descent pass
// The first pass of a two-pass launch when m <= 32: one query per 16 lanes
// (m <= 16, four per wavefront) or 32 lanes (two per wavefront), one
// neighbour row per lane.  A lane computes a whole distance with the AVX2
// kernel's 32 accumulators (value 32t + j into accumulator j, one fused
// multiply-add per element in t order, the 8-wide remainder into
// accumulators 0..7, x86.rs:47-77 / 96-129) and its lane combination
// s[l] = (a_l + a_l+8) + (a_l+16 + a_l+24),
// h = ((s0+s4)+(s1+s5))+((s2+s6)+(s3+s7)) (x86.rs:21-27), then the tail
// products, so every distance has the bits of the search kernel's quarter
// reduction.  The greedy steps are those of search_graph / greedy_closest
// (src/hnsw/search.rs:49-68,146-161): the first list position holding the
// smallest distance, taken when strictly below the best.

// Accumulator j's four fused multiply-adds of one full 4-block (values
// 32t + j, t in 4b..4b+3, at j*4 + t%4), in t order.
template <bool L2, class E>
__device__ __forceinline__ float acc4(float acc, const float *q, gptr<E> blk, int j) {
  float4 x;
  if constexpr (sizeof(E) == 4) {
    x = *(gptr<float4>)(blk + 4 * j);
  } else {
    const uint2 h = *(gptr<uint2>)(blk + 4 * j);
    x = make_float4(__uint_as_float(h.x << 16), __uint_as_float(h.x & 0xFFFF0000u),
                    __uint_as_float(h.y << 16), __uint_as_float(h.y & 0xFFFF0000u));
  }
  const float4 y = *reinterpret_cast<const float4 *>(q + 4 * j);
  if (L2) {
    float d = y.x - x.x;
    acc = __builtin_fmaf(d, d, acc);
    d = y.y - x.y;
    acc = __builtin_fmaf(d, d, acc);
    d = y.z - x.z;
    acc = __builtin_fmaf(d, d, acc);
    d = y.w - x.w;
    return __builtin_fmaf(d, d, acc);
  }
  acc = __builtin_fmaf(y.x, x.x, acc);
  acc = __builtin_fmaf(y.y, x.y, acc);
  acc = __builtin_fmaf(y.z, x.z, acc);
  return __builtin_fmaf(y.w, x.w, acc);
}

// The crate's combination of the 32 accumulators: s[l] = (a_l + a_{l+8}) +
// (a_{l+16} + a_{l+24}), h = ((s0+s4)+(s1+s5))+((s2+s6)+(s3+s7))
// (src/kernels/x86.rs:21-27, 70-73).
#define HH_COMBINE32(acc, out)                                                                 \
  do {                                                                                         \
    float sl_[8];                                                                              \
    _Pragma("unroll") for (int l_ = 0; l_ < 8; l_++) sl_[l_] =                                 \
        __fadd_rn(__fadd_rn(acc[l_], acc[l_ + 8]), __fadd_rn(acc[l_ + 16], acc[l_ + 24]));    \
    out = __fadd_rn(__fadd_rn(__fadd_rn(sl_[0], sl_[4]), __fadd_rn(sl_[1], sl_[5])),           \
                    __fadd_rn(__fadd_rn(sl_[2], sl_[6]), __fadd_rn(sl_[3], sl_[7])));          \
  } while (0)

// One lane's whole distance sum of an exact-layout row (D = 32 * TMAX) in
// the crate's order, the same bits as the quarter kernel's.
template <int TMAX, bool L2, class E>
__device__ __forceinline__ float lane_distance(const float *q, gptr<E> row) {
  float acc[32];
#pragma unroll
  for (int j = 0; j < 32; j++) acc[j] = 0.0f;
#pragma unroll
  for (int b = 0; b < TMAX / 4; b++) {
#pragma unroll
    for (int j = 0; j < 32; j++) acc[j] = acc4<L2, E>(acc[j], q + b * 128, row + b * 128, j);
  }
  float h;
  HH_COMBINE32(acc, h);
  return h;
}

// The last, partial block of TW < 4 values per accumulator (value 32t + j
// at j*TW + t%4).
template <int TW, bool L2, class E>
__device__ __forceinline__ void acc_partial(float (&acc)[32], const float *q, gptr<E> blk) {
#pragma unroll
  for (int j = 0; j < 32; j++) {
#pragma unroll
    for (int tt = 0; tt < TW; tt++) {
      const float x = to_f<E>(blk[j * TW + tt]), y = q[j * TW + tt];
      if (L2) {
        const float d = y - x;
        acc[j] = __builtin_fmaf(d, d, acc[j]);
      } else {
        acc[j] = __builtin_fmaf(y, x, acc[j]);
      }
    }
  }
}

// Any layout: the full 4-blocks, the last partial block (value 32t + j at
// j*tw + t%4, tw = T mod 4), the 8-wide remainder into accumulators 0..7,
// then the unfused tail products added after the combination, in order
// (x86.rs:70-77, 121-129) -- the quarter kernel's order, element by element.
template <bool L2, class E>
__device__ __forceinline__ float lane_distance_any(const float *q, gptr<E> row, const Shape &s) {
  float acc[32];
#pragma unroll
  for (int j = 0; j < 32; j++) acc[j] = 0.0f;
  const int full = s.T >> 2;
  for (int b = 0; b < full; b++) {
#pragma unroll
    for (int j = 0; j < 32; j++) acc[j] = acc4<L2, E>(acc[j], q + b * 128, row + b * 128, j);
  }
  const int tw = s.T & 3;
  if (tw == 1) acc_partial<1, L2, E>(acc, q + full * 128, row + full * 128);
  if (tw == 2) acc_partial<2, L2, E>(acc, q + full * 128, row + full * 128);
  if (tw == 3) acc_partial<3, L2, E>(acc, q + full * 128, row + full * 128);
  const int rest = 32 * s.T;
  for (int u = 0; u < s.U; u++) {
#pragma unroll
    for (int l = 0; l < 8; l++) {
      const float x = to_f<E>(row[rest + 8 * u + l]), y = q[rest + 8 * u + l];
      if (L2) {
        const float d = y - x;
        acc[l] = __builtin_fmaf(d, d, acc[l]);
      } else {
        acc[l] = __builtin_fmaf(y, x, acc[l]);
      }
    }
  }
  float h;
  HH_COMBINE32(acc, h);
  for (int r = 0; r < s.R; r++) {
    const int pos = rest + 8 * s.U + r;
    const float x = to_f<E>(row[pos]), y = q[pos];
    if (L2) {
      const float d = y - x;
      h = __fadd_rn(h, __fmul_rn(d, d));
    } else {
      h = __fadd_rn(h, __fmul_rn(y, x));
    }
  }
  return h;
}

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

// The traversal distance of the batch's precision (0 f32, 1 the crate's
// int8 codes, 2 bf16 rows, 3 per-dimension codes), with the bits the search
// kernel computes.
template <int TMAX, bool EXACT>
__device__ __forceinline__ float descent_distance(const SearchArgs &P, uint32_t prec,
                                                  const Shape &shp, const float *q,
                                                  const uint2 *qc, const QCodeInfo &qs,
                                                  uint32_t id) {
  if (prec == 1 || prec == 3) return lane_distance_i8<EXACT ? 4 * TMAX : 0>(P, prec, qs, qc, id);
  float s;
  if (prec == 2) {
    gptr<uint16_t> row = G(P.bf16) + (size_t)id * P.stride;
    if constexpr (EXACT)
      s = P.metric == 0 ? lane_distance<TMAX, true, uint16_t>(q, row)
                        : lane_distance<TMAX, false, uint16_t>(q, row);
    else
      s = P.metric == 0 ? lane_distance_any<true, uint16_t>(q, row, shp)
                        : lane_distance_any<false, uint16_t>(q, row, shp);
  } else {
    gptr<float> row = G(P.rows) + (size_t)id * P.stride;
    if constexpr (EXACT)
      s = P.metric == 0 ? lane_distance<TMAX, true, float>(q, row)
                        : lane_distance<TMAX, false, float>(q, row);
    else
      s = P.metric == 0 ? lane_distance_any<true, float>(q, row, shp)
                        : lane_distance_any<false, float>(q, row, shp);
  }
  return finish_metric(s, P.metric);
}

template <int TMAX, int LPQ, bool EXACT>
__global__ void __launch_bounds__(descent_threads<TMAX>()) descent_kernel(const SearchArgs P) {
  static_assert(!EXACT || TMAX % 4 == 0, "exact layout");
  constexpr int kStride = descent_stride<TMAX>();
  constexpr int kCodeWords = descent_code_words<TMAX>();
  constexpr int kQueries = descent_threads<TMAX>() / LPQ;
  __shared__ __align__(16) float qs[kQueries][kStride];
  __shared__ __align__(16) uint2 qcs[kQueries][kCodeWords];
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
      // the query (prepared, row layout) and its codes into this slot
      gptr<float4> src = (gptr<float4>)(P.qprep + qi * P.stride);
      float4 *dst = reinterpret_cast<float4 *>(qs[qslot]);
      for (int c = il; c < (int)(P.stride / 4); c += LPQ) dst[c] = src[c];
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
      const float *q = qs[qslot];
      const uint2 *qc = qcs[qslot];
      uint32_t cur = P.entry;
      uint64_t cur_off = G(P.upper_off)[cur];
      float best = descent_distance<TMAX, EXACT>(P, prec, shp, q, qc, qk, cur);
      up++;
      for (int layer = P.top; layer >= 1; layer--) {
        for (;;) {
          gptr<uint32_t> blk = G(P.upper) + cur_off + (size_t)(layer - 1) * (P.m + 1);
          uint32_t cnt = blk[0];
          if (cnt > P.m) cnt = P.m;
          const bool valid = (uint32_t)il < cnt;
          const uint32_t nid = valid ? blk[1 + il] : kInvalid;
          const uint64_t noff = valid ? (uint64_t)G(P.upper_off)[nid] : 0;
          float md = valid ? descent_distance<TMAX, EXACT>(P, prec, shp, q, qc, qk, nid) : 0.0f;
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
  // one count per query: its first lane adds the query's total
  if (il == 0 && up && P.stats) atomicAdd(&P.stats[2], up);
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

#define HH_SEARCH_INST(TM, EX, XA)                                                  \
  template __global__ void hnsw_search_kernel<TM, EX, XA, 1>(const SearchArgs);     \
  template __global__ void hnsw_search_kernel<TM, EX, XA, 2>(const SearchArgs);     \
  template __global__ void hnsw_search_kernel<TM, EX, XA, 4>(const SearchArgs);     \
  template __global__ void hnsw_search_kernel<TM, EX, XA, 8>(const SearchArgs);
HH_SEARCH_INST(4, false, false)
HH_SEARCH_INST(4, true, false)
HH_SEARCH_INST(4, false, true)
HH_SEARCH_INST(8, false, false)
HH_SEARCH_INST(8, true, false)
HH_SEARCH_INST(8, false, true)
HH_SEARCH_INST(32, false, false)
HH_SEARCH_INST(32, true, false)
HH_SEARCH_INST(32, false, true)

}  // namespace hhk
