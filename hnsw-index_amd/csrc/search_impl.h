// search_impl.h — the batched HNSW search kernel on gfx950, instantiated
// per precision class by search_bf16.hip, search_codes.hip and
// search_any.hip.
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
#pragma once

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
// (The write phase, given every element's merged position: ep for W's,
// ap for the batch's.)
template <int WPL>
__device__ __forceinline__ uint64_t w_write(uint64_t *W, const uint64_t (&e)[WPL],
                                            const uint32_t (&ep)[WPL], uint32_t ap,
                                            uint32_t &size, uint32_t cap, uint64_t amask,
                                            uint64_t akey) {
  const int lane = lane_id();
  const bool mine = (amask >> lane) & 1;
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
template <int WPL>
__device__ __forceinline__ uint64_t w_merge(uint64_t *W, const uint64_t (&e)[WPL], uint32_t &size,
                                            uint32_t cap, uint64_t amask, uint64_t akey) {
  uint32_t ep[WPL];
  const uint32_t ap = merge_ranks<WPL>(e, ep, size, amask, akey);
  return w_write<WPL>(W, e, ep, ap, size, cap, amask, akey);
}

// Beam steps expand up to kBeamMax lists, up to 256 neighbour slots: slot
// s = lane + 64 g (lane group g < kBeamGroups) holds list s / m0, position
// s % m0.
constexpr int kBeamMax = 8, kBeamGroups = 4;

// One of eight wave-uniform values, chosen per lane by j (explicit selects:
// a lane-indexed private array would live in scratch memory).
__device__ __forceinline__ uint32_t sel8(uint32_t j, const uint32_t (&x)[kBeamMax]) {
  uint32_t r = kInvalid;
#pragma unroll
  for (int t = 0; t < kBeamMax; t++) r = j == (uint32_t)t ? x[t] : r;
  return r;
}

// Adjacency prefetch of a beam step: the candidates at the front after a
// merge are the next step's pops (unless the pool refills C first), so their
// layer-0 lists are loaded as soon as the merge ranks are known, in the
// step's slot layout.  A hint only: the next step checks each list's id.
struct FrontFetch {
  const uint32_t *adj0;
  uint32_t m0, n, groups;  // lists per step, lane groups they take
  uint32_t sj[kBeamGroups], sp[kBeamGroups];  // this lane's list and position, per group
  uint32_t id[kBeamMax];
  uint32_t v[kBeamGroups];
};

__device__ __forceinline__ void front_fetch(FrontFetch &f) {
  gptr<uint32_t> adj = G(f.adj0);
#pragma unroll
  for (int g = 0; g < kBeamGroups; g++) {
    f.v[g] = kInvalid;
    if ((uint32_t)g < f.groups) {
      const uint32_t a = sel8(f.sj[g], f.id);
      if (f.sj[g] < f.n && a != kInvalid) f.v[g] = adj[(size_t)a * f.m0 + f.sp[g]];
    }
  }
}

// The merged order's first f.n entries, given C's first chunk (e0, its
// merged positions ep0, size entries) and the batch (amask, words akey at
// positions ap), and their lists fetched: list entries at positions below
// kBeamMax come from the first chunk, since positions never fall below
// indices.
__device__ __forceinline__ void front_capture(FrontFetch &f, uint64_t e0, uint32_t ep0, uint32_t ap,
                                              uint64_t amask, uint64_t akey, uint32_t size) {
  const int lane = lane_id();
  const bool mine = (amask >> lane) & 1;
#pragma unroll
  for (int p = 0; p < kBeamMax; p++) {
    f.id[p] = kInvalid;
    if (p < (int)f.n) {
      const uint64_t la = __ballot(mine && ap == (uint32_t)p);
      const uint64_t lb = __ballot(lane < (int)size && ep0 == (uint32_t)p);
      if (la) f.id[p] = rl((uint32_t)akey, __ffsll((unsigned long long)la) - 1);
      else if (lb) f.id[p] = rl((uint32_t)e0, __ffsll((unsigned long long)lb) - 1);
    }
  }
  front_fetch(f);
}

// One beam step's visit: the neighbours of the popped lists x[0..ne) in slot
// layout (each slot's id from the prefetch where its list matches, else
// loaded), probed in the visited set (LDS table or, in mode 1, the HBM
// bitset with the spill log), and the new ones staged in st_id[0..nn) in
// slot order.  Returns nn.  All 64 lanes must call it.
__device__ __forceinline__ uint32_t beam_visit(gptr<uint32_t> adj0, const uint32_t (&x)[kBeamMax],
                                               uint32_t ne, const FrontFetch &ff, uint32_t *table,
                                               uint32_t hcap, int mode, uint32_t *spill,
                                               gmut<uint32_t> slog, uint32_t slog_cap,
                                               uint32_t &slog_n, uint32_t *st_id) {
  const uint32_t groups = (ne * ff.m0 + 63) / 64;
  uint32_t nid[kBeamGroups];
  bool nw[kBeamGroups];
#pragma unroll
  for (int g = 0; g < kBeamGroups; g++) {
    nid[g] = kInvalid;
    nw[g] = false;
    if ((uint32_t)g < groups) {
      const uint32_t c = ff.sj[g] < ne ? sel8(ff.sj[g], x) : kInvalid;
      if (c != kInvalid)
        nid[g] = c == sel8(ff.sj[g], ff.id) ? ff.v[g] : adj0[(size_t)c * ff.m0 + ff.sp[g]];
    }
  }
  if (mode == 0) {
    lds_visit2(table, hcap, nid[0], nid[1], nw[0], nw[1]);
    if (groups > 2) lds_visit2(table, hcap, nid[2], nid[3], nw[2], nw[3]);
  } else {
#pragma unroll
    for (int g = 0; g < kBeamGroups; g++)
      if (nid[g] != kInvalid) nw[g] = bit_visit(spill, nid[g]);
  }
  uint64_t nm[kBeamGroups];
  uint32_t pre[kBeamGroups], nn = 0;
#pragma unroll
  for (int g = 0; g < kBeamGroups; g++) {
    nm[g] = __ballot(nw[g]);
    pre[g] = nn;
    nn += (uint32_t)popc64(nm[g]);
  }
  if (mode == 1 && nn > 0) {
    if (slog_n + nn <= slog_cap) {
#pragma unroll
      for (int g = 0; g < kBeamGroups; g++)
        if (nw[g]) slog[slog_n + pre[g] + mbcnt(nm[g])] = nid[g];
    }
    slog_n += nn;
  }
#pragma unroll
  for (int g = 0; g < kBeamGroups; g++)
    if (nw[g]) st_id[pre[g] + mbcnt(nm[g])] = nid[g];
  return nn;
}

// C[head..head+size) (held in e[]) merged with the batch elements in
// `amask` into C[0..): the merged order's first `keep` entries stay, keep =
// min(ccap, entries below pmin); the rest go to the pool unless dead
// (results full and distance above the worst), and pmin stays a lower
// bound of the pool.  With results full, dead entries are also trimmed from
// C's end.
template <int CPL>
__device__ __forceinline__ void c_write(uint64_t *C, const uint64_t (&e)[CPL],
                                        const uint32_t (&ep)[CPL], uint32_t ap, uint32_t &head,
                                        uint32_t &size, uint32_t ccap, uint64_t amask,
                                        uint64_t akey, bool wfull, float worst,
                                        gmut<uint64_t> pool, bool has_pool, uint32_t pool_cap,
                                        uint32_t &pn, uint64_t &pmin, unsigned long long &ovf,
                                        unsigned long long &moved, FrontFetch *ff = nullptr) {
  const int lane = lane_id();
  const bool mine = (amask >> lane) & 1;
  if (ff) front_capture(*ff, e[0], ep[0], ap, amask, akey, size);
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
template <int CPL>
__device__ __forceinline__ void c_merge(uint64_t *C, const uint64_t (&e)[CPL], uint32_t &head,
                                        uint32_t &size, uint32_t ccap, uint64_t amask,
                                        uint64_t akey, bool wfull, float worst,
                                        gmut<uint64_t> pool, bool has_pool, uint32_t pool_cap,
                                        uint32_t &pn, uint64_t &pmin, unsigned long long &ovf,
                                        unsigned long long &moved, FrontFetch *ff = nullptr) {
  uint32_t ep[CPL];
  const uint32_t ap = merge_ranks<CPL>(e, ep, size, amask, akey);
  c_write<CPL>(C, e, ep, ap, head, size, ccap, amask, akey, wfull, worst, pool, has_pool,
               pool_cap, pn, pmin, ovf, moved, ff);
}

// Ranks of one batch against C and W in a single pass over the batch (the
// beam steps): batch elements `amask` (words akey) go into C, the admitted
// ones `am` also into W.  Per element one broadcast; its rank in each list
// is counted with ballots and kept by its lane.  NC: the chunks of C
// that hold entries.
template <int CPL, int NC, int WPL>
__device__ __forceinline__ void joint_ranks_n(const uint64_t (&ce)[CPL], uint32_t (&cp)[CPL],
                                              uint32_t csize, const uint64_t (&we)[WPL],
                                              uint32_t (&wp)[WPL], uint32_t wsize, uint64_t amask,
                                              uint64_t am, uint64_t akey, uint32_t &apc,
                                              uint32_t &apw) {
  const int lane = lane_id();
  uint64_t vc[NC], vw[WPL];
#pragma unroll
  for (int c = 0; c < NC; c++) vc[c] = __ballot(c * 64 + lane < (int)csize);
#pragma unroll
  for (int c = 0; c < WPL; c++) vw[c] = __ballot(c * 64 + lane < (int)wsize);
#pragma unroll
  for (int c = 0; c < CPL; c++) cp[c] = c * 64 + lane;
#pragma unroll
  for (int c = 0; c < WPL; c++) wp[c] = c * 64 + lane;
  uint32_t bc = 0, bw = 0;       // batch elements (admitted ones) ordering before mine
  uint32_t rc = 0, rw = 0;       // my rank in C, in W (written by my element's step)
  for (uint64_t m = amask; m; m &= m - 1) {
    const int j = __ffsll((unsigned long long)m) - 1;
    const uint64_t kj = rl64(akey, j);
    const bool before = kj < akey;
    uint32_t r = 0;
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const bool lt = ce[c] < kj;
      r += (uint32_t)popc64(__ballot(lt) & vc[c]);
      cp[c] += lt ? 0u : 1u;
    }
    bc += before ? 1u : 0u;
    rc = lane == j ? r : rc;
    if ((am >> j) & 1) {
      uint32_t r2 = 0;
#pragma unroll
      for (int c = 0; c < WPL; c++) {
        const bool lt = we[c] < kj;
        r2 += (uint32_t)popc64(__ballot(lt) & vw[c]);
        wp[c] += lt ? 0u : 1u;
      }
      bw += before ? 1u : 0u;
      rw = lane == j ? r2 : rw;
    }
  }
  apc = rc + bc;
  apw = rw + bw;
}
// Ranks of a batch (amask, words akey) against the single list U of a
// beam step (entries e, expanded marks f, size entries): every element's
// merged position (ep for U's, the return value for the batch's), and for
// each batch element its rank among the merged order's unexpanded entries
// (ua: unexpanded U entries and batch entries before it).  N: the chunks of
// U that hold entries.
template <int SW, int N>
__device__ __forceinline__ uint32_t u_ranks_n(const uint64_t (&e)[SW], const uint32_t (&f)[SW],
                                              uint32_t (&ep)[SW], uint32_t size, uint64_t amask,
                                              uint64_t akey, uint32_t &ua) {
  const int lane = lane_id();
  uint64_t vc[N], uc[N];
#pragma unroll
  for (int c = 0; c < N; c++) {
    vc[c] = __ballot(c * 64 + lane < (int)size);
    uc[c] = __ballot(c * 64 + lane < (int)size && f[c] == 0u);
  }
#pragma unroll
  for (int c = 0; c < SW; c++) ep[c] = c * 64 + lane;
  uint32_t b = 0, rl_ = 0, ru = 0;
  for (uint64_t m = amask; m; m &= m - 1) {
    const int j = __ffsll((unsigned long long)m) - 1;
    const uint64_t kj = rl64(akey, j);
    uint32_t r = 0, u = 0;
#pragma unroll
    for (int c = 0; c < N; c++) {
      const uint64_t lt = __ballot(e[c] < kj);
      r += (uint32_t)popc64(lt & vc[c]);
      u += (uint32_t)popc64(lt & uc[c]);
      ep[c] += (e[c] < kj) ? 0u : 1u;
    }
    b += kj < akey ? 1u : 0u;
    rl_ = lane == j ? r : rl_;
    ru = lane == j ? u : ru;
  }
  ua = ru + b;
  return rl_ + b;
}
template <int SW>
__device__ __forceinline__ uint32_t u_ranks(const uint64_t (&e)[SW], const uint32_t (&f)[SW],
                                            uint32_t (&ep)[SW], uint32_t size, uint64_t amask,
                                            uint64_t akey, uint32_t &ua) {
  if (size <= 64) return u_ranks_n<SW, 1>(e, f, ep, size, amask, akey, ua);
  if constexpr (SW > 2) {
    if (size <= 128) return u_ranks_n<SW, 2>(e, f, ep, size, amask, akey, ua);
  }
  if constexpr (SW > 4) {
    if (size <= 256) return u_ranks_n<SW, 4>(e, f, ep, size, amask, akey, ua);
  }
  return u_ranks_n<SW, SW>(e, f, ep, size, amask, akey, ua);
}

template <int CPL, int WPL>
__device__ __forceinline__ void joint_ranks(const uint64_t (&ce)[CPL], uint32_t (&cp)[CPL],
                                            uint32_t csize, const uint64_t (&we)[WPL],
                                            uint32_t (&wp)[WPL], uint32_t wsize, uint64_t amask,
                                            uint64_t am, uint64_t akey, uint32_t &apc,
                                            uint32_t &apw) {
  if (csize <= 64) return joint_ranks_n<CPL, 1, WPL>(ce, cp, csize, we, wp, wsize, amask, am, akey, apc, apw);
  if constexpr (CPL > 2) {
    if (csize <= 128) return joint_ranks_n<CPL, 2, WPL>(ce, cp, csize, we, wp, wsize, amask, am, akey, apc, apw);
  }
  if constexpr (CPL > 4) {
    if (csize <= 256) return joint_ranks_n<CPL, 4, WPL>(ce, cp, csize, we, wp, wsize, amask, am, akey, apc, apw);
  }
  return joint_ranks_n<CPL, CPL, WPL>(ce, cp, csize, we, wp, wsize, amask, am, akey, apc, apw);
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

template <int TMAX, bool EXTRA, bool EXACT, class E, int B>
__device__ __forceinline__ void gather_round(const SearchArgs &P, const QLanes<TMAX, EXTRA> &q,
                                             const E *rows, const uint32_t *ids, int cnt,
                                             int base, float *out) {
  const int lane = lane_id(), i = lane & 15, quarter = lane >> 4;
  const Shape s{P.T, P.U, P.R, P.stride};
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

// bf16 rows of an exact layout (D = 32 * TMAX): B rows per quarter per
// round held as the loaded words (four per row at D = 128) and widened only
// as each product is formed, in the order of quarter_acc.
template <int TMAX, int B>
__device__ __forceinline__ void gather_round_bf16x(const SearchArgs &P, const QLanes<TMAX, false> &q,
                                                   const uint16_t *rows, const uint32_t *ids,
                                                   int cnt, int base, float *out) {
  const int lane = lane_id(), i = lane & 15, quarter = lane >> 4;
  constexpr int NB = TMAX / 4;
  const int groups = (cnt - base + 3) >> 2;
  uint2 ra[B][NB], rc[B][NB];
#pragma unroll
  for (int b = 0; b < B; b++) {
    const int row = base + 4 * b + quarter;
    const uint32_t rid = ids[row < cnt ? row : cnt - 1];
    gptr<uint16_t> p = G(rows) + (size_t)rid * P.stride;
#pragma unroll
    for (int k = 0; k < NB; k++) {
      if (b == 0 || b < groups) {
        ra[b][k] = *(gptr<uint2>)(p + k * 128 + 4 * i);
        rc[b][k] = *(gptr<uint2>)(p + k * 128 + 64 + 4 * i);
      } else {
        ra[b][k] = rc[b][k] = make_uint2(0u, 0u);
      }
    }
  }
#pragma unroll
  for (int b = 0; b < B; b++) {
    if (b > 0 && b >= groups) break;
    const int row = base + 4 * b + quarter;
    float lo = 0.0f, hi = 0.0f;
#pragma unroll
    for (int k = 0; k < NB; k++) {
      const uint32_t wa[2] = {ra[b][k].x, ra[b][k].y}, wc[2] = {rc[b][k].x, rc[b][k].y};
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const float xl = __uint_as_float(j & 1 ? wa[j >> 1] & 0xFFFF0000u : wa[j >> 1] << 16);
        const float xh = __uint_as_float(j & 1 ? wc[j >> 1] & 0xFFFF0000u : wc[j >> 1] << 16);
        if (P.metric == 0) {
          const float dl = q.lo[4 * k + j] - xl, dh = q.hi[4 * k + j] - xh;
          lo = __builtin_fmaf(dl, dl, lo);
          hi = __builtin_fmaf(dh, dh, hi);
        } else {
          lo = __builtin_fmaf(q.lo[4 * k + j], xl, lo);
          hi = __builtin_fmaf(q.hi[4 * k + j], xh, hi);
        }
      }
    }
    const float h = finish_metric(quarter_reduce<false>(lo, hi, 0.0f, 0, quarter), P.metric);
    if (i == 0 && row < cnt) out[row] = h;
  }
}

// The first round is straight-line code, so its loads issue right after
// the caller's (no loop-header wait for earlier loads in flight).
template <int TMAX, bool EXTRA, bool EXACT, class E, int B = rows_per_quarter<TMAX>()>
__device__ __forceinline__ void gather_rows(const SearchArgs &P, const QLanes<TMAX, EXTRA> &q,
                                            const E *rows, const uint32_t *ids, int cnt,
                                            float *out) {
  constexpr int R = 4 * B;
  gather_round<TMAX, EXTRA, EXACT, E, B>(P, q, rows, ids, cnt, 0, out);
  for (int base = R; base < cnt; base += R)
    gather_round<TMAX, EXTRA, EXACT, E, B>(P, q, rows, ids, cnt, base, out);
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

// Code rows a quarter gathers per round: in the code instances 16 (64 per
// round, a beam step's usual batch) for rows of one 128-byte chunk, 8 for
// two, 4 beyond; in the general instance 4 (its registers are shared with
// the f32 path).
template <int CC, int PC>
constexpr int code_rows_per_quarter() {
  return codes_class(PC) ? (CC <= 2 ? 8 : 4) : 4;
}

// Integer-code distances of rows ids[0..cnt): prec 1 the crate's codes
// (exact sums, src/kernels/int8.rs:13-48, rescaled as
// src/quantization.rs:85-92), prec 3 the per-dimension codes (exact
// sum_d qcode_d * code_d, then t = sigma * acc and L2 (qconst + rnorm) -
// (t + t), inner product -(qconst + t), cosine 1 - (qconst + t)).  A
// quarter takes several rows per round, every load issued first.
template <int CC, int PC>
__device__ __forceinline__ void gather_codes(const SearchArgs &P, const uint2 (&qc)[CC],
                                             const QCodeInfo &qs, int prec, const uint32_t *ids,
                                             int cnt, float *out, bool exact_layout) {
  const int lane = lane_id(), i = lane & 15, quarter = lane >> 4;
  constexpr int B = code_rows_per_quarter<CC, PC>();
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

// Integer-code distances of 128-byte code rows (dim <= 128) in the code
// instances: four lanes per row, 32 bytes each (two 16-byte loads), so one
// load instruction covers sixteen rows and B of them 16 * B rows per round;
// a row's exact sum is its four lanes' sums (two quad-permute adds).  The
// query code is held in the same layout: qv = bytes [32 t, 32 t + 32) for
// lane t of each quad.  Same sums and rescaling as gather_codes.
template <int B>
__device__ __forceinline__ void gather_codes4(const SearchArgs &P, const uint4 (&qv)[2],
                                              const QCodeInfo &qs, int prec, const uint32_t *ids,
                                              int cnt, float *out) {
  const int lane = lane_id(), t = lane & 3, r = lane >> 2;
  const bool pd = prec == 3;
  const int8_t *codes = pd ? P.pd_codes : P.codes;
  const bool l2 = P.metric == 0;
  const float s2 = __fmul_rn(P.sq_scale, P.sq_scale);
  for (int base = 0; base < cnt; base += 16 * B) {
    uint4 ca[B], cb[B];
    float rn[B];
#pragma unroll
    for (int b = 0; b < B; b++) {
      const int row = base + 16 * b + r;
      const uint32_t id = ids[row < cnt ? row : cnt - 1];
      gptr<uint4> cr = (gptr<uint4>)(codes + (size_t)id * 128 + 32 * t);
      const bool need = b == 0 || base + 16 * b < cnt;
      ca[b] = need ? (uint4)cr[0] : make_uint4(0u, 0u, 0u, 0u);
      cb[b] = need ? (uint4)cr[1] : make_uint4(0u, 0u, 0u, 0u);
      rn[b] = need && pd && l2 ? G(P.pd_rnorm)[id] : 0.0f;
    }
#pragma unroll
    for (int b = 0; b < B; b++) {
      if (b > 0 && base + 16 * b >= cnt) break;
      const int row = base + 16 * b + r;
      const int w[8] = {(int)ca[b].x, (int)ca[b].y, (int)ca[b].z, (int)ca[b].w,
                        (int)cb[b].x, (int)cb[b].y, (int)cb[b].z, (int)cb[b].w};
      const int qw[8] = {(int)qv[0].x, (int)qv[0].y, (int)qv[0].z, (int)qv[0].w,
                         (int)qv[1].x, (int)qv[1].y, (int)qv[1].z, (int)qv[1].w};
      int acc = 0;
      if (!pd && l2) {
        int ab = 0, bb = 0;
#pragma unroll
        for (int x = 0; x < 8; x++) {
          ab = __builtin_amdgcn_sdot4(qw[x], w[x], ab, false);
          bb = __builtin_amdgcn_sdot4(w[x], w[x], bb, false);
        }
        acc = bb - 2 * ab;
      } else {
#pragma unroll
        for (int x = 0; x < 8; x++) acc = __builtin_amdgcn_sdot4(qw[x], w[x], acc, false);
      }
      acc += dppi<kXor1>(acc);
      acc += dppi<kXor2>(acc);
      if (t == 0 && row < cnt) {
        float v;
        if (pd) {
          const float tt = __fmul_rn(qs.sigma, __int2float_rn(acc));
          if (l2) v = __fsub_rn(__fadd_rn(qs.qconst, rn[b]), __fadd_rn(tt, tt));
          else if (P.metric == 1) v = -__fadd_rn(qs.qconst, tt);
          else v = __fsub_rn(1.0f, __fadd_rn(qs.qconst, tt));
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
template <int TMAX, bool EXTRA, bool EXACT, int PC>
__device__ __forceinline__ void gather_dists(const SearchArgs &P, const QLanes<TMAX, EXTRA> &q,
                                             const uint2 (&qc)[code_chunks<TMAX, EXACT>()],
                                             const uint4 (&qv)[2], const QCodeInfo &qs, int prec,
                                             const uint32_t *ids, int cnt, float *out) {
  if (prec == 1 || prec == 3) {
    if constexpr (codes_class(PC) && TMAX <= 4) {
      if (P.cstride == 128) {
        gather_codes4<2>(P, qv, qs, prec, ids, cnt, out);
        return;
      }
    }
    gather_codes<code_chunks<TMAX, EXACT>(), PC>(P, qc, qs, prec, ids, cnt, out, EXACT);
    return;
  }
  // the bf16 instance takes twice the rows per round at D <= 128 (32: the
  // half-size rows of a beam step in one round trip)
  if constexpr (PC == kPrecBf16 && EXACT && TMAX <= 4) {
    if (prec == 2) {
      gather_round_bf16x<TMAX, 8>(P, q, P.bf16, ids, cnt, 0, out);
      for (int base = 32; base < cnt; base += 32)
        gather_round_bf16x<TMAX, 8>(P, q, P.bf16, ids, cnt, base, out);
      return;
    }
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

// EXACT: D == 32 * TMAX (branch-free row loads, see load_q).  PC: the
// precisions the instance handles (kPrecBf16, kPrecCodes for the
// per-dimension int8 codes, kPrecI8 for the crate's, kPrecF32, kPrecAny for
// every precision); a uniform batch runs the instance of its precision, so
// its loop carries no code or registers for the others.
template <int TMAX, bool EXTRA, bool EXACT, int SW, int PC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TMAX <= 8 ? 3 : 2)))
hnsw_search_kernel(const SearchArgs P) {
  extern __shared__ uint32_t lds[];
  const int lane = lane_id(), i16 = lane & 15;
  const int wave = threadIdx.x >> 6;
  const uint32_t slot = blockIdx.x * (blockDim.x >> 6) + wave;
  uint32_t *table = lds + (size_t)wave * P.lds_words;
  uint32_t *st_id = table + P.hash_cap;
  float *st_d = reinterpret_cast<float *>(st_id + P.st_cap);
  uint64_t *Ck = reinterpret_cast<uint64_t *>(st_d + P.st_cap);  // 128 * SW candidate words
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

  // The re-run launch follows every batch that may overflow a pool; with
  // nothing listed its waves leave before touching the shared counter.
  if (P.phase == 3 && rl(G(P.redo_count)[0], 0) == 0) return;

  // Phase 1 (descent only, nearly uniform work per query) takes queries
  // statically, without the shared counter.
  uint32_t static_next = slot;
  const uint32_t total_waves = gridDim.x * (blockDim.x >> 6);
  bool first = true;
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
      // each wave's first query is its slot; later ones come from the
      // shared counter, numbered past those (so the launch does not start
      // with every wave's atomic on one address)
      qi = first ? slot : total_waves + rl(atomicAdd(P.counter, lane == 0 ? 1u : 0u), 0);
    }
    first = false;
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
    if constexpr (PC == kPrecBf16) prec = 2;
    if constexpr (PC == kPrecCodes) prec = 3;
    if constexpr (PC == kPrecI8) prec = 1;
    if constexpr (PC == kPrecF32) prec = 0;
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

    // query, in the row layout (code instances load it for the re-rank
    // only), plus its int8 code words
    QLanes<TMAX, EXTRA> q;
#pragma unroll
    for (int t = 0; t < TMAX; t++) q.lo[t] = q.hi[t] = 0.0f;
    q.e[0] = q.r = 0.0f;
    if constexpr (EXTRA) q.e[1] = q.e[2] = 0.0f;
    if constexpr (!codes_class(PC)) load_q<TMAX, EXTRA, float>(P.qprep + (size_t)qi * P.stride, shp, i16, q);
    uint2 qc[code_chunks<TMAX, EXACT>()];
#pragma unroll
    for (int c = 0; c < code_chunks<TMAX, EXACT>(); c++) qc[c] = make_uint2(0, 0);
    QCodeInfo qs{0, 1.0f, 0.0f};
    // the code instances hold 128-byte query codes four lanes per row
    uint4 qv[2] = {make_uint4(0u, 0u, 0u, 0u), make_uint4(0u, 0u, 0u, 0u)};
    if (prec == 1 || prec == 3) {
      bool quad = false;
      if constexpr (codes_class(PC) && TMAX <= 4) quad = P.cstride == 128;
      if (quad) {
        gptr<uint4> cq = (gptr<uint4>)(P.qcode + (size_t)qi * 128 + 32 * (lane & 3));
        qv[0] = cq[0];
        qv[1] = cq[1];
        if (prec == 1 && P.metric == 0) {
          const int qw[8] = {(int)qv[0].x, (int)qv[0].y, (int)qv[0].z, (int)qv[0].w,
                             (int)qv[1].x, (int)qv[1].y, (int)qv[1].z, (int)qv[1].w};
          int aa = 0;
#pragma unroll
          for (int x = 0; x < 8; x++) aa = __builtin_amdgcn_sdot4(qw[x], qw[x], aa, false);
          aa += dppi<kXor1>(aa);
          aa += dppi<kXor2>(aa);
          qs.qaa = aa;
        }
      } else {
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
        }
      }
      if (prec == 3) {
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
      gather_dists<TMAX, EXTRA, EXACT, PC>(P, q, qc, qv, qs, (int)prec, st_id, 1, st_d);
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
            gather_dists<TMAX, EXTRA, EXACT, PC>(P, q, qc, qv, qs, (int)prec, st_id, c, st_d);
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
    // this query's layer-0 counters, added to the wave's at its end (or
    // dropped with the attempt when it is re-run)
    uint32_t q_dist = 0, q_exp = 0, q_spill = 0;
    unsigned long long q_ovf = 0, q_cspill = 0;
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
    // refill C from the pool in chunks of 64, each merged like a neighbour
    // batch; what does not fit goes back to the pool, which is compacted in
    // place (a chunk writes at most as many entries as it read)
    auto refill = [&]() {
      const uint32_t total = pn;
      const bool wfull = wsize >= ef;
      pn = 0;
      pmin = ~0ull;
      for (uint32_t base = 0; base < total; base += 64) {
        const uint32_t c = total - base < 64 ? total - base : 64;
        const uint64_t x = lane < (int)c ? (uint64_t)pool[base + lane] : ~0ull;
        const uint64_t live = __ballot(lane < (int)c && !(wfull && ckey_d(x) > worst));
        uint64_t ce[2 * SW];
        list_load<2 * SW>(Ck + head, csize, ce);
        c_merge<2 * SW>(Ck, ce, head, csize, ccap, live, x, wfull, worst, pool, has_pool,
                        P.cpool_cap, pn, pmin, q_ovf, q_cspill);
      }
    };
    // Moves the LDS visited table into the wave's HBM bitset (the query
    // continues there, logging what it marks).
    auto leave_lds = [&]() {
      for (uint32_t t = lane; t < hcap; t += 64) {
        const uint32_t v = table[t];
        if (v != kInvalid)
          __hip_atomic_fetch_or(spill + (v >> 5), 1u << (v & 31), __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
      }
      __threadfence();
      mode = 1;
      q_spill++;
    };

    if (beam_class(PC) || ((prec == 2 || prec == 3) && P.beam > 1)) {
      // ---- beam steps (bf16 and per-dimension int8 traversals)
      const uint32_t m0 = P.m0, beam = P.beam;
      FrontFetch ff;
      ff.adj0 = P.adj0;
      ff.m0 = m0;
      ff.n = beam;
      ff.groups = (beam * m0 + 63) / 64;
#pragma unroll
      for (int g = 0; g < kBeamGroups; g++) {
        ff.sj[g] = (uint32_t)(lane + 64 * g) / m0;
        ff.sp[g] = (uint32_t)(lane + 64 * g) - ff.sj[g] * m0;
        ff.v[g] = kInvalid;
      }
#pragma unroll
      for (int t = 0; t < kBeamMax; t++) ff.id[t] = kInvalid;
      uint32_t x[kBeamMax];
      if (allow == nullptr && P.deleted_bits == nullptr) {
        // -- every visited node is admissible: one list U (= W) of the ef
        // smallest seen, each entry marked once expanded; the candidates are
        // its unexpanded entries (C and the pool stay unused)
        uint32_t *F = reinterpret_cast<uint32_t *>(Ck);  // marks, by U position
        uint64_t ue[SW];
        uint32_t uf[SW];
        list_load<SW>(Wk, wsize, ue);
#pragma unroll
        for (int c = 0; c < SW; c++) uf[c] = 0u;
        // the first `beam` unexpanded entries, marked: their ids in x
        auto take_front = [&]() {
          uint32_t got = 0;
#pragma unroll
          for (int t = 0; t < kBeamMax; t++) x[t] = kInvalid;
#pragma unroll
          for (int c = 0; c < SW; c++) {
            if (got >= beam || c * 64 >= (int)wsize) break;
            const bool cand = c * 64 + lane < (int)wsize && uf[c] == 0u;
            const uint32_t r = got + (uint32_t)mbcnt(__ballot(cand));
            const bool pick = cand && r < beam;
            for (uint64_t m = __ballot(pick); m; m &= m - 1) {
              const uint32_t id = rl((uint32_t)ue[c], __ffsll((unsigned long long)m) - 1);
#pragma unroll
              for (int t = 0; t < kBeamMax; t++) x[t] = got == (uint32_t)t ? id : x[t];
              got++;
            }
            uf[c] = pick ? 1u : uf[c];
          }
          return got;
        };
        uint32_t ne = take_front();
        while (ne > 0) {
          q_exp += ne;
          // the table never fills: leave it before a step could pass its limit
          if (mode == 0 && spill && vcount + ne * m0 > P.hash_limit) leave_lds();
          const uint32_t nn = beam_visit(adj0, x, ne, ff, table, hcap, mode, spill, slog,
                                         P.spill_log_cap, slog_n, st_id);
          vcount += nn;
          q_dist += nn;
          bool fetched = false;
          if (nn > 0) {
            wave_sync();
            gather_dists<TMAX, EXTRA, EXACT, PC>(P, q, qc, qv, qs, (int)prec, st_id, (int)nn, st_d);
            wave_sync();
            // every new neighbour against U as it stood at the step's start
            const bool wfull0 = wsize >= ef;
            const float worst0 = worst;
            for (uint32_t base = 0; base < nn; base += 64) {
              const uint32_t i = base + (uint32_t)lane;
              const uint32_t myid = i < nn ? st_id[i] : kInvalid;
              const float myd = i < nn ? st_d[i] : 0.0f;
              const bool surv = i < nn && (!wfull0 || myd < worst0);
              const uint64_t amask = __ballot(surv);
              if (!amask) continue;
              const uint64_t akey = surv ? ckey(myd, myid) : ~0ull;
              uint32_t wp[SW], ua;
              const uint32_t ap = u_ranks<SW>(ue, uf, wp, wsize, amask, akey, ua);
              const uint32_t total = wsize + (uint32_t)popc64(amask);
              const uint32_t nsize = total < ef ? total : ef;
              if (base + 64 >= nn) {
                // the merged order's first `beam` unexpanded entries are the
                // next step's pops: their lists load during the writes.  A U
                // entry's unexpanded rank counts the unexpanded U entries and
                // the batch entries before it; a batch entry's is ua.
#pragma unroll
                for (int t = 0; t < kBeamMax; t++) ff.id[t] = kInvalid;
                uint32_t uprev = 0;
#pragma unroll
                for (int c = 0; c < SW; c++) {
                  if (c > 0 && c * 64 >= (int)wsize) break;
                  const bool un = c * 64 + lane < (int)wsize && uf[c] == 0u;
                  const uint64_t um = __ballot(un);
                  const uint32_t ur =
                      uprev + (uint32_t)mbcnt(um) + (wp[c] - (uint32_t)(c * 64 + lane));
#pragma unroll
                  for (int t = 0; t < kBeamMax; t++) {
                    const uint64_t hit =
                        __ballot(un && wp[c] < nsize && ur == (uint32_t)t && t < (int)beam);
                    if (hit) ff.id[t] = rl((uint32_t)ue[c], __ffsll((unsigned long long)hit) - 1);
                  }
                  uprev += (uint32_t)popc64(um);
                }
#pragma unroll
                for (int t = 0; t < kBeamMax; t++) {
                  const uint64_t hit = __ballot(surv && ap < nsize && ua == (uint32_t)t && t < (int)beam);
                  if (hit) ff.id[t] = rl((uint32_t)akey, __ffsll((unsigned long long)hit) - 1);
                }
                front_fetch(ff);
                fetched = true;
              }
#pragma unroll
              for (int c = 0; c < SW; c++) {
                if (c > 0 && c * 64 >= (int)wsize) break;
                if (c * 64 + lane < (int)wsize && wp[c] < nsize) {
                  Wk[wp[c]] = ue[c];
                  F[wp[c]] = uf[c];
                }
              }
              if (surv && ap < nsize) {
                Wk[ap] = akey;
                F[ap] = 0u;
              }
              wave_sync();
              wsize = nsize;
              list_load<SW>(Wk, wsize, ue);
#pragma unroll
              for (int c = 0; c < SW; c++) {
                const uint32_t xi = c * 64 + lane;
                uf[c] = xi < wsize ? F[xi] : 1u;
              }
              worst = ckey_d(Wk[wsize - 1]);
              wave_sync();
            }
          }
          ne = take_front();
          if (!fetched) {
#pragma unroll
            for (int t = 0; t < kBeamMax; t++) ff.id[t] = x[t];
            front_fetch(ff);
          }
        }
      } else
      for (;;) {
        // -- two lists (filters or tombstones: candidates that fail the
        // filter stay traversable)
        if (csize < beam && pn > 0) refill();
        if (csize == 0) break;
        // pop up to `beam` candidates while the stop rule does not hold
        const uint32_t avail = csize < beam ? csize : beam;
        const uint64_t fe = lane < (int)avail ? Ck[head + lane] : ~0ull;
        const uint64_t stop =
            __ballot(lane < (int)avail && wsize >= ef && ckey_d(fe) > worst);
        const uint32_t ne = stop ? (uint32_t)(__ffsll((unsigned long long)stop) - 1) : avail;
        if (ne == 0) break;
#pragma unroll
        for (int t = 0; t < kBeamMax; t++) x[t] = rl((uint32_t)fe, t);
        head += ne;
        csize -= ne;
        q_exp += ne;
        // the table never fills: leave it before a step could pass its limit
        if (mode == 0 && spill && vcount + ne * m0 > P.hash_limit) leave_lds();
        const uint32_t nn = beam_visit(adj0, x, ne, ff, table, hcap, mode, spill, slog,
                                       P.spill_log_cap, slog_n, st_id);
        vcount += nn;
        q_dist += nn;
        bool fetched = false;
        if (nn > 0) {
          wave_sync();
          gather_dists<TMAX, EXTRA, EXACT, PC>(P, q, qc, qv, qs, (int)prec, st_id, (int)nn, st_d);
          wave_sync();
          // every new neighbour against the results at the step's start
          const bool wfull0 = wsize >= ef;
          const float worst0 = worst;
          for (uint32_t base = 0; base < nn; base += 64) {
            const uint32_t i = base + (uint32_t)lane;
            const uint32_t myid = i < nn ? st_id[i] : kInvalid;
            const float myd = i < nn ? st_d[i] : 0.0f;
            const bool surv = i < nn && (!wfull0 || myd < worst0);
            const uint64_t amask = __ballot(surv);
            if (!amask) continue;
            const uint64_t am = __ballot(surv && admit(P, allow, myid));
            const uint64_t akey = surv ? ckey(myd, myid) : ~0ull;
            // one pass over the survivors ranks them in C and W; W is written
            // first so C drops what the new worst makes dead
            uint64_t ce[2 * SW], we[SW];
            list_load<2 * SW>(Ck + head, csize, ce);
            list_load<SW>(Wk, wsize, we);
            uint32_t cp[2 * SW], wp[SW], apc, apw;
            joint_ranks<2 * SW, SW>(ce, cp, csize, we, wp, wsize, amask, am, akey, apc, apw);
            // C's new front is known: its lists load during the writes
            front_capture(ff, ce[0], cp[0], apc, amask, akey, csize);
            fetched = true;
            if (am) worst = ckey_d(w_write<SW>(Wk, we, wp, apw, wsize, ef, am, akey));
            c_write<2 * SW>(Ck, ce, cp, apc, head, csize, ccap, amask, akey, wsize >= ef, worst,
                            pool, has_pool, P.cpool_cap, pn, pmin, q_ovf, q_cspill);
          }
        }
        if (!fetched) {
          // C's front is unchanged: fetch its lists for the next step
          const uint64_t nf = lane < (int)beam && (uint32_t)lane < csize ? Ck[head + lane] : ~0ull;
#pragma unroll
          for (int t = 0; t < kBeamMax; t++) ff.id[t] = rl((uint32_t)nf, t);
          front_fetch(ff);
        }
      }
    } else
    for (;;) {
      if (csize == 0) {
        if (pn == 0) break;
        head = 0;
        refill();
        if (csize == 0) break;
      }
      const uint64_t top = Ck[head];
      const float cd = ckey_d(top);
      const uint32_t cid = (uint32_t)top;
      head++;
      csize--;
      if (wsize >= ef && cd > worst) break;
      q_exp++;
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
        gather_dists<TMAX, EXTRA, EXACT, PC>(P, q, qc, qv, qs, (int)prec, st_id, nn, st_d);
        wave_sync();
        const float myd = lane < nn ? st_d[lane] : 0.0f;
        const uint64_t am = __ballot(adm);
        vcount += nn;
        q_dist += nn;
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
          // one pass over the accepted neighbours ranks them in C and (the
          // admitted ones) in W
          uint64_t ce[2 * SW];
          list_load<2 * SW>(Ck + head, csize, ce);
          uint32_t cp[2 * SW], wp[SW], apc, apw;
          joint_ranks<2 * SW, SW>(ce, cp, csize, we, wp, wsize, amask, amask & am, akey, apc, apw);
          if (amask & am) worst = ckey_d(w_write<SW>(Wk, we, wp, apw, wsize, ef, amask & am, akey));
          c_write<2 * SW>(Ck, ce, cp, apc, head, csize, ccap, amask, akey, wsize >= ef, worst,
                          pool, has_pool, P.cpool_cap, pn, pmin, q_ovf, q_cspill);
        }
        if (mode == 0 && vcount > P.hash_limit && spill) leave_lds();
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

    if (P.redo && P.phase != 3 && q_ovf) {
      // the pool dropped a candidate: list the query for the phase-3 launch
      // (which starts from the same layer-0 entry) and drop this attempt
      st.redo++;
      seed_query(P, qi, ep_id, ep_d);
      const uint32_t at = rl(atomicAdd(P.redo_count, lane == 0 ? 1u : 0u), 0);
      if (at < P.nq) GM(P.redo)[at] = qi;
      wave_sync();
      continue;
    }

    st.dist += q_dist;
    st.exp += q_exp;
    st.spill += q_spill;
    st.ovf += q_ovf;
    st.cspill += q_cspill;

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
    if constexpr (codes_class(PC)) load_q<TMAX, EXTRA, float>(P.qprep + (size_t)qi * P.stride, shp, i16, q);
    if (prec != 0) {
      for (int base = 0; base < wsize; base += 64) {
        const int c = wsize - base < 64 ? wsize - base : 64;
        gather_dists<TMAX, EXTRA, EXACT, PC>(P, q, qc, qv, qs, 0, res_id + base, c, res_d + base);
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
  // this wave's line of the counters; zero counts are not added
  if (lane == 0 && P.stats) {
    unsigned long long *sl = P.stats + (size_t)(slot % kStatLines) * kStatStride;
    if (st.q) atomicAdd(&sl[0], st.q);
    if (st.dist) atomicAdd(&sl[1], st.dist);
    if (st.up) atomicAdd(&sl[2], st.up);
    if (st.exp) atomicAdd(&sl[3], st.exp);
    if (st.rr) atomicAdd(&sl[4], st.rr);
    if (st.ovf) atomicAdd(&sl[5], st.ovf);
    if (st.spill) atomicAdd(&sl[6], st.spill);
    if (st.cspill) atomicAdd(&sl[7], st.cspill);
    if (st.redo) atomicAdd(&sl[8], st.redo);
  }
}

}  // namespace hhk

// Explicit instances of one precision class (used by the search_*.hip files).
#define HH_SEARCH_INST_SW(TM, EX, XA, PC)                                                    \
  template __global__ void hhk::hnsw_search_kernel<TM, EX, XA, 1, PC>(const hhk::SearchArgs); \
  template __global__ void hhk::hnsw_search_kernel<TM, EX, XA, 2, PC>(const hhk::SearchArgs); \
  template __global__ void hhk::hnsw_search_kernel<TM, EX, XA, 4, PC>(const hhk::SearchArgs); \
  template __global__ void hhk::hnsw_search_kernel<TM, EX, XA, 8, PC>(const hhk::SearchArgs);
#define HH_SEARCH_INST_CLASS(PC)             \
  HH_SEARCH_INST_SW(4, false, false, PC)     \
  HH_SEARCH_INST_SW(4, true, false, PC)      \
  HH_SEARCH_INST_SW(4, false, true, PC)      \
  HH_SEARCH_INST_SW(8, false, false, PC)     \
  HH_SEARCH_INST_SW(8, true, false, PC)      \
  HH_SEARCH_INST_SW(8, false, true, PC)      \
  HH_SEARCH_INST_SW(32, false, false, PC)    \
  HH_SEARCH_INST_SW(32, true, false, PC)     \
  HH_SEARCH_INST_SW(32, false, true, PC)
