// serve.hip — continuous-batching request loop over a device-resident
// index (no crate counterpart: the crate serves one `search` call per
// thread, src/hnsw/mod.rs:412-455; this is the batched form of that loop).
//
// Requests (one query, its hh_query_spec, optionally a registered filter)
// queue on the host.  A server thread keeps up to `max_inflight` launches
// on the GPU, each on its own stream and workspace: whenever a launch
// slot is free it admits up to `max_batch` waiting requests, copies them
// into pinned staging, launches the mixed-spec search and the copy of the
// results back, and records an event.  Requests arriving while batches
// run are admitted as soon as a slot frees, so their work overlaps the
// running batches on the device.
//
// Admission follows the device budget of a launch.  A launch's LDS shape
// (per-wave lists and visited table) is set by its largest ef, so requests
// wait in one queue per ef bucket (max(ef, k) <= 64, 128, 256, 512) and a
// launch takes requests of one bucket only: the bucket of the oldest
// waiting request, up to the wavefronts the kernel keeps resident at that
// bucket's shape.  Small-ef requests therefore never run at the occupancy
// of a large-ef one, and requests are served in arrival order across
// buckets.  Device memory is fixed at open: each slot's buffers hold the
// largest batch, and the index's workspaces are reserved for it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "hh_internal.h"

using hh::fail;

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  uint64_t ticket;
  std::vector<float> query;
  hh_query_spec spec;
  Clock::time_point arrival;
};

struct Result {
  std::vector<uint64_t> keys;
  std::vector<float> dist;
  uint32_t count = 0;
  int32_t status = 0;
  double latency_ms = 0.0;
};

template <class T>
int halloc(T **p, size_t n) {
  *p = nullptr;
  if (hipHostMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T), hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    return fail(HH_OUT_OF_MEMORY, "pinned host allocation failed");
  }
  return HH_OK;
}
template <class T>
int dalloc(T **p, size_t n) {
  *p = nullptr;
  if (hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
    (void)hipGetLastError();
    return fail(HH_OUT_OF_MEMORY, "device allocation failed");
  }
  return HH_OK;
}

// One launch slot: stream, pinned staging, device buffers, the batch it
// carries while in flight.
struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  float *h_q = nullptr, *d_q = nullptr, *h_d = nullptr, *d_d = nullptr;
  hh_query_spec *h_s = nullptr, *d_s = nullptr;
  uint64_t *h_k = nullptr, *d_k = nullptr;
  uint32_t *h_c = nullptr, *d_c = nullptr;
  int32_t *h_st = nullptr, *d_st = nullptr;
  std::vector<Request> batch;
  bool busy = false;
};

// Queue of a request: the list-size bucket of max(ef, k).
int bucket_of(const hh_query_spec &sp) {
  const uint32_t e = std::max(sp.ef, sp.k);
  return e <= 64 ? 0 : e <= 128 ? 1 : e <= 256 ? 2 : 3;
}

}  // namespace

struct hh_server {
  hh_index *idx = nullptr;
  int device = 0;
  uint32_t dim = 0, max_batch = 0, k_max = 0, max_inflight = 0;
  uint32_t bucket_batch[4] = {0, 0, 0, 0};  // launch size per ef bucket
  uint64_t n_words = 0;  // words per registered filter bitmap
  std::vector<std::unique_ptr<Slot>> slots;
  uint32_t *d_filters = nullptr;  // registered filters, filter_cap bitmaps
  uint32_t filter_cap = 0, nfilters = 0;
  std::mutex mu;
  std::condition_variable cv, done_cv;
  std::deque<Request> pending[4];  // one queue per ef bucket
  uint64_t npending = 0;
  std::unordered_map<uint64_t, Result> results;
  uint64_t next_ticket = 1;
  uint64_t served = 0, batches = 0, max_queue = 0;
  bool stop = false;
  int error = HH_OK;
  std::string error_msg;
  std::thread worker;

  void run();
  int launch(Slot &s, uint32_t nfilt);
  void finish(Slot &s);
  // Moves the next launch's requests into s.batch: the bucket of the
  // oldest waiting request, up to that bucket's launch size.
  void admit(Slot &s);
};

void hh_server::admit(Slot &s) {
  int b = -1;
  for (int i = 0; i < 4; i++)
    if (!pending[i].empty() && (b < 0 || pending[i].front().ticket < pending[b].front().ticket))
      b = i;
  s.batch.clear();
  if (b < 0) return;
  auto &q = pending[b];
  while (!q.empty() && s.batch.size() < bucket_batch[b]) {
    s.batch.push_back(std::move(q.front()));
    q.pop_front();
  }
  npending -= s.batch.size();
}

void hh_server::run() {
  (void)hipSetDevice(device);
  std::unique_lock<std::mutex> lock(mu);
  for (;;) {
    // collect finished launches
    bool progressed = false;
    for (auto &sp : slots) {
      Slot &s = *sp;
      if (s.busy && hipEventQuery(s.done) == hipSuccess) {
        finish(s);
        progressed = true;
      }
    }
    if (progressed) done_cv.notify_all();
    // admit waiting requests into free slots
    for (auto &sp : slots) {
      Slot &s = *sp;
      if (s.busy || npending == 0) continue;
      admit(s);
      // the filter table as registered so far (register_filter appends
      // under the lock; launches read only entries below this count)
      const uint32_t nfilt = nfilters;
      lock.unlock();
      const int st = launch(s, nfilt);
      lock.lock();
      if (st) {
        error = st;
        error_msg = hh::last_error();
        for (auto &r : s.batch) {
          Result res;
          res.status = st;
          results[r.ticket] = std::move(res);
        }
        s.batch.clear();
        done_cv.notify_all();
      } else {
        s.busy = true;
      }
    }
    bool any_busy = false;
    for (auto &sp : slots) any_busy |= sp->busy;
    if (stop && !any_busy && npending == 0) break;
    if (any_busy) {
      // poll the in-flight events without holding the lock for long
      lock.unlock();
      std::this_thread::sleep_for(std::chrono::microseconds(20));
      lock.lock();
    } else {
      cv.wait(lock, [&] { return stop || npending > 0; });
    }
  }
}

int hh_server::launch(Slot &s, uint32_t nfilt) {
  const size_t nb = s.batch.size();
  uint32_t ef_max = 1;
  for (size_t i = 0; i < nb; i++) {
    std::memcpy(s.h_q + i * dim, s.batch[i].query.data(), dim * sizeof(float));
    s.h_s[i] = s.batch[i].spec;
    ef_max = std::max(ef_max, std::max(s.h_s[i].ef, s.h_s[i].k));
  }
  // the launch is sized for this batch's largest ef, not the largest
  // supported one (larger specs fail per request)
  ef_max = std::min<uint32_t>(ef_max, 512);
  if (hipMemcpyAsync(s.d_q, s.h_q, nb * dim * 4, hipMemcpyHostToDevice, s.stream) != hipSuccess ||
      hipMemcpyAsync(s.d_s, s.h_s, nb * sizeof(hh_query_spec), hipMemcpyHostToDevice, s.stream) != hipSuccess)
    return fail(HH_DEVICE_ERROR, "staging copy failed");
  int st = hh::search_spec_device_strided(idx, s.d_q, nb, s.d_s, k_max, ef_max,
                                          nfilt ? d_filters : nullptr, nfilt, n_words, s.d_k,
                                          s.d_d, s.d_c, s.d_st, s.stream);
  if (st) return st;
  if (hipMemcpyAsync(s.h_k, s.d_k, nb * k_max * 8, hipMemcpyDeviceToHost, s.stream) != hipSuccess ||
      hipMemcpyAsync(s.h_d, s.d_d, nb * k_max * 4, hipMemcpyDeviceToHost, s.stream) != hipSuccess ||
      hipMemcpyAsync(s.h_c, s.d_c, nb * 4, hipMemcpyDeviceToHost, s.stream) != hipSuccess ||
      hipMemcpyAsync(s.h_st, s.d_st, nb * 4, hipMemcpyDeviceToHost, s.stream) != hipSuccess ||
      hipEventRecord(s.done, s.stream) != hipSuccess)
    return fail(HH_DEVICE_ERROR, "result copy failed");
  return HH_OK;
}

// Called with the lock held, once the slot's event has completed.
void hh_server::finish(Slot &s) {
  const auto now = Clock::now();
  for (size_t i = 0; i < s.batch.size(); i++) {
    Result r;
    r.count = s.h_c[i];
    r.status = s.h_st[i];
    r.keys.assign(s.h_k + i * k_max, s.h_k + i * k_max + r.count);
    r.dist.assign(s.h_d + i * k_max, s.h_d + i * k_max + r.count);
    r.latency_ms = std::chrono::duration<double, std::milli>(now - s.batch[i].arrival).count();
    results[s.batch[i].ticket] = std::move(r);
  }
  served += s.batch.size();
  batches++;
  s.batch.clear();
  s.busy = false;
}

extern "C" {

int hh_serve_open(hh_index *idx, const hh_serve_config *cfg, hh_server **out) {
  if (!idx || !cfg || !out) return fail(HH_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (cfg->k_max == 0 || cfg->k_max > 512) return fail(HH_INVALID_PARAMETER, "k_max must be in 1..512");
  std::unique_ptr<hh_server> s(new hh_server());
  s->idx = idx;
  s->device = hh::index_device(idx);
  s->dim = hh::index_dim(idx);
  const uint64_t n = hh::index_nodes(idx);
  s->n_words = (n + 31) / 32;
  s->k_max = cfg->k_max;
  s->max_inflight = std::max<uint32_t>(1, std::min<uint32_t>(cfg->max_inflight ? cfg->max_inflight : 2, 4));
  s->max_batch = cfg->max_batch ? cfg->max_batch : hh::index_resident_waves(idx);
  s->max_batch = std::max<uint32_t>(1, s->max_batch);
  for (int b = 0; b < 4; b++) {
    // auto: the resident wavefronts at the bucket's launch shape
    const uint32_t waves = hh::index_resident_waves(idx, 64u << b);
    s->bucket_batch[b] = cfg->max_batch ? s->max_batch : std::min(s->max_batch, std::max(1u, waves));
  }
  s->max_queue = cfg->max_queue ? cfg->max_queue : (uint64_t)s->max_batch * 64;
  s->filter_cap = cfg->max_filters;
  if (hipSetDevice(s->device) != hipSuccess) return fail(HH_DEVICE_ERROR, "hipSetDevice failed");
  if (s->filter_cap) {
    int st = dalloc(&s->d_filters, s->filter_cap * std::max<uint64_t>(1, s->n_words));
    if (st) return st;
  }
  const size_t B = s->max_batch, K = s->k_max;
  for (uint32_t i = 0; i < s->max_inflight; i++) {
    std::unique_ptr<Slot> sl(new Slot());
    if (hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&sl->done, hipEventDisableTiming) != hipSuccess)
      return fail(HH_DEVICE_ERROR, "stream creation failed");
    int st = halloc(&sl->h_q, B * s->dim);
    if (!st) st = dalloc(&sl->d_q, B * s->dim);
    if (!st) st = halloc(&sl->h_s, B);
    if (!st) st = dalloc(&sl->d_s, B);
    if (!st) st = halloc(&sl->h_k, B * K);
    if (!st) st = dalloc(&sl->d_k, B * K);
    if (!st) st = halloc(&sl->h_d, B * K);
    if (!st) st = dalloc(&sl->d_d, B * K);
    if (!st) st = halloc(&sl->h_c, B);
    if (!st) st = dalloc(&sl->d_c, B);
    if (!st) st = halloc(&sl->h_st, B);
    if (!st) st = dalloc(&sl->d_st, B);
    s->slots.push_back(std::move(sl));
    if (st) return st;
  }
  {
    // every workspace a launch can take (the index keeps up to four) sized
    // now, not on the first large batch
    int st = hh::index_reserve(idx, s->max_batch, 4);
    if (st) return st;
  }
  hh_server *raw = s.get();
  hh::index_serving(idx, +1);
  raw->worker = std::thread([raw] { raw->run(); });
  *out = s.release();
  return HH_OK;
}

int hh_serve_register_filter(hh_server *s, const uint32_t *allow_bits, uint32_t *filter_id) {
  if (!s || !allow_bits || !filter_id) return fail(HH_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> lock(s->mu);
  if (s->nfilters >= s->filter_cap) return fail(HH_CAPACITY_EXCEEDED, "filter table is full");
  if (hipSetDevice(s->device) != hipSuccess ||
      hipMemcpy(s->d_filters + (size_t)s->nfilters * s->n_words, allow_bits, s->n_words * 4,
                hipMemcpyHostToDevice) != hipSuccess)
    return fail(HH_DEVICE_ERROR, "filter upload failed");
  *filter_id = s->nfilters++;
  return HH_OK;
}

int hh_serve_submit(hh_server *s, const float *query, const hh_query_spec *spec, uint64_t *ticket) {
  if (!s || !query || !spec || !ticket) return fail(HH_INVALID_ARGUMENT, "null argument");
  if (spec->k > s->k_max) return fail(HH_INVALID_PARAMETER, "k above the server's k_max");
  Request r;
  r.query.assign(query, query + s->dim);
  r.spec = *spec;
  r.arrival = Clock::now();
  {
    std::lock_guard<std::mutex> lock(s->mu);
    if (s->stop) return fail(HH_INVALID_ARGUMENT, "server is closing");
    if (s->npending >= s->max_queue)
      return fail(HH_CAPACITY_EXCEEDED, "request queue is full");
    r.ticket = s->next_ticket++;
    *ticket = r.ticket;
    s->pending[bucket_of(r.spec)].push_back(std::move(r));
    s->npending++;
  }
  s->cv.notify_one();
  return HH_OK;
}

int hh_serve_poll(hh_server *s, uint64_t ticket, int wait, uint64_t *keys, float *dist,
                  uint32_t *count, int32_t *status, double *latency_ms, int *done) {
  if (!s || !done) return fail(HH_INVALID_ARGUMENT, "null argument");
  std::unique_lock<std::mutex> lock(s->mu);
  auto it = s->results.find(ticket);
  if (it == s->results.end() && wait) {
    s->done_cv.wait(lock, [&] { return s->results.count(ticket) > 0 || (s->stop && s->npending == 0); });
    it = s->results.find(ticket);
  }
  if (it == s->results.end()) {
    *done = 0;
    return HH_OK;
  }
  const Result &r = it->second;
  if (count) *count = r.count;
  if (status) *status = r.status;
  if (latency_ms) *latency_ms = r.latency_ms;
  if (keys) std::memcpy(keys, r.keys.data(), r.count * 8);
  if (dist) std::memcpy(dist, r.dist.data(), r.count * 4);
  *done = 1;
  s->results.erase(it);
  return HH_OK;
}

int hh_serve_stats(hh_server *s, uint64_t *served, uint64_t *batches, uint64_t *queued) {
  if (!s) return fail(HH_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> lock(s->mu);
  if (served) *served = s->served;
  if (batches) *batches = s->batches;
  if (queued) *queued = s->npending;
  return s->error;
}

int hh_serve_replay(hh_server *s, const float *queries, const hh_query_spec *specs, uint64_t n,
                    const double *arrival_us, double *latency_ms, uint64_t *out_keys,
                    float *out_dist, uint32_t *out_count, int32_t *out_status) {
  if (!s || !queries || !specs || !arrival_us || !latency_ms)
    return fail(HH_INVALID_ARGUMENT, "null argument");
  std::vector<uint64_t> tickets(n, 0);
  std::atomic<uint64_t> submitted{0};
  std::atomic<int> err{HH_OK};
  std::string err_msg;
  const auto t0 = Clock::now();
  // submit on time from this thread; collect from a second one
  std::thread collector([&] {
    std::vector<uint64_t> keys(s->k_max);
    std::vector<float> dist(s->k_max);
    for (uint64_t i = 0; i < n; i++) {
      while (submitted.load(std::memory_order_acquire) <= i) {
        if (err.load()) return;
        std::this_thread::yield();
      }
      uint32_t cnt = 0;
      int32_t st = 0;
      double lat = 0;
      int done = 0;
      if (hh_serve_poll(s, tickets[i], 1, keys.data(), dist.data(), &cnt, &st, &lat, &done) || !done) {
        err = HH_DEVICE_ERROR;
        return;
      }
      latency_ms[i] = lat;
      if (out_count) out_count[i] = cnt;
      if (out_status) out_status[i] = st;
      if (out_keys) std::memcpy(out_keys + i * s->k_max, keys.data(), cnt * 8);
      if (out_dist) std::memcpy(out_dist + i * s->k_max, dist.data(), cnt * 4);
    }
  });
  for (uint64_t i = 0; i < n; i++) {
    const auto due = t0 + std::chrono::nanoseconds((int64_t)(arrival_us[i] * 1000.0));
    while (Clock::now() < due) {
    }
    int st;
    for (;;) {
      st = hh_serve_submit(s, queries + i * s->dim, specs + i, &tickets[i]);
      if (st != HH_CAPACITY_EXCEEDED) break;
      std::this_thread::yield();  // queue full: admission back-pressure
    }
    if (st) {
      err = st;
      err_msg = hh::last_error();
      break;
    }
    submitted.store(i + 1, std::memory_order_release);
  }
  collector.join();
  if (err.load()) return fail(err.load(), err_msg.empty() ? "replay failed" : err_msg);
  return HH_OK;
}

int hh_serve_close(hh_server *s) {
  if (!s) return HH_OK;
  {
    std::lock_guard<std::mutex> lock(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  if (s->worker.joinable()) s->worker.join();
  (void)hipSetDevice(s->device);
  for (auto &sl : s->slots) {
    (void)hipStreamSynchronize(sl->stream);
    (void)hipHostFree(sl->h_q);
    (void)hipHostFree(sl->h_s);
    (void)hipHostFree(sl->h_k);
    (void)hipHostFree(sl->h_d);
    (void)hipHostFree(sl->h_c);
    (void)hipHostFree(sl->h_st);
    (void)hipFree(sl->d_q);
    (void)hipFree(sl->d_s);
    (void)hipFree(sl->d_k);
    (void)hipFree(sl->d_d);
    (void)hipFree(sl->d_c);
    (void)hipFree(sl->d_st);
    (void)hipEventDestroy(sl->done);
    (void)hipStreamDestroy(sl->stream);
  }
  if (s->d_filters) (void)hipFree(s->d_filters);
  hh::index_serving(s->idx, -1);
  const int err = s->error;
  delete s;
  return err;
}

}  // extern "C"
