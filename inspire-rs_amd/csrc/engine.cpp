// Continuous-batching front end for one server.  The reference answers one
// keyword query per TCP round trip (server.rs:116-158 -> keyword.rs:636-640);
// here callers submit serialized queries from any thread and a dispatcher
// thread packs whatever is pending (up to max_batch, optionally waiting up to
// an arrival window) into one inspire_respond_batch_device call.
//
// Two batch buffer sets let a full batch's upload (own stream) overlap the
// running batch under saturation; below that one batch is in flight.  A
// completer thread waits for each batch's download, stamps its queries done
// and wakes their waiters.  Admission: max_batch is capped so that the
// server's batch buffers and the engine's device buffers fit in the free HBM
// at creation.
#include <hip/hip_runtime.h>

#include <chrono>
#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/inspire_hip.h"
#include "runtime.hpp"

using namespace inspire;
using Clock = std::chrono::steady_clock;

namespace {

enum ReqState { kFree = 0, kQueued, kRunning, kDone, kFailed };

struct Request {
  uint64_t ticket = 0;
  ReqState state = kFree;
  Clock::time_point t_submit, t_done;
  std::string error;
  uint64_t gen = 0;  // state generation the request ran on
};

// The stash of a state generation that was swapped out while some of its
// answers were still to be collected (responses carry the stash of the state
// they were computed on).
struct StashSnapshot {
  uint32_t entry = 0, count = 0;
  std::vector<uint8_t> bytes;
  int64_t unwaited = 0;
};

struct InFlight {
  int set;
  std::vector<int> slots;
};

constexpr int kSets = 2;
constexpr size_t kReserveBytes = (size_t)1 << 30;  // HBM left free at admission

}  // namespace

struct inspire_engine {
  inspire_server_t* srv = nullptr;
  ServerInfo info{};
  int nq = 0, max_batch = 0, capacity = 0;
  double window_ms = 0;
  size_t q_stride = 0;
  hipStream_t up = nullptr, compute = nullptr, down = nullptr;
  uint8_t* h_query = nullptr;         // pinned, capacity x query_bytes
  uint8_t* h_out[kSets] = {};         // pinned, max_batch x blob per set
  uint8_t* d_query[kSets] = {};       // max_batch x q_stride per set
  uint8_t* d_out[kSets] = {};         // max_batch x blob per set
  hipEvent_t ev_up[kSets] = {}, ev_comp[kSets] = {}, ev_down[kSets] = {};
  std::vector<uint8_t> out_store;     // capacity x blob: finished responses
  std::vector<Request> reqs;
  std::vector<int> free_slots;
  std::deque<int> pending;
  std::unordered_map<uint64_t, int> by_ticket;
  std::deque<InFlight> inflight;
  bool set_busy[kSets] = {false, false};
  uint64_t next_ticket = 1;
  int64_t batches = 0, queries = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  bool stop = false;
  bool dispatching = false;  // a batch has left `pending` but is not yet in `inflight`
  bool paused = false;       // a state swap holds new batches back
  uint64_t gen = 0;          // state generation (incremented by each swap)
  std::unordered_map<uint64_t, int64_t> gen_unwaited;      // dispatched, not yet waited, per generation
  std::unordered_map<uint64_t, StashSnapshot> old_stash;  // generations swapped out with answers pending
  std::thread dispatcher, completer;

  void dispatch_loop();
  void complete_loop();
  void fail_batch(const std::vector<int>& slots, const std::string& why);
};

void inspire_engine::fail_batch(const std::vector<int>& slots, const std::string& why) {
  std::lock_guard<std::mutex> lk(mu);
  for (int sl : slots) {
    reqs[sl].state = kFailed;
    reqs[sl].error = why;
    reqs[sl].t_done = Clock::now();
  }
  cv_done.notify_all();
}

void inspire_engine::dispatch_loop() {
  (void)hipSetDevice(info.device);
  for (;;) {
    int set = -1;
    std::vector<int> batch;
    {
      std::unique_lock<std::mutex> lk(mu);
      auto free_set = [&]() {
        for (int k = 0; k < kSets; ++k)
          if (!set_busy[k]) return k;
        return -1;
      };
      // A second batch goes in flight only when a full one is waiting; below
      // that the next batch forms while the running one finishes, so it takes
      // in every query that arrived meanwhile instead of queueing behind a
      // small early batch.
      auto busy = [&]() {
        int n = 0;
        for (int k = 0; k < kSets; ++k) n += set_busy[k] ? 1 : 0;
        return n;
      };
      cv_work.wait(lk, [&] {
        return stop || (!paused && !pending.empty() && free_set() >= 0 &&
                        (busy() == 0 || (int)pending.size() >= max_batch));
      });
      if (stop) return;
      if (window_ms > 0 && (int)pending.size() < max_batch) {
        // arrival window: give the batch until the oldest query has waited window_ms
        const auto deadline =
            reqs[pending.front()].t_submit + std::chrono::duration_cast<Clock::duration>(
                                                 std::chrono::duration<double, std::milli>(window_ms));
        cv_work.wait_until(lk, deadline, [&] { return stop || paused || (int)pending.size() >= max_batch; });
        if (stop) return;
        // a state swap started during the window holds this batch back too
        if (paused) continue;
      }
      set = free_set();
      while (!pending.empty() && (int)batch.size() < max_batch) {
        const int sl = pending.front();
        pending.pop_front();
        reqs[sl].state = kRunning;
        reqs[sl].gen = gen;
        ++gen_unwaited[gen];
        batch.push_back(sl);
      }
      set_busy[set] = true;
      dispatching = true;
    }
    const int count = (int)batch.size();
    std::string why;
    hipError_t e = hipSuccess;
    for (int i = 0; i < count && e == hipSuccess; ++i)
      e = hipMemcpyAsync(d_query[set] + (size_t)i * q_stride, h_query + (size_t)batch[i] * info.query_bytes,
                         info.query_bytes, hipMemcpyHostToDevice, up);
    if (e == hipSuccess) e = hipEventRecord(ev_up[set], up);
    if (e == hipSuccess) e = hipStreamWaitEvent(compute, ev_up[set], 0);
    if (e != hipSuccess) why = std::string("engine upload: ") + hipGetErrorString(e);
    if (why.empty()) {
      const int rc = inspire_respond_batch_device(srv, d_query[set], q_stride, info.query_bytes, nq, count,
                                                  d_out[set], info.blob, compute);
      if (rc != INSPIRE_OK) why = inspire_last_error();
    }
    if (why.empty()) {
      e = hipEventRecord(ev_comp[set], compute);
      if (e == hipSuccess) e = hipStreamWaitEvent(down, ev_comp[set], 0);
      if (e == hipSuccess)
        e = hipMemcpyAsync(h_out[set], d_out[set], (size_t)count * info.blob, hipMemcpyDeviceToHost, down);
      if (e == hipSuccess) e = hipEventRecord(ev_down[set], down);
      if (e != hipSuccess) why = std::string("engine download: ") + hipGetErrorString(e);
    }
    if (!why.empty()) {
      // nothing of this batch is pending on the streams past the failed call
      (void)hipStreamSynchronize(up);
      (void)hipStreamSynchronize(compute);
      (void)hipStreamSynchronize(down);
      fail_batch(batch, why);
      std::lock_guard<std::mutex> lk(mu);
      set_busy[set] = false;
      dispatching = false;
      cv_work.notify_all();
      cv_done.notify_all();
      continue;
    }
    std::lock_guard<std::mutex> lk(mu);
    inflight.push_back(InFlight{set, std::move(batch)});
    dispatching = false;
    ++batches;
    queries += count;
    cv_done.notify_all();
  }
}

void inspire_engine::complete_loop() {
  (void)hipSetDevice(info.device);
  for (;;) {
    InFlight f;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_done.wait(lk, [&] { return !inflight.empty() || stop; });
      if (inflight.empty()) return;  // stopping: destroy waited until nothing was queued or in flight
      f = inflight.front();
    }
    const hipError_t e = hipEventSynchronize(ev_down[f.set]);
    const auto now = Clock::now();
    if (e == hipSuccess)
      for (size_t i = 0; i < f.slots.size(); ++i)
        std::memcpy(out_store.data() + (size_t)f.slots[i] * info.blob, h_out[f.set] + i * info.blob, info.blob);
    std::lock_guard<std::mutex> lk(mu);
    inflight.pop_front();
    for (int sl : f.slots) {
      reqs[sl].t_done = now;
      if (e == hipSuccess) {
        reqs[sl].state = kDone;
      } else {
        reqs[sl].state = kFailed;
        reqs[sl].error = std::string("engine batch: ") + hipGetErrorString(e);
      }
    }
    set_busy[f.set] = false;
    cv_done.notify_all();
    cv_work.notify_all();
  }
}

namespace {

void release(inspire_engine_t* g) {
  for (int k = 0; k < kSets; ++k) {
    if (g->h_out[k]) (void)hipHostFree(g->h_out[k]);
    if (g->d_query[k]) (void)hipFree(g->d_query[k]);
    if (g->d_out[k]) (void)hipFree(g->d_out[k]);
    if (g->ev_up[k]) (void)hipEventDestroy(g->ev_up[k]);
    if (g->ev_comp[k]) (void)hipEventDestroy(g->ev_comp[k]);
    if (g->ev_down[k]) (void)hipEventDestroy(g->ev_down[k]);
  }
  if (g->h_query) (void)hipHostFree(g->h_query);
  if (g->up) (void)hipStreamDestroy(g->up);
  if (g->compute) (void)hipStreamDestroy(g->compute);
  if (g->down) (void)hipStreamDestroy(g->down);
  delete g;
}

#define ENGINE_TRY(call, what)                          \
  do {                                                  \
    hipError_t e_ = (call);                             \
    if (e_ != hipSuccess) {                             \
      release(g);                                       \
      return hip_error(e_, what);                       \
    }                                                   \
  } while (0)

}  // namespace

extern "C" int inspire_engine_create(inspire_server_t* srv, int nq, int max_batch, double window_ms, int capacity,
                                     inspire_engine_t** out) {
  if (!srv || !out || max_batch < 1 || max_batch > 1024 || capacity < 1 || capacity > (1 << 20) || window_ms < 0)
    return set_error(INSPIRE_ERR_INVALID, "bad engine argument");
  *out = nullptr;
  ServerInfo info{};
  int rc = server_info(srv, nq, &info);
  if (rc) return rc;
  if (nq > 2) return set_error(INSPIRE_ERR_INVALID, "hash queries per query");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return set_error(INSPIRE_ERR_HIP, "hipGetDevice failed");
  if (dev != info.device && hipSetDevice(info.device) != hipSuccess)
    return set_error(INSPIRE_ERR_HIP, "hipSetDevice failed");

  // Admission by HBM budget: the largest batch whose server batch buffers and
  // engine buffers fit in the free memory less a reserve.
  const size_t q_stride = (info.query_bytes + 255) / 256 * 256;
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return set_error(INSPIRE_ERR_HIP, "hipMemGetInfo failed");
  const size_t budget = free_b > kReserveBytes ? free_b - kReserveBytes : 0;
  auto need = [&](int b) {
    const size_t server = b > info.batch_max ? server_batch_bytes(srv, b) : 0;
    return server + (size_t)kSets * b * (q_stride + info.blob);
  };
  int fit = max_batch;
  while (fit > 0 && need(fit) > budget) --fit;
  if (fit < 1) return set_error(INSPIRE_ERR_NOMEM, "no batch fits in the free device memory");
  if (fit > info.batch_max) {
    rc = inspire_server_set_max_batch(srv, fit);
    if (rc) return rc;
    rc = server_info(srv, nq, &info);
    if (rc) return rc;
  }

  inspire_engine_t* g = new inspire_engine();
  g->srv = srv;
  g->info = info;
  g->nq = nq;
  g->max_batch = fit;
  g->capacity = capacity;
  g->window_ms = window_ms;
  g->q_stride = q_stride;
  ENGINE_TRY(hipStreamCreateWithFlags(&g->up, hipStreamNonBlocking), "engine stream");
  ENGINE_TRY(hipStreamCreateWithFlags(&g->compute, hipStreamNonBlocking), "engine stream");
  ENGINE_TRY(hipStreamCreateWithFlags(&g->down, hipStreamNonBlocking), "engine stream");
  ENGINE_TRY(hipHostMalloc((void**)&g->h_query, (size_t)capacity * info.query_bytes, hipHostMallocDefault),
             "engine pinned queries");
  for (int k = 0; k < kSets; ++k) {
    ENGINE_TRY(hipHostMalloc((void**)&g->h_out[k], (size_t)fit * info.blob, hipHostMallocDefault),
               "engine pinned responses");
    ENGINE_TRY(hipMalloc((void**)&g->d_query[k], (size_t)fit * q_stride), "engine device queries");
    ENGINE_TRY(hipMalloc((void**)&g->d_out[k], (size_t)fit * info.blob), "engine device responses");
    ENGINE_TRY(hipEventCreateWithFlags(&g->ev_up[k], hipEventDisableTiming), "engine event");
    ENGINE_TRY(hipEventCreateWithFlags(&g->ev_comp[k], hipEventDisableTiming), "engine event");
    ENGINE_TRY(hipEventCreateWithFlags(&g->ev_down[k], hipEventDisableTiming), "engine event");
  }
  g->out_store.resize((size_t)capacity * info.blob);
  g->reqs.resize(capacity);
  g->free_slots.reserve(capacity);
  for (int i = capacity - 1; i >= 0; --i) g->free_slots.push_back(i);
  g->dispatcher = std::thread([g] { g->dispatch_loop(); });
  g->completer = std::thread([g] { g->complete_loop(); });
  *out = g;
  return INSPIRE_OK;
}

extern "C" int inspire_engine_submit(inspire_engine_t* g, const uint8_t* query, size_t len, uint64_t* ticket) {
  if (!g || !query || !ticket) return set_error(INSPIRE_ERR_INVALID, "null argument");
  const auto t0 = Clock::now();
  if (len < g->info.query_bytes) return set_error(INSPIRE_ERR_INVALID, "keyword query truncated");
  if (query[0] != g->nq) return set_error(INSPIRE_ERR_INVALID, "query's hash-query count differs from the engine's");
  int sl = -1;
  {
    std::lock_guard<std::mutex> lk(g->mu);
    if (g->stop) return set_error(INSPIRE_ERR_STATE, "engine stopped");
    if (g->free_slots.empty()) return set_error(INSPIRE_ERR_CAPACITY, "engine queue full");
    sl = g->free_slots.back();
    g->free_slots.pop_back();
  }
  std::memcpy(g->h_query + (size_t)sl * g->info.query_bytes, query, g->info.query_bytes);
  std::lock_guard<std::mutex> lk(g->mu);
  Request& r = g->reqs[sl];
  r.ticket = g->next_ticket++;
  r.state = kQueued;
  r.t_submit = t0;
  r.error.clear();
  g->by_ticket[r.ticket] = sl;
  g->pending.push_back(sl);
  *ticket = r.ticket;
  g->cv_work.notify_all();
  return INSPIRE_OK;
}

extern "C" int inspire_engine_wait(inspire_engine_t* g, uint64_t ticket, uint8_t* resp, size_t cap, size_t* resp_len,
                                   double* latency_ms) {
  if (!g || !resp_len) return set_error(INSPIRE_ERR_INVALID, "null argument");
  std::unique_lock<std::mutex> lk(g->mu);
  auto it = g->by_ticket.find(ticket);
  if (it == g->by_ticket.end()) return set_error(INSPIRE_ERR_INVALID, "unknown ticket");
  const int sl = it->second;
  Request& r = g->reqs[sl];
  // the largest framing this request can get: its generation's stash, or the
  // current one if it has not run yet
  auto framing = [&](const StashSnapshot** old) {
    *old = nullptr;
    if (r.state != kQueued) {
      auto o = g->old_stash.find(r.gen);
      if (o != g->old_stash.end()) *old = &o->second;
    }
    return *old ? response_bytes_with_stash(g->nq, g->info.blob / (size_t)g->nq, (*old)->entry, (*old)->count)
                : inspire_server_response_bytes(g->srv, g->nq);
  };
  const StashSnapshot* old = nullptr;
  *resp_len = std::max(framing(&old), inspire_server_response_bytes(g->srv, g->nq));
  if (!resp || cap < *resp_len) return set_error(INSPIRE_ERR_CAPACITY, "response buffer too small");
  g->cv_done.wait(lk, [&] { return r.state == kDone || r.state == kFailed; });
  int rc = INSPIRE_OK;
  // A state swap may have run while this request waited (queued, it then ran
  // on the new state, whose stash can be larger): size the framing again
  // under the lock that swaps hold, and leave the ticket collectible if the
  // caller's buffer no longer fits.
  const size_t need = framing(&old);
  if (r.state == kDone && cap < need) {
    *resp_len = need;
    return set_error(INSPIRE_ERR_CAPACITY, "response buffer too small for the state the query ran on");
  }
  if (r.state == kFailed) {
    rc = set_error(INSPIRE_ERR_HIP, r.error);
  } else {
    const uint8_t* blobs = g->out_store.data() + (size_t)sl * g->info.blob;
    *resp_len = old ? write_response(g->nq, g->info.blob / (size_t)g->nq, blobs, old->entry, old->count,
                                     old->bytes.data(), resp)
                    : server_write_response(g->srv, g->nq, blobs, resp);
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r.t_done - r.t_submit).count();
  }
  if (r.state == kDone || r.state == kFailed) {
    // this request's generation has one answer fewer to collect
    auto u = g->gen_unwaited.find(r.gen);
    if (u != g->gen_unwaited.end() && --u->second == 0) {
      g->gen_unwaited.erase(u);
      g->old_stash.erase(r.gen);
    }
  }
  g->by_ticket.erase(it);
  r.state = kFree;
  g->free_slots.push_back(sl);
  return rc;
}

extern "C" int inspire_engine_stats(inspire_engine_t* g, int64_t* batches, int64_t* queries, int* max_batch) {
  if (!g) return set_error(INSPIRE_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(g->mu);
  if (batches) *batches = g->batches;
  if (queries) *queries = g->queries;
  if (max_batch) *max_batch = g->max_batch;
  return INSPIRE_OK;
}

// Live DB update under the engine: new batches are held back, the running
// ones finish, the server exchanges its state with `staged`
// (inspire_server_swap_state), and the held queries run on the new state.
extern "C" int inspire_engine_swap_state(inspire_engine_t* g, inspire_server_t* staged) {
  if (!g || !staged) return set_error(INSPIRE_ERR_INVALID, "null argument");
  std::unique_lock<std::mutex> lk(g->mu);
  if (g->paused) return set_error(INSPIRE_ERR_INVALID, "a state swap is already running on this engine");
  g->paused = true;
  g->cv_done.wait(lk, [&] { return g->inflight.empty() && !g->dispatching; });
  StashSnapshot snap;
  server_stash(g->srv, &snap.entry, &snap.count, &snap.bytes);
  int rc = inspire_server_swap_state(g->srv, staged);
  if (rc == INSPIRE_OK) rc = server_info(g->srv, g->nq, &g->info);
  if (rc == INSPIRE_OK) {
    // answers computed on the old state keep its stash
    if (g->gen_unwaited.count(g->gen)) g->old_stash[g->gen] = std::move(snap);
    ++g->gen;
  }
  g->paused = false;
  g->cv_work.notify_all();
  return rc;
}

extern "C" void inspire_engine_destroy(inspire_engine_t* g) {
  if (!g) return;
  {
    std::unique_lock<std::mutex> lk(g->mu);
    // let queued work finish first
    g->cv_done.wait(lk, [&] { return g->pending.empty() && g->inflight.empty() && !g->dispatching; });
    g->stop = true;
    g->cv_work.notify_all();
    g->cv_done.notify_all();
  }
  g->dispatcher.join();
  g->completer.join();
  release(g);
}
