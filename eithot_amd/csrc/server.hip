// Serving: continuous batching of in-place transpose requests over
// device-resident tensors.
//
// Requests are queued on submit and admitted in arrival order whenever an
// in-flight slot and enough of the shared scratch pool are free (the
// in-place method needs no output buffer, so scratch is the only memory a
// request consumes).  Admitted requests are enqueued on one of several
// streams, so passes of different requests overlap on the device; a host
// callback behind each request's last pass stamps its completion time.
//
// Requests on overlapping bytes (a tensor transposed and later transposed
// back) run in submission order: admission is FIFO, and an admitted request
// whose bytes overlap a running one joins that request's stream and waits on
// the completion event of every other running request it overlaps.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/eithot.h"
#include "kernels.hpp"
#include "plan.hpp"
#include "runtime.hpp"
#include "scratch_pool.hpp"

using namespace eithot;

namespace {

using Clock = std::chrono::steady_clock;

// workgroup floor per pass of a served request (a pass launched alone
// spreads over at least 1024 workgroups; requests share the GPU)
constexpr int64_t kServedMinGroups = 128;

struct Request {
  int64_t id = 0;
  char* dev = nullptr;
  uint64_t nbytes = 0;   // the tensor occupies [dev, dev + nbytes)
  PlanPtr plan;          // held until the request completes
  hipEvent_t done_ev = nullptr;  // recorded after the request's last pass
  int status = 0;  // 0 queued, 1 running, 2 done, <0 error
  Clock::time_point t_submit;
  std::atomic<int64_t> done_ns{0};  // completion stamp (ns since t_submit), 0 = pending
  int64_t off = -1;                 // scratch pool offset
  int64_t bytes = 0;
  int stream = 0;
  double latency_ms = 0;
};

using Pool = ScratchPool;

void on_done(void* p) {
  Request* r = static_cast<Request*>(p);
  const int64_t ns =
      std::chrono::duration_cast<std::chrono::nanoseconds>(Clock::now() - r->t_submit).count();
  r->done_ns.store(ns > 0 ? ns : 1, std::memory_order_release);
}

}  // namespace

struct eithot_server {
  Device* dev = nullptr;
  char* pool_mem = nullptr;
  Pool pool;
  int max_inflight = 1;
  std::vector<hipStream_t> streams;
  std::vector<int> stream_load;
  std::mutex mu;
  int64_t next_id = 0;
  std::map<int64_t, std::unique_ptr<Request>> reqs;
  std::deque<Request*> queue;
  std::vector<Request*> running;  // in admission order
  std::vector<hipEvent_t> free_events;
  double alpha = 0.05;

  ~eithot_server() {
    for (hipStream_t s : streams) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
    for (Request* r : running)
      if (r->done_ev) free_events.push_back(r->done_ev);
    for (hipEvent_t e : free_events) (void)hipEventDestroy(e);
    if (pool_mem) (void)hipFree(pool_mem);
  }

  hipEvent_t take_event() {
    if (!free_events.empty()) {
      hipEvent_t e = free_events.back();
      free_events.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
    return e;
  }

  static bool overlaps(const Request* a, const Request* b) {
    return a->dev < b->dev + b->nbytes && b->dev < a->dev + a->nbytes;
  }

  // retire finished requests; returns how many completed
  int retire() {
    int n = 0;
    for (size_t i = 0; i < running.size();) {
      Request* r = running[i];
      const int64_t ns = r->done_ns.load(std::memory_order_acquire);
      if (!ns) {
        ++i;
        continue;
      }
      r->latency_ms = ns / 1e6;
      r->status = 2;
      pool.release(r->off, r->bytes);
      --stream_load[r->stream];
      free_events.push_back(r->done_ev);
      r->done_ev = nullptr;
      r->plan.reset();
      running.erase(running.begin() + (long)i);  // keeps admission order
      ++n;
    }
    return n;
  }

  void admit() {
    while (!queue.empty() && (int)running.size() < max_inflight) {
      Request* r = queue.front();
      const int64_t need = r->plan->scratch_bytes;
      int64_t off = 0;
      if (need > 0) {
        off = pool.alloc(need);
        if (off < 0) break;  // FIFO: wait for scratch to free up
      }
      r->off = off;
      r->bytes = need;
      r->done_ev = take_event();
      if (!r->done_ev) {
        pool.release(off, need);
        break;  // retried on the next poll
      }
      // the least-loaded stream, unless the request's bytes overlap a
      // running request: then the stream of the last such request
      int s = 0;
      for (int k = 1; k < (int)streams.size(); ++k)
        if (stream_load[k] < stream_load[s]) s = k;
      for (const Request* o : running)
        if (overlaps(r, o)) s = o->stream;
      r->stream = s;
      ++stream_load[s];
      queue.pop_front();
      hipStream_t st = streams[s];
      int rc = EITHOT_OK;
      for (const Request* o : running)
        if (overlaps(r, o) && o->stream != s && hipStreamWaitEvent(st, o->done_ev, 0) != hipSuccess)
          rc = EITHOT_ERR_DEVICE;
      if (rc == EITHOT_OK) {
        // launches are exclusive with other threads' table eviction
        std::lock_guard<std::mutex> lk(dev->enqueue_mu);
        TableAccount acct(&dev->cached_tables);
        // requests share the GPU, so a pass need not fill it alone: at most
        // 128 workgroups' floor per pass (2 % more tensors/s at saturation
        // than the single-launch 1024)
        rc = run_plan(dev, *r->plan, r->dev, need > 0 ? pool_mem + off : nullptr, need, st,
                      nullptr, kServedMinGroups);
      }
      if (rc == EITHOT_OK && hipEventRecord(r->done_ev, st) != hipSuccess) rc = EITHOT_ERR_DEVICE;
      if (rc == EITHOT_OK && hipLaunchHostFunc(st, on_done, r) != hipSuccess) rc = EITHOT_ERR_DEVICE;
      if (rc != EITHOT_OK) {
        // passes already enqueued may still run: wait for them before the
        // scratch and the event are reused
        (void)hipStreamSynchronize(st);
        r->status = rc;
        pool.release(r->off, r->bytes);
        free_events.push_back(r->done_ev);
        r->done_ev = nullptr;
        r->plan.reset();
        --stream_load[s];
        continue;
      }
      r->status = 1;
      running.push_back(r);
    }
  }
};

extern "C" {

eithot_server* eithot_server_create(int64_t scratch_pool_bytes, int max_inflight, int nstreams) {
  int err = 0;
  Device* d = device_ctx(&err);
  if (!d || scratch_pool_bytes < 0 || max_inflight < 1 || nstreams < 1 || nstreams > 64)
    return nullptr;
  auto* s = new eithot_server;
  s->dev = d;
  s->max_inflight = max_inflight;
  if (scratch_pool_bytes > 0 &&
      hipMalloc((void**)&s->pool_mem, (size_t)scratch_pool_bytes) != hipSuccess) {
    (void)hipGetLastError();
    delete s;
    return nullptr;
  }
  s->pool.init(scratch_pool_bytes);
  for (int k = 0; k < nstreams; ++k) {
    hipStream_t st;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
      delete s;
      return nullptr;
    }
    s->streams.push_back(st);
    s->stream_load.push_back(0);
  }
  return s;
}

int64_t eithot_server_submit(eithot_server* s, void* dev, const int64_t* dims, const int* perm,
                             int order, size_t elem_bytes) {
  if (!s) return fail(EITHOT_ERR_INVALID, "null server");
  int rc = validate(dev, dims, perm, order, elem_bytes);
  if (rc) return rc;
  PlanPtr plan = s->dev->plan_for(dims, perm, order, elem_bytes, 1, s->alpha, &rc);
  if (!plan) return rc ? rc : fail(EITHOT_ERR_INVALID, "planning failed");
  if (!s->pool.fits(plan->scratch_bytes))
    return fail(EITHOT_ERR_ALPHA, "request needs %lld scratch bytes, pool holds %lld",
                (long long)Pool::granule(plan->scratch_bytes), (long long)s->pool.cap);
  std::lock_guard<std::mutex> lk(s->mu);
  auto r = std::make_unique<Request>();
  r->id = s->next_id++;
  r->dev = (char*)dev;
  r->nbytes = plan->tensor_bytes;
  r->plan = std::move(plan);
  r->t_submit = Clock::now();
  Request* raw = r.get();
  s->reqs[raw->id] = std::move(r);
  s->queue.push_back(raw);
  s->admit();
  return raw->id;
}

int eithot_server_poll(eithot_server* s, int wait) {
  if (!s) return fail(EITHOT_ERR_INVALID, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  int n = s->retire();
  s->admit();
  if (wait && n == 0 && !s->running.empty()) {
    // block until the oldest in-flight request's stream drains past it
    Request* r = s->running.front();
    while (!r->done_ns.load(std::memory_order_acquire))
      std::this_thread::sleep_for(std::chrono::microseconds(10));
    n = s->retire();
    s->admit();
  }
  return n;
}

int eithot_server_status(eithot_server* s, int64_t id, double* latency_ms) {
  if (!s) return fail(EITHOT_ERR_INVALID, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return fail(EITHOT_ERR_INVALID, "unknown request");
  if (latency_ms) *latency_ms = it->second->latency_ms;
  return it->second->status;
}

int eithot_server_drain(eithot_server* s) {
  if (!s) return fail(EITHOT_ERR_INVALID, "null server");
  for (;;) {
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->retire();
      s->admit();
      if (s->queue.empty() && s->running.empty()) return EITHOT_OK;
      // nothing runs and the head still cannot start: it never will
      if (s->running.empty())
        return fail(EITHOT_ERR_DEVICE, "request %lld cannot be admitted",
                    (long long)s->queue.front()->id);
    }
    int rc = eithot_server_poll(s, 1);
    if (rc < 0) return rc;
  }
}

void eithot_server_destroy(eithot_server* s) { delete s; }

}  // extern "C"
