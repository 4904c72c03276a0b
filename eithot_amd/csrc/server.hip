// Serving: continuous batching of in-place transpose requests over
// device-resident tensors.
//
// Requests are queued on submit and admitted in arrival order whenever an
// in-flight slot and enough of the shared scratch pool are free (the
// in-place method needs no output buffer, so scratch is the only memory a
// request consumes).  Admitted requests are enqueued on one of several
// streams, so passes of different requests overlap on the device; a host
// callback behind each request's last pass stamps its completion time.
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

using namespace eithot;

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  int64_t id = 0;
  char* dev = nullptr;
  const Plan* plan = nullptr;
  int status = 0;  // 0 queued, 1 running, 2 done, <0 error
  Clock::time_point t_submit;
  std::atomic<int64_t> done_ns{0};  // completion stamp (ns since t_submit), 0 = pending
  int64_t off = -1;                 // scratch pool offset
  int64_t bytes = 0;
  int stream = 0;
  double latency_ms = 0;
};

// first-fit allocator over [0, cap)
struct Pool {
  int64_t cap = 0;
  std::map<int64_t, int64_t> free_;  // offset -> length
  void init(int64_t c) {
    cap = c;
    free_.clear();
    if (c > 0) free_[0] = c;
  }
  int64_t alloc(int64_t n) {
    if (n <= 0) return 0;
    n = (n + 255) & ~(int64_t)255;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= n) {
        const int64_t off = it->first, len = it->second;
        free_.erase(it);
        if (len > n) free_[off + n] = len - n;
        return off;
      }
    }
    return -1;
  }
  void release(int64_t off, int64_t n) {
    if (n <= 0 || off < 0) return;
    n = (n + 255) & ~(int64_t)255;
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

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
  std::vector<Request*> running;
  double alpha = 0.05;

  ~eithot_server() {
    for (hipStream_t s : streams) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
    if (pool_mem) (void)hipFree(pool_mem);
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
      running[i] = running.back();
      running.pop_back();
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
      int s = 0;
      for (int k = 1; k < (int)streams.size(); ++k)
        if (stream_load[k] < stream_load[s]) s = k;
      r->stream = s;
      ++stream_load[s];
      queue.pop_front();
      hipStream_t st = streams[s];
      int rc = run_plan(dev, *r->plan, r->dev, need > 0 ? pool_mem + off : nullptr, need, st, nullptr);
      if (rc == EITHOT_OK && hipLaunchHostFunc(st, on_done, r) != hipSuccess) rc = EITHOT_ERR_DEVICE;
      if (rc != EITHOT_OK) {
        r->status = rc;
        pool.release(r->off, r->bytes);
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
  const Plan* plan = s->dev->plan_for(dims, perm, order, elem_bytes, 1, s->alpha, &rc);
  if (!plan) return rc ? rc : fail(EITHOT_ERR_INVALID, "planning failed");
  if (plan->scratch_bytes > s->pool.cap)
    return fail(EITHOT_ERR_ALPHA, "request needs %lld scratch bytes, pool holds %lld",
                (long long)plan->scratch_bytes, (long long)s->pool.cap);
  std::lock_guard<std::mutex> lk(s->mu);
  auto r = std::make_unique<Request>();
  r->id = s->next_id++;
  r->dev = (char*)dev;
  r->plan = plan;
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
    }
    int rc = eithot_server_poll(s, 1);
    if (rc < 0) return rc;
  }
}

void eithot_server_destroy(eithot_server* s) { delete s; }

}  // extern "C"
