// C-ABI of libeithot.so: device contexts, scratch arena, plan cache and the
// synchronous / stream-ordered transpose entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/eithot.h"
#include "kernels.hpp"
#include "plan.hpp"
#include "runtime.hpp"

namespace eithot {

namespace {
thread_local std::string g_err;
thread_local size_t g_peak = 0;
thread_local std::vector<float> g_pass_ms;
int g_timing = 0;
}  // namespace

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

void set_peak(size_t b) { g_peak = b; }

Device* device_ctx(int* err) {
  static std::mutex mu;
  // never destroyed: device memory is not freed from static destructors at
  // process exit, when the HIP runtime may already be gone
  static auto* devs = new std::vector<std::unique_ptr<Device>>();
  int id = 0;
  if (hipGetDevice(&id) != hipSuccess) {
    if (err) *err = fail(EITHOT_ERR_NO_DEVICE, "no HIP device available");
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(mu);
  if ((int)devs->size() <= id) devs->resize(id + 1);
  if (!(*devs)[id]) {
    auto d = std::make_unique<Device>();
    d->id = id;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, id) != hipSuccess) {
      if (err) *err = fail(EITHOT_ERR_NO_DEVICE, "hipGetDeviceProperties failed");
      return nullptr;
    }
    d->prof.cus = prop.multiProcessorCount;
    // one workgroup may use the whole LDS of a CU on gfx950 (160 KiB)
    int64_t lds = (int64_t)prop.maxSharedMemoryPerMultiProcessor;
    if (lds <= 0) lds = (int64_t)prop.sharedMemPerBlock;
    d->prof.lds_max = std::min<int64_t>(lds, 160 * 1024);
    d->prof.lds_target = std::min<int64_t>(64 * 1024, d->prof.lds_max);
    hipError_t e = configure_kernels(d->prof.lds_max);
    if (e != hipSuccess) {
      // fall back to the per-block limit the runtime reports
      d->prof.lds_max = (int64_t)prop.sharedMemPerBlock;
      d->prof.lds_target = std::min<int64_t>(d->prof.lds_target, d->prof.lds_max);
      (void)hipGetLastError();
    }
    d->cfg.lds_target = d->prof.lds_target;
    (*devs)[id] = std::move(d);
  }
  return (*devs)[id].get();
}

DeviceProfile Device::profile() const {
  DeviceProfile p = prof;
  tuning_apply(tuning_get(), p);
  return p;
}

LaunchCfg Device::launch_cfg() const {
  const Tuning t = tuning_get();
  LaunchCfg c = cfg;
  c.lds_target = std::min<int64_t>(t.lds_target, prof.lds_max);
  c.threads = (int)t.threads;
  c.threads_big = (int)t.threads_big;
  c.xcd_remap = (int)t.xcd_remap;
  c.cycle_variant = (int)t.cycle_variant;
  c.segtile_threads = (int)t.segtile_threads;
  c.lds_threads_auto = (int)t.lds_threads_auto;
  c.segtile_nt = (int)t.segtile_nt;
  c.cycle_u_depth = (int)t.cycle_u_depth;
  c.min_groups = t.min_groups;
  return c;
}

std::string plan_key(const int64_t* dims, const int* perm1, int rank, size_t elem, int n_sub,
                     double alpha) {
  std::ostringstream k;
  k << tuning_get().epoch << "#";
  for (int i = 0; i < rank; ++i) k << dims[i] << ",";
  k << "|";
  for (int i = 0; i < rank; ++i) k << perm1[i] << ",";
  k << "|" << elem << "|" << n_sub << "|" << alpha;
  return k.str();
}

PlanPtr Device::plan_for(const int64_t* dims, const int* perm1, int rank, size_t elem, int n_sub,
                         double alpha, int* rc) {
  const DeviceProfile pr = profile();
  const std::string key = plan_key(dims, perm1, rank, elem, n_sub, alpha);
  {
    std::lock_guard<std::mutex> lk(mu);
    if (PlanPtr p = plans.get(key)) {
      *rc = EITHOT_OK;
      return p;
    }
  }
  // plan without the lock (a large shape takes seconds); a plan another
  // thread cached meanwhile wins
  auto p = std::make_shared<Plan>();
  *rc = make_plan(dims, perm1, rank, elem, n_sub, alpha, pr, pr.reorder, *p);
  if (*rc != EITHOT_OK) return nullptr;
  std::vector<PlanPtr> evicted;  // destroyed after the lock is released
  PlanPtr out;
  {
    std::lock_guard<std::mutex> lk(mu);
    if ((out = plans.get(key))) return out;
    out = p;
    evicted = plans.put(key, out, (uint64_t)std::max<int64_t>(p->table_bytes, 0));
  }
  return out;
}

int64_t Device::held_bytes() { return arena_bytes + cached_tables.load(std::memory_order_relaxed); }

void Device::trim_locked(int64_t budget, const Plan* keep) {
  // what this call will hold: the resident tables, those of `keep` still to
  // be uploaded, and the arena it ends with (the current one, or one grown
  // to its need); the plans are walked only when that exceeds the budget
  const int64_t need = keep ? keep->scratch_bytes : 0;
  const int64_t keep_res = keep ? keep->device_table_bytes(id) : 0;
  const int64_t keep_tables = keep ? std::max(keep->table_bytes, keep_res) : 0;
  int64_t others = cached_tables.load(std::memory_order_relaxed) - keep_res;
  auto held_after = [&] { return std::max(arena_bytes, need) + keep_tables + others; };
  if (held_after() <= budget) return;
  std::vector<PlanPtr> order;  // most recently used first
  {
    std::lock_guard<std::mutex> lk(mu);
    plans.for_each([&](const std::string&, const PlanPtr& p) { order.push_back(p); });
  }
  // an arena larger than this call needs is given back first
  if (arena && arena_bytes > need) {
    if (arena_ready) (void)hipEventSynchronize(arena_ready);
    (void)hipFree(arena);
    arena = nullptr;
    arena_bytes = 0;
  }
  // then the tables of other plans, least recently used first
  for (auto it = order.rbegin(); it != order.rend(); ++it) {
    if (held_after() <= budget) break;
    if (it->get() == keep) continue;
    const int64_t b = (*it)->device_table_bytes(id);
    if (!b) continue;
    (*it)->drop_device_tables(id);
    others -= b;
  }
}

int Device::reserve_arena_locked(int64_t bytes, hipStream_t stream) {
  if (bytes <= 0) return EITHOT_OK;
  if (arena_ready) {  // earlier async users of the arena must finish first
    if (hipStreamWaitEvent(stream, arena_ready, 0) != hipSuccess)
      return fail(EITHOT_ERR_DEVICE, "hipStreamWaitEvent failed");
  }
  if (arena_bytes >= bytes) return EITHOT_OK;
  if (arena) {
    // every earlier user is done once its event has completed (users
    // enqueue under the same lock)
    if (arena_ready) (void)hipEventSynchronize(arena_ready);
    if (hipFree(arena) != hipSuccess) (void)hipGetLastError();
    arena = nullptr;
    arena_bytes = 0;
  }
  if (hipMalloc((void**)&arena, (size_t)bytes) != hipSuccess) {
    (void)hipGetLastError();
    arena = nullptr;
    return fail(EITHOT_ERR_NOMEM, "scratch allocation of %lld bytes failed", (long long)bytes);
  }
  arena_bytes = bytes;
  return EITHOT_OK;
}

int Device::release_arena_after_locked(hipStream_t stream) {
  if (!arena_ready && hipEventCreateWithFlags(&arena_ready, hipEventDisableTiming) != hipSuccess)
    return fail(EITHOT_ERR_DEVICE, "hipEventCreate failed");
  if (hipEventRecord(arena_ready, stream) != hipSuccess)
    return fail(EITHOT_ERR_DEVICE, "hipEventRecord failed");
  return EITHOT_OK;
}

// Timing events of this thread on each device, reused from call to call
// (creating and destroying them per call costs more than the launches).
std::vector<hipEvent_t>& timing_events(size_t n) {
  thread_local std::map<int, std::vector<hipEvent_t>> pools;
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::vector<hipEvent_t>& v = pools[dev];
  while (v.size() < n) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) break;
    v.push_back(e);
  }
  return v;
}

// The segmented tile's block order (xcd_remap 1: each XCD takes a
// contiguous range of matrices; 0: the hardware's round-robin) wins by 3-7 %
// one way or the other depending on the shape, with no static rule found.
// A plan's first call runs each segmented-tile pass with order 1 untimed
// (it may be cold); its second runs order 1 and its third order 0, both
// between timing events; a later call whose events have completed keeps the
// faster order for good.  The calls do their real work, so the trial costs
// no extra transposes.
namespace {
struct SegOrderTrial {
  int stage = -1;  // -1: untimed first call, 0: time order 1 next, 1: time order 0 next,
                   // 2: both timed, 3: decided
  int choice = 1;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  ~SegOrderTrial() {
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};
struct PlanLaunchState {
  std::vector<SegOrderTrial> seg;  // one per pass (used by segmented tiles)
};
std::mutex g_launch_state_mu;

// Block order of each pass of this call and the events to record around
// it ({nullptr, nullptr}: none).  A stream under graph capture takes the
// decided order (else 1) and neither records nor queries events.
void seg_orders(const Plan& plan, int base_remap, hipStream_t stream, std::vector<int>& remap,
                std::vector<std::pair<hipEvent_t, hipEvent_t>>& rec) {
  remap.assign(plan.ops.size(), base_remap);
  rec.assign(plan.ops.size(), {nullptr, nullptr});
  if (base_remap != 1) return;  // a forced order is kept
  bool any = false;
  for (const Op& o : plan.ops) any = any || o.kind == OP_SEGTILE;
  if (!any) return;
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(stream, &cap) != hipSuccess) {
    (void)hipGetLastError();
    cap = hipStreamCaptureStatusActive;  // unknown: treat as captured
  }
  std::lock_guard<std::mutex> lk(g_launch_state_mu);
  if (cap != hipStreamCaptureStatusNone) {
    auto st = std::static_pointer_cast<PlanLaunchState>(plan.launch_state);
    for (size_t i = 0; i < plan.ops.size(); ++i)
      if (plan.ops[i].kind == OP_SEGTILE && st && st->seg[i].stage == 3) remap[i] = st->seg[i].choice;
    return;
  }
  auto st = std::static_pointer_cast<PlanLaunchState>(plan.launch_state);
  if (!st) {
    st = std::make_shared<PlanLaunchState>();
    st->seg.resize(plan.ops.size());
    plan.launch_state = st;
  }
  for (size_t i = 0; i < plan.ops.size(); ++i) {
    if (plan.ops[i].kind != OP_SEGTILE) continue;
    SegOrderTrial& t = st->seg[i];
    if (t.stage == -1) {
      remap[i] = 1;
      t.stage = 0;
    } else if (t.stage == 0) {
      bool ok = true;
      for (hipEvent_t& e : t.ev)
        if (!e && hipEventCreate(&e) != hipSuccess) ok = false;
      if (!ok) {
        (void)hipGetLastError();
        t.stage = 3;  // no events: keep order 1
      } else {
        rec[i] = {t.ev[0], t.ev[1]};
        t.stage = 1;
      }
      remap[i] = 1;
    } else if (t.stage == 1) {
      rec[i] = {t.ev[2], t.ev[3]};
      remap[i] = 0;
      t.stage = 2;
    } else if (t.stage == 2) {
      float a = 0, b = 0;
      if (hipEventElapsedTime(&a, t.ev[0], t.ev[1]) == hipSuccess &&
          hipEventElapsedTime(&b, t.ev[2], t.ev[3]) == hipSuccess) {
        t.choice = a <= b ? 1 : 0;
        t.stage = 3;
      } else {
        (void)hipGetLastError();  // not complete yet: ask again on a later call
      }
      remap[i] = t.stage == 3 ? t.choice : 1;
    } else {
      remap[i] = t.choice;
    }
  }
}
}  // namespace

int run_plan(Device* d, const Plan& plan, char* data, char* scratch, int64_t scratch_bytes,
             hipStream_t stream, std::vector<float>* times, int64_t min_groups_cap) {
  std::vector<hipEvent_t> ev;
  if (times) {
    const std::vector<hipEvent_t>& pool = timing_events(plan.ops.size() + 1);
    if (pool.size() >= plan.ops.size() + 1 && hipEventRecord(pool[0], stream) == hipSuccess) {
      ev.assign(pool.begin(), pool.begin() + (long)(plan.ops.size() + 1));
    } else {
      times->clear();
      times = nullptr;  // no events to time with
    }
  }
  int rc = EITHOT_OK;
  LaunchCfg cfg = d->launch_cfg();
  if (min_groups_cap > 0) cfg.min_groups = std::min(cfg.min_groups, min_groups_cap);
  std::vector<int> remap;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> rec;
  seg_orders(plan, cfg.xcd_remap, stream, remap, rec);
  for (size_t i = 0; i < plan.ops.size(); ++i) {
    LaunchCfg c = cfg;
    c.xcd_remap = remap[i];
    if (rec[i].first && hipEventRecord(rec[i].first, stream) != hipSuccess) (void)hipGetLastError();
    hipError_t e = launch_op(plan.ops[i], data, scratch, scratch_bytes, c, stream);
    if (e == hipSuccess && rec[i].second && hipEventRecord(rec[i].second, stream) != hipSuccess)
      (void)hipGetLastError();
    if (e != hipSuccess) {
      rc = fail(EITHOT_ERR_DEVICE, "pass %zu launch failed: %s", i, hipGetErrorString(e));
      break;
    }
    if (times && hipEventRecord(ev[i + 1], stream) != hipSuccess) {
      rc = fail(EITHOT_ERR_DEVICE, "pass %zu timing event failed", i);
      break;
    }
  }
  if (times) {
    times->clear();
    if (rc == EITHOT_OK) {
      if (hipEventSynchronize(ev.back()) != hipSuccess)
        return fail(EITHOT_ERR_DEVICE, "pass timing synchronisation failed");
      for (size_t i = 0; i < plan.ops.size(); ++i) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, ev[i], ev[i + 1]) != hipSuccess) {
          times->clear();
          return fail(EITHOT_ERR_DEVICE, "pass %zu elapsed time unavailable", i);
        }
        times->push_back(ms);
      }
    }
  }
  return rc;
}

int validate(const void* dev, const int64_t* dims, const int* perm, int order, size_t elem) {
  if (!dev) return fail(EITHOT_ERR_INVALID, "null tensor pointer");
  if (!dims || !perm || order < 1 || order > 64)
    return fail(EITHOT_ERR_INVALID, "order must be in 1..64 with dims and perm given");
  if (elem != 1 && elem != 2 && elem != 4 && elem != 8)
    return fail(EITHOT_ERR_ELEM, "elem_bytes must be 1, 2, 4 or 8 (got %zu)", elem);
  std::vector<int> seen(order + 1, 0);
  for (int i = 0; i < order; ++i) {
    if (dims[i] < 1) return fail(EITHOT_ERR_INVALID, "dims[%d] = %lld < 1", i, (long long)dims[i]);
    if (perm[i] < 1 || perm[i] > order || seen[perm[i]])
      return fail(EITHOT_ERR_INVALID, "perm is not a 1-based permutation of 1..%d", order);
    seen[perm[i]] = 1;
  }
  if (((uintptr_t)dev) % elem) return fail(EITHOT_ERR_INVALID, "tensor pointer misaligned");
  return EITHOT_OK;
}

static int transpose_impl(void* dev, const int64_t* dims, const int* perm, int order, size_t elem,
                          int n_sub, double alpha, hipStream_t stream, bool sync) {
  g_err.clear();
  g_peak = 0;
  int rc = validate(dev, dims, perm, order, elem);
  if (rc) return rc;
  if (!(alpha > 0)) return fail(EITHOT_ERR_INVALID, "alpha must be > 0");
  Device* d = device_ctx(&rc);
  if (!d) return rc;
  hipPointerAttribute_t attr;
  if (hipPointerGetAttributes(&attr, dev) != hipSuccess) {
    (void)hipGetLastError();
    return fail(EITHOT_ERR_INVALID, "pointer is not device memory");
  }
  PlanPtr plan = d->plan_for(dims, perm, order, elem, n_sub, alpha, &rc);
  if (!plan) {
    if (!g_err.size()) fail(rc, "planning failed (%d)", rc);
    return rc;
  }
  {
    std::lock_guard<std::mutex> lk(d->enqueue_mu);
    // the library's device bytes stay within alpha of this tensor, or this
    // plan's own need when one footprint exceeds it, or the holding floor
    // (which spares calls on small tensors from evicting and re-uploading
    // the tables of the others)
    const int64_t own = plan->scratch_bytes + plan->table_bytes;
    const int64_t budget = std::max<int64_t>(
        std::max<int64_t>((int64_t)(alpha * (double)plan->tensor_bytes), own),
        tuning_get().hold_floor);
    d->trim_locked(budget, plan.get());
    if (plan->scratch_bytes > 0) {
      rc = d->reserve_arena_locked(plan->scratch_bytes, stream);
      if (rc) return rc;
    }
    std::vector<float>* times = g_timing ? &g_pass_ms : nullptr;
    TableAccount acct(&d->cached_tables);
    rc = run_plan(d, *plan, (char*)dev, d->arena, plan->scratch_bytes, stream, times);
    if (rc) return rc;
    if (plan->scratch_bytes > 0) {
      rc = d->release_arena_after_locked(stream);
      if (rc) return rc;
    }
  }
  g_peak = (size_t)d->held_bytes();
  if (sync) {
    hipError_t e = hipStreamSynchronize(stream);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) return fail(EITHOT_ERR_DEVICE, "device error: %s", hipGetErrorString(e));
  }
  return EITHOT_OK;
}

}  // namespace eithot

using namespace eithot;

extern "C" {

int eithot_transpose(void* dev, const int64_t* dims, const int* perm, int order,
                     size_t elem_bytes, int n_sub, double alpha) {
  return transpose_impl(dev, dims, perm, order, elem_bytes, n_sub, alpha, nullptr, true);
}

int eithot_transpose_async(void* dev, const int64_t* dims, const int* perm, int order,
                           size_t elem_bytes, int n_sub, double alpha, void* stream) {
  return transpose_impl(dev, dims, perm, order, elem_bytes, n_sub, alpha, (hipStream_t)stream,
                        false);
}

int eithot_inplace_transpose(void* data, int source, int rank, const int* dim,
                             const int* permutation, size_t sizeofType, int num_block,
                             double large_ratio) {
  if (rank < 1 || rank > 64 || !dim || !permutation || source < 0)
    return fail(EITHOT_ERR_INVALID, "invalid rank, dims, permutation or source");
  std::vector<int64_t> d(rank);
  std::vector<int> p(rank);
  for (int i = 0; i < rank; ++i) {
    d[i] = dim[i];
    p[i] = permutation[i] + 1;
  }
  char* base = (char*)data + (size_t)source * sizeofType;
  return eithot_transpose(base, d.data(), p.data(), rank, sizeofType, num_block, large_ratio);
}

size_t eithot_last_peak_extra_bytes(void) { return g_peak; }

const char* eithot_last_error(void) { return g_err.c_str(); }

void eithot_release(void) {
  int err = 0;
  Device* d = device_ctx(&err);
  if (!d) return;
  std::vector<PlanPtr> dropped;
  std::lock_guard<std::mutex> lk(d->enqueue_mu);
  (void)hipDeviceSynchronize();
  if (d->arena) (void)hipFree(d->arena);
  d->arena = nullptr;
  d->arena_bytes = 0;
  {
    std::lock_guard<std::mutex> lk2(d->mu);
    dropped = d->plans.clear();
  }
  // device tables are freed now; a plan still held elsewhere (a queued
  // server request) uploads its tables again when it runs
  for (const PlanPtr& p : dropped) p->drop_device_tables(d->id);
}

int eithot_estimate(const int64_t* dims, const int* perm, int order, size_t elem_bytes,
                    double alpha, int* n_sub_out, double* alpha_out, size_t* extra_bytes_out) {
  DeviceProfile prof;
  tuning_apply(tuning_get(), prof);
  int rc = estimate_native(dims, perm, order, elem_bytes, alpha, prof, n_sub_out, alpha_out,
                           extra_bytes_out);
  if (rc) return fail(rc, "estimate failed (%d)", rc);
  return EITHOT_OK;
}

int eithot_estimate_reference(const int64_t* dims, const int* perm, int order, size_t elem_bytes,
                              double alpha, const eithot_ref_profile* prof, double* init_bytes,
                              double* new_ratio, double* final_bytes, double* final_ratio,
                              int* nb, double* alpha_out, int* flags) {
  RefProfile p;
  if (prof) {
    p.smem_per_block = prof->smem_per_block;
    p.sms = prof->sms;
    p.max_threads_per_sm = prof->max_threads_per_sm;
    p.padding_blocks_per_sm = prof->padding_blocks_per_sm;
  }
  RefEstimate e;
  int rc = estimate_reference(dims, perm, order, elem_bytes, alpha, p, e);
  if (rc) return fail(rc, "reference estimate failed (%d)", rc);
  if (init_bytes) *init_bytes = e.init_bytes;
  if (new_ratio) *new_ratio = e.new_ratio;
  if (final_bytes) *final_bytes = e.final_bytes;
  if (final_ratio) *final_ratio = e.final_ratio;
  if (nb) *nb = e.nb;
  if (alpha_out) *alpha_out = e.alpha;
  if (flags) *flags = (e.printed_new_ratio ? 1 : 0) | (e.decomposed ? 2 : 0);
  return EITHOT_OK;
}

int eithot_decompose(const int64_t* dims, const int* perm, int order, int reorder,
                     int* stage_rank, int* stage_perm, int64_t* stage_dims, int max_stages) {
  std::vector<Stage> st;
  if (!dims || !perm || !decompose(dims, perm, order, reorder, st))
    return fail(EITHOT_ERR_INVALID, "invalid permutation");
  const int n = (int)st.size();
  for (int s = 0; s < n && s < max_stages; ++s) {
    if (stage_rank) stage_rank[s] = st[s].rank;
    for (int i = 0; i < 4; ++i) {
      if (stage_perm) stage_perm[4 * s + i] = st[s].perm[i];
      if (stage_dims) stage_dims[4 * s + i] = st[s].dims[i];
    }
  }
  return n;
}

int eithot_plan_describe(const int64_t* dims, const int* perm, int order, size_t elem_bytes,
                         int n_sub, double alpha, int64_t lds_max, eithot_op_desc* ops,
                         int max_ops, int64_t* scratch_bytes) {
  DeviceProfile prof;
  if (lds_max > 0) prof.lds_max = lds_max;
  tuning_apply(tuning_get(), prof);
  Plan plan;
  int rc = make_plan(dims, perm, order, elem_bytes, n_sub, alpha, prof, prof.reorder, plan);
  if (rc) return fail(rc, "planning failed (%d)", rc);
  if (scratch_bytes) *scratch_bytes = plan.scratch_bytes + plan.table_bytes;
  const int n = (int)plan.ops.size();
  for (int i = 0; i < n && i < max_ops && ops; ++i) {
    const Op& o = plan.ops[i];
    eithot_op_desc& x = ops[i];
    x.kind = o.kind;
    x.map = o.map;
    x.staged = o.staged;
    x.stage = o.stage;
    x.base = o.base;
    x.B = o.B;
    x.R = o.R;
    x.C = o.C;
    x.U = o.U;
    x.us = o.us;
    x.W = o.W;
    x.m = o.m;
    x.n = o.n;
    x.c = o.c;
    x.a = o.a;
    x.b = o.b;
    x.k = o.k;
    x.binv = o.binv;
  }
  return n;
}

int eithot_plan_partition(const int64_t* dims, const int* perm, int order, size_t elem_bytes,
                          int n_sub, double alpha, int* n_sub_used, int* dim, int64_t* part_len,
                          int64_t* last_len) {
  DeviceProfile prof;
  tuning_apply(tuning_get(), prof);
  Plan plan;
  int rc = make_plan(dims, perm, order, elem_bytes, n_sub, alpha, prof, prof.reorder, plan);
  if (rc) return fail(rc, "planning failed (%d)", rc);
  if (n_sub_used) *n_sub_used = plan.n_sub;
  if (dim) *dim = plan.partition_dim;
  if (part_len) *part_len = plan.part_len;
  if (last_len) *last_len = plan.last_len;
  return EITHOT_OK;
}

int64_t eithot_plan_block_table(const int64_t* dims, const int* perm, int order,
                                size_t elem_bytes, int n_sub, double alpha, int64_t lds_max,
                                int op_index, uint32_t* src, int64_t max_cells) {
  DeviceProfile prof;
  if (lds_max > 0) prof.lds_max = lds_max;
  tuning_apply(tuning_get(), prof);
  Plan plan;
  int rc = make_plan(dims, perm, order, elem_bytes, n_sub, alpha, prof, prof.reorder, plan);
  if (rc) return fail(rc, "planning failed (%d)", rc);
  if (op_index < 0 || op_index >= (int)plan.ops.size() || !plan.ops[op_index].blk)
    return fail(EITHOT_ERR_INVALID, "pass %d is not a fused block permutation", op_index);
  const BlockTable& t = *plan.ops[op_index].blk;
  for (int64_t i = 0; i < t.cells && i < max_cells && src; ++i) src[i] = t.src[i];
  return t.cells;
}

int eithot_plan_cycle_maps(const int64_t* dims, const int* perm, int order, size_t elem_bytes,
                           int n_sub, double alpha, int64_t lds_max, int op_index,
                           int64_t* maps, int max_maps) {
  DeviceProfile prof;
  if (lds_max > 0) prof.lds_max = lds_max;
  tuning_apply(tuning_get(), prof);
  Plan plan;
  int rc = make_plan(dims, perm, order, elem_bytes, n_sub, alpha, prof, prof.reorder, plan);
  if (rc) return fail(rc, "planning failed (%d)", rc);
  if (op_index < 0 || op_index >= (int)plan.ops.size() || !plan.ops[op_index].cyc)
    return fail(EITHOT_ERR_INVALID, "pass %d is not a cycle-following pass", op_index);
  const std::vector<CycleMap>& mv = plan.ops[op_index].cyc->maps;
  for (int k = 0; k < (int)mv.size() && k < max_maps && maps; ++k) {
    maps[4 * k] = mv[k].bc;
    maps[4 * k + 1] = mv[k].R;
    maps[4 * k + 2] = mv[k].C;
    maps[4 * k + 3] = mv[k].m;
  }
  return (int)mv.size();
}

int eithot_plan_segments(const int64_t* dims, const int* perm, int order, size_t elem_bytes,
                         int n_sub, double alpha, int64_t lds_max, int op_index, int64_t* segb,
                         int64_t* inner, int64_t* outer) {
  DeviceProfile prof;
  if (lds_max > 0) prof.lds_max = lds_max;
  tuning_apply(tuning_get(), prof);
  Plan plan;
  int rc = make_plan(dims, perm, order, elem_bytes, n_sub, alpha, prof, prof.reorder, plan);
  if (rc) return fail(rc, "planning failed (%d)", rc);
  if (op_index < 0 || op_index >= (int)plan.ops.size() ||
      plan.ops[op_index].kind != OP_SEGTILE)
    return fail(EITHOT_ERR_INVALID, "pass %d is not a segmented tile pass", op_index);
  const Op& o = plan.ops[op_index];
  if (segb) *segb = o.segb;
  for (int k = 0; k < 6; ++k) {
    if (inner) {
      inner[2 * k] = k < o.sin_n ? o.sin_radix[k] : 1;
      inner[2 * k + 1] = k < o.sin_n ? o.sin_stride[k] : 0;
    }
    if (outer) {
      outer[2 * k] = k < o.sout_n ? o.sout_radix[k] : 1;
      outer[2 * k + 1] = k < o.sout_n ? o.sout_stride[k] : 0;
    }
  }
  return o.sin_n + 16 * o.sout_n;
}

int eithot_set_tuning(const char* key, int64_t value) {
  int rc = tuning_set(key, value);
  if (rc) return fail(rc, "unknown tuning key or value out of range");
  return EITHOT_OK;
}

int eithot_get_tuning(const char* key, int64_t* value) {
  int rc = tuning_value(key, value);
  if (rc) return fail(rc, "unknown tuning key");
  return EITHOT_OK;
}

void eithot_set_pass_timing(int enable) { g_timing = enable ? 1 : 0; }

int eithot_last_pass_times(float* ms, int max) {
  const int n = (int)g_pass_ms.size();
  for (int i = 0; i < n && i < max && ms; ++i) ms[i] = g_pass_ms[i];
  return n;
}

int eithot_fill_pattern(void* dev, uint64_t n, size_t elem_bytes, uint64_t seed, int mode,
                        uint64_t first, void* stream) {
  if (!dev) return fail(EITHOT_ERR_INVALID, "null pointer");
  hipError_t e = launch_fill((char*)dev, n, (int)elem_bytes, seed, mode, first, (hipStream_t)stream);
  if (e != hipSuccess) return fail(EITHOT_ERR_DEVICE, "fill failed: %s", hipGetErrorString(e));
  return EITHOT_OK;
}

int eithot_verify_pattern(const void* dev, const int64_t* dims, const int* perm, int order,
                          size_t elem_bytes, uint64_t seed, int mode,
                          unsigned long long* mismatches) {
  int rc = validate(dev, dims, perm, order, elem_bytes);
  if (rc) return rc;
  if (order > 16) return fail(EITHOT_ERR_UNSUPPORTED, "verify supports order <= 16");
  std::vector<int> p0(order);
  for (int i = 0; i < order; ++i) p0[i] = perm[i] - 1;
  unsigned long long* d_cnt = nullptr;
  if (hipMalloc((void**)&d_cnt, sizeof(unsigned long long)) != hipSuccess)
    return fail(EITHOT_ERR_NOMEM, "counter allocation failed");
  hipError_t e = hipMemset(d_cnt, 0, sizeof(unsigned long long));
  if (e == hipSuccess)
    e = launch_verify((const char*)dev, dims, p0.data(), order, (int)elem_bytes, seed, mode,
                      d_cnt, nullptr);
  unsigned long long h = 0;
  if (e == hipSuccess) e = hipMemcpy(&h, d_cnt, sizeof(h), hipMemcpyDeviceToHost);
  const hipError_t fe = hipFree(d_cnt);
  if (e == hipSuccess) e = fe;
  if (e != hipSuccess) return fail(EITHOT_ERR_DEVICE, "verify failed: %s", hipGetErrorString(e));
  if (mismatches) *mismatches = h;
  return EITHOT_OK;
}

}  // extern "C"
