// Batched transposition: many independent tensors (orders, shapes and
// element sizes may all differ) transposed with one persistent launch per
// decomposition-stage index, captured once in a HIP graph and replayed.
//
// Launch k runs pass k of every tensor whose plan has one; passes that
// cannot join a persistent launch (scratch-staged or 64-bit indexed) run as
// their own launches right after launch k.  Every tensor's passes therefore
// keep their order, and tensors never share footprints.
//
// The autotuner times candidate launch configurations (footprint packing,
// workgroup size, persistent grid) on the real batch: each candidate runs
// the forward batch and then the inverse batch, so the tensors end where
// they started and the chosen configuration is replayed from the graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <vector>

#include "../../include/eithot.h"
#include "kernels.hpp"
#include "plan.hpp"
#include "runtime.hpp"

using namespace eithot;

namespace {

struct TensorReq {
  char* dev;
  std::vector<int64_t> dims;
  std::vector<int> perm;
  size_t elem;
};

struct LaunchStage {
  BatchLaunch* persistent = nullptr;
  std::vector<std::pair<const Op*, char*>> side;  // run with launch_op after the launch
};

struct Candidate {
  int64_t lds_target;
  int threads;
  int grid_per_cu;
};

void free_stages(std::vector<LaunchStage>& v) {
  for (LaunchStage& s : v) batch_launch_free(s.persistent);
  v.clear();
}

}  // namespace

struct eithot_batch {
  std::vector<TensorReq> tensors;
  std::vector<Plan> fwd, inv;
  std::vector<LaunchStage> stages;
  char* scratch = nullptr;
  int64_t scratch_bytes = 0;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  Device* dev = nullptr;
  LaunchCfg cfg;
  int grid_per_cu = 0;
  bool built = false;

  ~eithot_batch() {
    free_stages(stages);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (scratch) (void)hipFree(scratch);
    if (stream) (void)hipStreamDestroy(stream);
  }

  int build_stages(const std::vector<Plan>& plans, const LaunchCfg& c, int gpc,
                   std::vector<LaunchStage>& out) {
    size_t depth = 0;
    for (const Plan& p : plans) depth = std::max(depth, p.ops.size());
    out.assign(depth, LaunchStage());
    for (size_t k = 0; k < depth; ++k) {
      std::vector<std::pair<const Op*, char*>> ops;
      for (size_t t = 0; t < plans.size(); ++t)
        if (k < plans[t].ops.size()) ops.push_back({&plans[t].ops[k], tensors[t].dev});
      std::vector<size_t> rejected;
      int rc = batch_launch_build(ops, c, gpc, dev->prof.cus, scratch, scratch_bytes,
                                  &out[k].persistent, &rejected);
      if (rc) return rc;
      for (size_t i : rejected) out[k].side.push_back(ops[i]);
    }
    return EITHOT_OK;
  }

  int enqueue(const std::vector<LaunchStage>& st, const LaunchCfg& c, hipStream_t s) {
    for (const LaunchStage& x : st) {
      if (batch_launch_enqueue(x.persistent, s) != hipSuccess)
        return fail(EITHOT_ERR_DEVICE, "batched launch failed");
      for (const auto& so : x.side) {
        const hipError_t e = launch_op(*so.first, so.second, scratch, scratch_bytes, c, s);
        if (e != hipSuccess)
          return fail(EITHOT_ERR_DEVICE, "batched side pass (kind %d) failed: %s", so.first->kind,
                      hipGetErrorString(e));
      }
    }
    return EITHOT_OK;
  }
};

extern "C" {

eithot_batch* eithot_batch_create(void) {
  int err = 0;
  Device* d = device_ctx(&err);
  if (!d) return nullptr;
  auto* b = new eithot_batch;
  b->dev = d;
  b->cfg = d->cfg;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return nullptr;
  }
  return b;
}

int eithot_batch_add(eithot_batch* b, void* dev, const int64_t* dims, const int* perm, int order,
                     size_t elem_bytes) {
  if (!b) return fail(EITHOT_ERR_INVALID, "null batch");
  if (b->built) return fail(EITHOT_ERR_INVALID, "batch already built");
  int rc = validate(dev, dims, perm, order, elem_bytes);
  if (rc) return rc;
  b->tensors.push_back({(char*)dev, std::vector<int64_t>(dims, dims + order),
                        std::vector<int>(perm, perm + order), elem_bytes});
  return (int)b->tensors.size() - 1;
}

int eithot_batch_build(eithot_batch* b, double alpha, int tune) {
  if (!b) return fail(EITHOT_ERR_INVALID, "null batch");
  if (b->built) return fail(EITHOT_ERR_INVALID, "batch already built");
  if (!(alpha > 0)) return fail(EITHOT_ERR_INVALID, "alpha must be > 0");
  const size_t n = b->tensors.size();
  DeviceProfile prof = b->dev->profile();
  prof.cycle_shift = 0;  // batched cycle items move narrow cells word by word
  b->cfg = b->dev->launch_cfg();
  b->fwd.resize(n);
  b->inv.resize(n);
  for (size_t t = 0; t < n; ++t) {
    const TensorReq& r = b->tensors[t];
    const int order = (int)r.dims.size();
    int rc = make_plan(r.dims.data(), r.perm.data(), order, r.elem, 1, alpha, prof, prof.reorder,
                       b->fwd[t]);
    if (rc) return fail(rc, "planning tensor %zu failed (%d)", t, rc);
    b->scratch_bytes = std::max(b->scratch_bytes, b->fwd[t].scratch_bytes);
    if (tune) {
      std::vector<int64_t> od(order);
      std::vector<int> ip(order);
      for (int k = 0; k < order; ++k) {
        od[k] = r.dims[r.perm[k] - 1];
        ip[r.perm[k] - 1] = k + 1;
      }
      rc = make_plan(od.data(), ip.data(), order, r.elem, 1, alpha, prof, prof.reorder,
                     b->inv[t]);
      if (rc) return fail(rc, "planning inverse of tensor %zu failed (%d)", t, rc);
      b->scratch_bytes = std::max(b->scratch_bytes, b->inv[t].scratch_bytes);
    }
  }
  // the batched cycle passes of one stage run together: each keeps its own
  // boundary cells, so a stage needs the sum of theirs (side passes run
  // after the batched launch and reuse the buffer from its start)
  for (const std::vector<Plan>* plans : {&b->fwd, &b->inv}) {
    size_t depth = 0;
    for (const Plan& p : *plans) depth = std::max(depth, p.ops.size());
    for (size_t k = 0; k < depth; ++k) {
      int64_t sum = 0;
      for (const Plan& p : *plans)
        if (k < p.ops.size()) sum += batch_boundary_bytes(p.ops[k]);
      b->scratch_bytes = std::max(b->scratch_bytes, sum);
    }
  }
  if (b->scratch_bytes > 0 && hipMalloc((void**)&b->scratch, (size_t)b->scratch_bytes) != hipSuccess) {
    (void)hipGetLastError();
    return fail(EITHOT_ERR_NOMEM, "batch scratch allocation failed");
  }
  std::vector<Candidate> cands = {{b->cfg.lds_target, b->cfg.threads, 0}};
  if (tune) {
    cands.clear();
    for (int64_t lt : {32 * 1024, 64 * 1024, 128 * 1024})
      for (int th : {256, 512})
        for (int gpc : {0, 4}) cands.push_back({std::min<int64_t>(lt, b->dev->prof.lds_max), th, gpc});
  }
  float best_ms = 1e30f;
  Candidate best = cands[0];
  if (cands.size() > 1) {
    hipEvent_t e0, e1, e2;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventCreate(&e2);
    for (const Candidate& c : cands) {
      LaunchCfg cfg = b->cfg;
      cfg.lds_target = c.lds_target;
      cfg.threads = c.threads;
      cfg.threads_big = std::max(c.threads, 512);
      std::vector<LaunchStage> f, v;
      int rc = b->build_stages(b->fwd, cfg, c.grid_per_cu, f);
      if (!rc) rc = b->build_stages(b->inv, cfg, c.grid_per_cu, v);
      if (!rc) {
        (void)hipEventRecord(e0, b->stream);
        rc = b->enqueue(f, cfg, b->stream);
        (void)hipEventRecord(e1, b->stream);
        if (!rc) rc = b->enqueue(v, cfg, b->stream);
        (void)hipEventRecord(e2, b->stream);
        if (hipStreamSynchronize(b->stream) != hipSuccess) rc = EITHOT_ERR_DEVICE;
      }
      float ms = 0;
      if (!rc && hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms < best_ms) {
        best_ms = ms;
        best = c;
      }
      free_stages(f);
      free_stages(v);
      if (rc) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipEventDestroy(e2);
        return rc;
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
  }
  b->cfg.lds_target = best.lds_target;
  b->cfg.threads = best.threads;
  b->cfg.threads_big = std::max(best.threads, 512);
  b->grid_per_cu = best.grid_per_cu;
  b->inv.clear();
  int rc = b->build_stages(b->fwd, b->cfg, b->grid_per_cu, b->stages);
  if (rc) return rc;
  for (const LaunchStage& st : b->stages)
    for (const auto& so : st.side)
      if (prepare_op(*so.first) != hipSuccess)
        return fail(EITHOT_ERR_DEVICE, "uploading a pass table failed");
  // capture the forward launches once; replays reuse the instantiated graph
  if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return fail(EITHOT_ERR_DEVICE, "graph capture failed to start");
  rc = b->enqueue(b->stages, b->cfg, b->stream);
  hipError_t e = hipStreamEndCapture(b->stream, &b->graph);
  if (rc) return rc;
  if (e != hipSuccess) return fail(EITHOT_ERR_DEVICE, "graph capture failed");
  if (hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess)
    return fail(EITHOT_ERR_DEVICE, "graph instantiation failed");
  b->built = true;
  return EITHOT_OK;
}

int eithot_batch_run(eithot_batch* b, void* stream, int sync) {
  if (!b || !b->built) return fail(EITHOT_ERR_INVALID, "batch not built");
  hipStream_t s = stream ? (hipStream_t)stream : b->stream;
  if (hipGraphLaunch(b->exec, s) != hipSuccess) return fail(EITHOT_ERR_DEVICE, "graph launch failed");
  if (sync && hipStreamSynchronize(s) != hipSuccess)
    return fail(EITHOT_ERR_DEVICE, "batch execution failed");
  return EITHOT_OK;
}

int eithot_batch_info(const eithot_batch* b, int* launches, int64_t* scratch_bytes,
                      int* graph_captured) {
  if (!b) return fail(EITHOT_ERR_INVALID, "null batch");
  int n = 0;
  for (const auto& s : b->stages) n += batch_launch_count(s.persistent) + (int)s.side.size();
  if (launches) *launches = n;
  if (scratch_bytes) *scratch_bytes = b->scratch_bytes;
  if (graph_captured) *graph_captured = b->exec != nullptr;
  return EITHOT_OK;
}

void eithot_batch_destroy(eithot_batch* b) { delete b; }

}  // extern "C"
