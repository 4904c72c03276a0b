// Transposition that allocates no device memory: for callers that cannot
// call hipMalloc (tight serving loops, harnesses that forbid allocation).
//
// The scratch a plan needs (staged footprints, the boundary cells of cut
// cycles) and the device copies of its pass tables live in one fixed static
// device array, `scratch`, of kNoAllocBytes; every GPU the library runs on
// holds its own copy until the process ends.  Its device address is looked
// up once per device with hipGetSymbolAddress; host code uses that address,
// never the array's name.  The kernels never rely on the array's previous
// contents (every byte they read from it was written earlier in the same
// call), so nothing is cleared.  Calls that overlap on one GPU share the
// array and race; each call returns only when its work is done.
#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/eithot.h"
#include "kernels.hpp"
#include "plan.hpp"
#include "runtime.hpp"

namespace {

constexpr size_t kNoAllocBytes = 64ull << 20;
constexpr size_t kAlign = 256;

static __device__ __align__(256) char scratch[kNoAllocBytes];

struct DevState {
  char* ptr = nullptr;                   // device address of `scratch` on this device
  std::weak_ptr<const eithot::Plan> resident;  // plan whose tables the array holds
};

std::mutex g_mu;
std::map<int, DevState>& states() {
  static auto* m = new std::map<int, DevState>();  // never destroyed (see device_ctx)
  return *m;
}

// Plans of this entry point, apart from the device cache of the others:
// their tables are only ever copied into the static array, never uploaded
// to memory of their own, so evicting one frees host memory only.
std::mutex g_plan_mu;
eithot::LruCache<eithot::Plan>& plans() {
  static auto* c = new eithot::LruCache<eithot::Plan>(64, 1ull << 30);
  return *c;
}

eithot::PlanPtr plan_noalloc(const eithot::Device* d, const int64_t* dims, const int* perm,
                             int order, size_t elem, int n_sub, double alpha, int* rc) {
  const std::string key = std::to_string(d->id) + "@" +
                          eithot::plan_key(dims, perm, order, elem, n_sub, alpha);
  {
    std::lock_guard<std::mutex> lk(g_plan_mu);
    if (eithot::PlanPtr p = plans().get(key)) {
      *rc = EITHOT_OK;
      return p;
    }
  }
  const eithot::DeviceProfile pr = d->profile();
  auto p = std::make_shared<eithot::Plan>();
  *rc = eithot::make_plan(dims, perm, order, elem, n_sub, alpha, pr, pr.reorder, *p);
  if (*rc != EITHOT_OK) return nullptr;
  std::vector<eithot::PlanPtr> evicted;
  std::lock_guard<std::mutex> lk(g_plan_mu);
  if (eithot::PlanPtr q = plans().get(key)) return q;
  evicted = plans().put(key, p, (uint64_t)std::max<int64_t>(p->table_bytes, 0));
  return p;
}

size_t align_up(size_t n) { return (n + kAlign - 1) / kAlign * kAlign; }

// Device table bytes of one op (0 when it reads none), laid out as the
// launcher uploads them.
size_t table_bytes(const eithot::Op& op) {
  if (op.kind == eithot::OP_CYCLE && op.cyc && !op.cyc->segs.empty())
    return op.cyc->segs.size() * sizeof(eithot::CycleSeg) + op.cyc->slots.size() * 4;
  if (op.kind == eithot::OP_BLOCK && op.blk && !op.blk->ndig)
    return op.blk->src.size() * (op.blk->cells <= 65536 ? 2 : 4);
  return 0;
}

}  // namespace

using namespace eithot;

extern "C" int eithot_transpose_noalloc(void* dev, const int64_t* dims, const int* perm, int order,
                                        size_t elem_bytes, int n_sub, double alpha) {
  set_peak(0);
  int rc = validate(dev, dims, perm, order, elem_bytes);
  if (rc) return rc;
  if (!(alpha > 0)) return fail(EITHOT_ERR_INVALID, "alpha must be > 0");
  Device* d = device_ctx(&rc);
  if (!d) return rc;
  // set-up on the first call on this device: the array's address here
  char* base = nullptr;
  std::unique_lock<std::mutex> lk(g_mu);
  DevState& st = states()[d->id];
  if (!st.ptr) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(EITHOT_ERR_DEVICE, "hipGetSymbolAddress of the static scratch failed: %s",
                  hipGetErrorString(e));
    st.ptr = static_cast<char*>(p);
  }
  base = st.ptr;
  lk.unlock();
  PlanPtr plan = plan_noalloc(d, dims, perm, order, elem_bytes, n_sub, alpha, &rc);
  if (!plan) return fail(rc, "planning failed (%d)", rc);
  // layout: the pass tables (each on a 256-byte boundary), then the scratch
  std::vector<size_t> toff(plan->ops.size(), 0);
  size_t used = 0;
  for (size_t i = 0; i < plan->ops.size(); ++i) {
    const size_t tb = table_bytes(plan->ops[i]);
    if (!tb) continue;
    toff[i] = used;
    used += align_up(tb);
  }
  const size_t scratch_off = used;
  used += align_up((size_t)std::max<int64_t>(plan->scratch_bytes, 0));
  if (used > kNoAllocBytes)
    return fail(EITHOT_ERR_NOMEM,
                "the plan needs %zu bytes of scratch and pass tables; the static scratch of "
                "eithot_transpose_noalloc holds %zu",
                used, kNoAllocBytes);
  hipStream_t stream = nullptr;
  // the tables, unless the array already holds this plan's from the last call
  lk.lock();
  const bool fresh = st.resident.lock() != plan;
  lk.unlock();
  if (fresh) {
    std::vector<uint16_t> idx16;
    for (size_t i = 0; i < plan->ops.size(); ++i) {
      const Op& op = plan->ops[i];
      if (!table_bytes(op)) continue;
      char* dst = base + toff[i];
      hipError_t e = hipSuccess;
      if (op.kind == OP_CYCLE) {
        const size_t sb = op.cyc->segs.size() * sizeof(CycleSeg);
        e = hipMemcpy(dst, op.cyc->segs.data(), sb, hipMemcpyHostToDevice);
        if (e == hipSuccess && !op.cyc->slots.empty())
          e = hipMemcpy(dst + sb, op.cyc->slots.data(), op.cyc->slots.size() * 4,
                        hipMemcpyHostToDevice);
      } else if (op.blk->cells <= 65536) {
        idx16.assign(op.blk->src.begin(), op.blk->src.end());
        e = hipMemcpy(dst, idx16.data(), idx16.size() * 2, hipMemcpyHostToDevice);
      } else {
        e = hipMemcpy(dst, op.blk->src.data(), op.blk->src.size() * 4, hipMemcpyHostToDevice);
      }
      if (e != hipSuccess) {
        lk.lock();
        st.resident.reset();
        return fail(EITHOT_ERR_DEVICE, "copying pass tables failed: %s", hipGetErrorString(e));
      }
    }
    lk.lock();
    st.resident = plan;
    lk.unlock();
  }
  const LaunchCfg cfg = d->launch_cfg();
  char* arena = plan->scratch_bytes > 0 ? base + scratch_off : nullptr;
  for (size_t i = 0; i < plan->ops.size(); ++i) {
    const Op& op = plan->ops[i];
    const void* table = table_bytes(op) ? (const void*)(base + toff[i]) : nullptr;
    hipError_t e = launch_op(op, (char*)dev, arena, plan->scratch_bytes, cfg, stream, table);
    if (e != hipSuccess)
      return fail(EITHOT_ERR_DEVICE, "pass %zu launch failed: %s", i, hipGetErrorString(e));
  }
  hipError_t e = hipStreamSynchronize(stream);
  if (e == hipSuccess) e = hipGetLastError();
  if (e != hipSuccess) return fail(EITHOT_ERR_DEVICE, "device error: %s", hipGetErrorString(e));
  set_peak(used);
  return EITHOT_OK;
}
