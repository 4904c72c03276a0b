// Internal runtime state shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "plan.hpp"
#include "plan_cache.hpp"

namespace eithot {

using PlanPtr = std::shared_ptr<const Plan>;

// Per-device state.  Two locks:
//   mu         the plan cache;
//   enqueue_mu held from the first use of the arena or of pass tables to the
//              event recorded after a call's last pass: arena growth or
//              shrinking, table eviction and every launch that uses them are
//              mutually exclusive, so no thread launches into memory another
//              thread frees.  Lock order: enqueue_mu, then mu.
struct Device {
  int id = 0;
  DeviceProfile prof;
  LaunchCfg cfg;
  std::mutex mu;
  LruCache<Plan> plans{kPlanCacheEntries, kPlanCacheHostBytes};
  std::mutex enqueue_mu;
  char* arena = nullptr;             // scratch for staged passes and cycle boundaries
  int64_t arena_bytes = 0;
  hipEvent_t arena_ready = nullptr;  // recorded after the last pass that used the arena
  // device bytes of the pass tables uploaded for plans of this cache (and
  // not yet freed; an evicted plan a request still holds counts until freed)
  std::atomic<int64_t> cached_tables{0};

  // plans cached per device, and the host bytes of their pass tables
  static constexpr size_t kPlanCacheEntries = 8192;
  static constexpr uint64_t kPlanCacheHostBytes = 16ull << 30;

  // The cached plan for these arguments, planned on first sight (outside
  // the cache lock).  The caller's reference keeps the plan alive even if
  // the cache evicts it.
  PlanPtr plan_for(const int64_t* dims, const int* perm1, int rank, size_t elem, int n_sub,
                   double alpha, int* rc);
  // hardware profile / launch configuration with the tuning knobs applied
  DeviceProfile profile() const;
  LaunchCfg launch_cfg() const;
  // With enqueue_mu held: keeps the device bytes the library holds (arena
  // plus the resident tables of cached plans) within `budget` by freeing the
  // tables of other cached plans, least recently used first, and shrinking
  // an arena larger than `keep` needs.
  void trim_locked(int64_t budget, const Plan* keep);
  // With enqueue_mu held: arena of at least `bytes`, ordered after its last
  // user on `stream`.
  int reserve_arena_locked(int64_t bytes, hipStream_t stream);
  int release_arena_after_locked(hipStream_t stream);
  // Device bytes the library holds for plain transposes on this device.
  int64_t held_bytes();
};

// Cache key of a plan: the arguments and the tuning epoch.
std::string plan_key(const int64_t* dims, const int* perm1, int rank, size_t elem, int n_sub,
                     double alpha);
Device* device_ctx(int* err);
int fail(int code, const char* fmt, ...);
void set_peak(size_t bytes);
int validate(const void* dev, const int64_t* dims, const int* perm, int order, size_t elem);
// Enqueues the plan's passes on `stream`.  min_groups_cap > 0 lowers the
// launch configuration's workgroup floor per pass to at most that (callers
// whose passes share the GPU with other work).
int run_plan(Device* d, const Plan& plan, char* data, char* scratch, int64_t scratch_bytes,
             hipStream_t stream, std::vector<float>* times, int64_t min_groups_cap = 0);

}  // namespace eithot
