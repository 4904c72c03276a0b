// Internal runtime state shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "plan.hpp"

namespace eithot {

struct Device {
  int id = 0;
  DeviceProfile prof;
  LaunchCfg cfg;
  std::mutex mu;
  std::map<std::string, std::unique_ptr<Plan>> plans;
  char* arena = nullptr;       // scratch for staged passes, grown on demand
  int64_t arena_bytes = 0;
  hipEvent_t arena_ready = nullptr;  // last stream use of the arena

  const Plan* plan_for(const int64_t* dims, const int* perm1, int rank, size_t elem, int n_sub,
                       double alpha, int* rc);
  // hardware profile / launch configuration with the tuning knobs applied
  DeviceProfile profile() const;
  LaunchCfg launch_cfg() const;
  int reserve_arena(int64_t bytes, hipStream_t stream);
  int release_arena_after(hipStream_t stream);
};

Device* device_ctx(int* err);
int fail(int code, const char* fmt, ...);
void set_peak(size_t bytes);
int validate(const void* dev, const int64_t* dims, const int* perm, int order, size_t elem);
int run_plan(Device* d, const Plan& plan, char* data, char* scratch, int64_t scratch_bytes,
             hipStream_t stream, std::vector<float>* times);

}  // namespace eithot
