// Launch interface of the gfx950 pass kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <atomic>
#include <memory>
#include <utility>
#include <vector>

#include "plan.hpp"

namespace eithot {

struct LaunchCfg {
  int64_t lds_target = 64 * 1024;  // footprints are packed into one workgroup up to this
  int threads = 256;
  int threads_big = 512;           // workgroups whose LDS footprint exceeds 96 KiB
  int xcd_remap = 1;               // keep consecutive footprints on one XCD
  int cycle_variant = 0;           // cycle-following kernel shape (0..3)
  int segtile_threads = 0;         // > 0: workgroup size of segmented tile passes
  int lds_threads_auto = 1;        // 1: tile/column/row workgroups sized for one load round
  int segtile_nt = 3;              // non-temporal segmented tile loads (bit 0) and stores (bit 1)
  int cycle_u_depth = 4;           // positions in flight per item of the shifted cycle kernel
  int64_t min_groups = 1024;       // workgroups a pass's footprints spread over at least (fewer
                                   // footprints per workgroup below that); batched launches,
                                   // whose passes share the GPU, use 1
};

// One pass of `op` over the tensor at `data`, on `stream`.  Staged ops use
// `scratch` (scratch_bytes >= one footprint).  `table`, when given, is the
// device copy of the op's cycle or block table to read (laid out as the
// launcher's own upload: segments then slots; uint16 indices when a block
// has at most 65536 cells, else uint32); otherwise the launcher uploads the
// table on first use.  Returns a hipError_t.
hipError_t launch_op(const Op& op, char* data, char* scratch, int64_t scratch_bytes,
                     const LaunchCfg& cfg, hipStream_t stream, const void* table = nullptr);

// Uploads what an op needs on the current device before its first launch
// (the cycle table of OP_CYCLE; nothing for other kinds).  Call before
// capturing the op into a graph.
hipError_t prepare_op(const Op& op);

// While in scope, pass tables uploaded by this thread are charged to `*c`
// (added on upload, subtracted when freed); the counter must outlive them.
struct TableAccount {
  explicit TableAccount(std::atomic<int64_t>* c);
  ~TableAccount();
  TableAccount(const TableAccount&) = delete;
  TableAccount& operator=(const TableAccount&) = delete;
  std::atomic<int64_t>* prev;
};
// Owner of device memory `p` (`bytes`) on device `dev` whose deleter waits
// for the device to go idle before hipFree; the bytes stay charged to the
// thread's TableAccount at upload time until then.
std::shared_ptr<void> device_owned(void* p, int dev, int64_t bytes);

// Bytes of LDS the op's LDS kernel uses per workgroup with cfg (0 if staged).
int64_t op_lds_bytes(const Op& op, const LaunchCfg& cfg);

// Fill n elements of elem bytes with the counter-based pattern v(seed, first+i)
// (mode 0: v = i, mode 1: splitmix64 of (seed, i)).
hipError_t launch_fill(char* data, uint64_t n, int elem, uint64_t seed, int mode, uint64_t first,
                       hipStream_t stream);

// Count output elements that differ from the permuted pattern: out[dest(i)]
// must equal v(seed, i) for every input linear index i.
hipError_t launch_verify(const char* data, const int64_t* dims, const int* perm0, int rank,
                         int elem, uint64_t seed, int mode, unsigned long long* d_mismatch,
                         hipStream_t stream);

// Persistent batched launch: one pass (op) of many tensors in one kernel.
// ops pairs each op with its tensor's device pointer; staged or 64-bit-index
// ops are not batched and are returned in `rejected` (indices into ops), as
// are cycle passes whose boundary cells no longer fit `scratch` (each cycle
// pass gets batch_boundary_bytes of it).  grid_per_cu > 0 sizes a persistent
// grid (workgroups per CU), 0 launches one workgroup per item.
struct BatchLaunch;
int64_t batch_boundary_bytes(const Op& op);
int batch_launch_build(const std::vector<std::pair<const Op*, char*>>& ops, const LaunchCfg& cfg,
                       int grid_per_cu, int cus, char* scratch, int64_t scratch_bytes,
                       BatchLaunch** out, std::vector<size_t>* rejected);
hipError_t batch_launch_enqueue(const BatchLaunch* L, hipStream_t stream);
void batch_launch_free(BatchLaunch* L);
int64_t batch_launch_items(const BatchLaunch* L);
int batch_launch_count(const BatchLaunch* L);  // kernel launches per enqueue

// Raise the dynamic-LDS limit of every pass kernel to `bytes` (once per device).
hipError_t configure_kernels(int64_t bytes);

}  // namespace eithot
