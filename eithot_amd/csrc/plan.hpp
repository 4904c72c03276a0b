// Host-side planner for in-place tensor transposition on MI355X.
//
// Pipeline: N-d permutation -> decomposer rounds (low-order stages, each a
// swap of two adjacent dim blocks) -> every stage is a batched 2-D transpose
// of U-byte units -> lowered to device passes (one-pass LDS tile transpose,
// Catanzaro C2R/R2C column/row passes, or hierarchical splits) with a cost
// model written for gfx950 (160 KB LDS per CU, 128-byte lines).
#pragma once

#include <cstddef>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace eithot {

// ---- decomposer ------------------------------------------------------------
// One low-order stage: the tensor viewed as (P, X, Y, S) [fastest first]
// becomes (P, Y, X, S).  `rank`/`perm`/`dims` give the reference's low-order
// form (21, 213, 132 or 1324 with merged dims).
struct Stage {
  int rank = 0;
  int perm[4] = {0, 0, 0, 0};
  int64_t dims[4] = {1, 1, 1, 1};
  int64_t P = 1, X = 1, Y = 1, S = 1;
};

enum Reorder { REORDER_MIX = 0, REORDER_PRE = 1, REORDER_POST = 2 };

// perm1 is 1-based, dims fastest first.  Returns false on an invalid perm.
bool decompose(const int64_t* dims, const int* perm1, int rank, int reorder,
               std::vector<Stage>& out);

// ---- device passes ---------------------------------------------------------
enum OpKind : int { OP_TILE = 1, OP_COL = 2, OP_ROW = 3, OP_CYCLE = 4, OP_BLOCK = 5, OP_SEGTILE = 6 };

// Index maps of the passes, all in gather form (dest <- source).
enum MapKind : int {
  MAP_NONE = 0,
  MAP_ROT = 1,         // C2R column rotate:   i' = (i + floor(j/b)) mod m
  MAP_ROTINV = 2,      // R2C column rotate:   i' = (i - floor(j/b)) mod m
  MAP_COLSH = 3,       // C2R column shuffle:  i' = (i*n + j - floor(i/a)) mod m
  MAP_COLSHINV = 4,    // R2C column shuffle:  inverse of MAP_COLSH
  MAP_ROWG = 5,        // C2R row shuffle (gather form)
  MAP_ROWD = 6,        // R2C row shuffle (gather form)
};

// Device copies of a host table, one per device id, made by the launcher on
// first use.  Each copy is a shared_ptr whose deleter waits for the device
// to go idle before freeing it, so dropping a copy never frees memory that a
// pass already enqueued still reads.
struct DeviceCopies {
  mutable std::mutex mu;
  mutable std::map<int, std::shared_ptr<void>> by_dev;
  bool has(int dev) const {
    std::lock_guard<std::mutex> lk(mu);
    return by_dev.count(dev) != 0;
  }
  // Drops the copy on `dev`; it is freed when the last holder lets go.
  void drop(int dev) const {
    std::shared_ptr<void> hold;
    {
      std::lock_guard<std::mutex> lk(mu);
      auto it = by_dev.find(dev);
      if (it == by_dev.end()) return;
      hold = std::move(it->second);
      by_dev.erase(it);
    }
  }
};

// Cycle structure of the R x C transpose of cells (OP_CYCLE).  In gather
// form output position p = j*R + i takes input position src(p) = i*C + j;
// following p -> src(p) walks a cycle.  Cycles longer than `lmax` are cut
// into segments; the first position of every cut segment gets a boundary
// slot whose old cell is saved before the main pass, because the segment
// before it in the cycle reads it last.
struct CycleSeg {
  uint32_t start;    // first position written
  uint32_t len;      // positions written: start, src(start), ...
  int32_t src_slot;  // slot holding the old cell the last write takes; -1: the
                     // segment is a whole cycle and takes old(start)
  uint32_t last;     // whole cycle: its last position (the one src maps to start)
};
// One cell map of a cycle pass: cells in blocks of bc cells; a block is an
// R x C matrix of m-cell units, transposed: cell q = (j*R + i)*m + x of a
// block takes cell (i*C + j)*m + x.
struct CycleMap {
  int64_t bc, R, C, m;
};
constexpr int kMaxCycleMaps = 5;  // maps one cycle pass composes
struct CycleTable {
  int64_t lmax = 0;
  // maps applied in pass order; the pass's source of position p is
  // src_1(src_2(...src_k(p))) (gather composition)
  std::vector<CycleMap> maps;
  std::vector<CycleSeg> segs;
  std::vector<uint32_t> slots;  // slot -> position whose old cell it holds
  int64_t bytes() const { return (int64_t)(segs.size() * sizeof(CycleSeg) + slots.size() * 4); }
  // device copies (segs then slots), one per device id, made by the launcher
  DeviceCopies dev;
};
// Builds the table (fixed points are skipped).  R*C must be below 2^32.
std::shared_ptr<CycleTable> build_cycle_table(int64_t R, int64_t C, int64_t lmax);
// Same over the composition of several cell maps on N cells (N < 2^32).
std::shared_ptr<CycleTable> build_cycle_table_maps(const std::vector<CycleMap>& maps, int64_t N,
                                                   int64_t lmax);

// Gather table of a fused block permutation (OP_BLOCK): consecutive
// contiguous tile passes whose matrices nest inside blocks of `cells` cells
// of `cell_bytes` bytes compose into one permutation of every block; output
// cell o of a block takes input cell src[o].
struct BlockTable {
  int64_t cells = 0, cell_bytes = 0;
  int pad_log2 = 0;  // > 0: LDS holds 4 pad bytes after every 2^pad_log2 bytes of a block
  std::vector<uint32_t> src;
  // When the permutation is a digit permutation (it is for nested
  // transposes): with o's mixed-radix digits d_k (radix[k], fastest first),
  // src[o] = sum d_k * stride[k].  ndig = 0: use the table.
  int ndig = 0;
  uint32_t radix[8] = {0}, stride[8] = {0};
  int64_t lds_bytes() const {
    const int64_t F = cells * cell_bytes;
    return pad_log2 ? F + (F >> pad_log2) * 4 + 4 : F;
  }
  // device bytes: the gather table, unless the digit form replaces it
  int64_t bytes() const { return ndig ? 0 : (int64_t)src.size() * (cells <= 65536 ? 2 : 4); }
  DeviceCopies dev;  // device copy (uint16 or uint32 indices), only in table form
};

struct Op {
  int kind = 0;
  int map = MAP_NONE;
  int staged = 0;      // 1: footprint exceeds LDS, executed through scratch
  int stage = -1;      // index of the decomposer stage it belongs to (-1: partition/join)
  uint64_t base = 0;   // byte offset of batch 0
  int64_t B = 1, R = 1, C = 1;  // B batches of an R x C grid of units
  int64_t U = 1;       // unit bytes
  int64_t us = 1;      // slice bytes (U % us == 0)
  int64_t W = 1;       // OP_COL: columns per panel
  // Catanzaro constants of the m x n grid (m = R, n = C of this op's grid)
  int64_t m = 1, n = 1, c = 1, a = 1, b = 1, k = 0, binv = 0;

  std::vector<CycleMap> cmaps;            // OP_CYCLE: cell maps (in units of U-byte cells)
  std::shared_ptr<const CycleTable> cyc;  // OP_CYCLE: segments shared by all B regions
  std::shared_ptr<const BlockTable> blk;  // OP_BLOCK: B blocks of R cells of U bytes
  std::vector<int64_t> bparts;            // OP_BLOCK: the fused tiles, (R, C, U) each
  // OP_SEGTILE: B matrices of R x C units of U bytes (C fastest), each
  // transposed in place.  Matrix b's bytes, in row-major order, lie in
  // R*C*U/segb segments of segb bytes; segment s of matrix b starts at byte
  // outer(b) + inner(s), where outer/inner sum digit * stride over the
  // mixed-radix digits (fastest first) of b and of s.  The transposed matrix
  // is written back over the same segments in the same order.
  int64_t segb = 0;
  int sin_n = 0, sout_n = 0;
  int64_t sin_radix[6] = {0}, sin_stride[6] = {0};
  int64_t sout_radix[6] = {0}, sout_stride[6] = {0};

  int64_t footprint_units() const;   // units staged per footprint
  int64_t footprint_bytes() const { return footprint_units() * us; }
  int64_t num_footprints() const;
  uint64_t bytes() const { return (uint64_t)B * R * C * U; }
};

// Lowering options of one batched 2-D transpose (the "lower_mask" knob).
enum LowerOption : int64_t {
  LOWER_TILE = 1,         // LDS tile, whole units or slices
  LOWER_CATANZARO = 2,    // Catanzaro column / row passes (LDS or staged)
  LOWER_SPLIT = 4,        // hierarchical split into two smaller transposes
  LOWER_STAGED_TILE = 8,  // tile through scratch
  LOWER_CYCLE = 16,       // cycle following of whole cells
  LOWER_ALL = 31
};

struct DeviceProfile {
  int64_t lds_max = 160 * 1024;   // bytes of LDS one workgroup may use (gfx950)
  int64_t lds_target = 64 * 1024; // per-workgroup LDS the launcher packs footprints to
  int cus = 256;
  int64_t min_segment = 64;       // smallest contiguous run worth a footprint slice
  // Tile footprints above this are sliced (slices of at least slice_min
  // bytes) so that several workgroups share a CU and overlap their load and
  // store phases; 0 keeps whole units whenever they fit lds_max.
  int64_t tile_fp_target = 64 * 1024;
  int64_t slice_min = 128;
  int64_t stage_chunk = 0;        // > 0: bytes of scratch one staged group may fill
  int64_t cycle_min_cell = 128;   // smallest cell moved by cycle following (0: off)
  int64_t fuse_max_block = 64 * 1024;  // largest fused block permutation (0: no fusion)
  int64_t block_digits = 1;       // fused blocks index cells arithmetically when they can
  int64_t block_pad = 1;          // fused blocks pad LDS against bank conflicts (0: never)
  int64_t cycle_word_cost = 1;    // cycle passes of cells in narrow words cost more
  int reorder = 0;                // 0 cheapest of pre/post (mix breaks ties), 1 pre, 2 post, 3 mix
  int64_t regroup = 1;            // with reorder 0: also plan cell pass / tile / cell pass
  int64_t cycle_lmax = 0;         // > 0: cycle segments of at most this many positions
  int64_t cycle_cell_pref = 0;    // 1: cycle passes of larger cells cost less
  int64_t lower_mask = LOWER_ALL; // lowering options the planner may use
  int cycle_shift = 1;            // cycle passes of narrow cells run the shifted 16-byte
                                  // kernel (0: word by word, as in batched launches)
  double staged_mult = 1.0;       // planner-internal: weight of staged passes in the search
};

// Process-wide tuning knobs (eithot_set_tuning); applied over the hardware
// profile of every device and over the host-only planning entry points.
struct Tuning {
  int64_t lds_target = 64 * 1024;
  int64_t tile_fp_target = 64 * 1024;
  int64_t slice_min = 128;
  int64_t stage_chunk = 0;
  int64_t cycle_min_cell = 128;
  int64_t fuse_max_block = 64 * 1024;
  int64_t block_digits = 1;
  int64_t block_pad = 1;
  int64_t cycle_word_cost = 1;
  int64_t threads = 256;
  int64_t threads_big = 512;
  int64_t xcd_remap = 1;
  int64_t cycle_variant = 0;
  int64_t reorder = 0;
  int64_t regroup = 1;
  int64_t cycle_lmax = 0;
  int64_t cycle_cell_pref = 0;
  int64_t segtile_threads = 0;
  int64_t lds_threads_auto = 1;
  int64_t segtile_nt = 3;
  int64_t lower_mask = LOWER_ALL;
  int64_t hold_floor = 64ll << 20;  // device bytes the library may keep whatever alpha allows
  int64_t cycle_u_depth = 4;        // positions in flight per item of the shifted cycle kernel (4, 6, 8)
  int64_t min_groups = 1024;        // workgroups a single-launch pass spreads its footprints over
  uint64_t epoch = 0;  // bumped on every change (plan caches key on it)
};
Tuning tuning_get();
int tuning_set(const char* key, int64_t value);  // 0 or EITHOT_ERR_INVALID
int tuning_value(const char* key, int64_t* value);
void tuning_apply(const Tuning& t, DeviceProfile& prof);

struct Plan {
  std::vector<Stage> stages;
  std::vector<Op> ops;
  int n_sub = 1;              // sub-tensors actually used
  int partition_dim = -1;     // 0-based input dim split into n_sub blocks
  int64_t part_len = 0;       // indices of the split dim in each of the first n_sub - 1
  int64_t last_len = 0;       // indices in the last sub-tensor (== part_len unless ragged)
  int64_t scratch_bytes = 0;  // device scratch the plan needs (0 if all LDS)
  int64_t table_bytes = 0;    // device cycle tables of its OP_CYCLE passes
  uint64_t tensor_bytes = 0;
  std::string describe() const;
  // Bytes of this plan's pass tables that are resident on device `dev`.
  int64_t device_table_bytes(int dev) const;
  // Frees this plan's pass tables on device `dev` (uploaded again on next use).
  void drop_device_tables(int dev) const;
  // The runtime's launch choices learned on this plan's calls (the
  // segmented tile's block order), created on its first run; the runtime
  // serialises access to it.
  mutable std::shared_ptr<void> launch_state;
};

// Full plan.  perm1 is 1-based.  alpha bounds the scratch (scratch <= alpha *
// tensor bytes, at least one footprint); n_sub > 1 requests the reference's
// sub-tensor partitioning along the largest dim when that dim exceeds
// floor(32/alpha): min(n_sub, dim) sub-tensors, the last one ragged when the
// count does not divide the dim.  reorder: 0 plans both decomposer
// orders and keeps the cheaper (the reference's mix rule breaks ties), 1
// pre-order, 2 post-order, 3 the mix rule.  Returns 0 or a negative
// EITHOT_ERR_*.
int make_plan(const int64_t* dims, const int* perm1, int rank, size_t elem, int n_sub,
              double alpha, const DeviceProfile& prof, int reorder, Plan& plan);

// Lowering of one batched 2-D transpose (exposed for tests).
void lower_transpose2d(int64_t B, int64_t R, int64_t C, int64_t U, uint64_t base, int stage,
                       const DeviceProfile& prof, std::vector<Op>& ops);

// ---- estimators -------------------------------------------------------------
// Reference-parity model of Nd_estimation (Nd_memory_estimation/src/
// test_inplace.cpp:18-147) on a CUDA device profile.
struct RefProfile {
  int64_t smem_per_block = 49152;   // RTX 3090
  int sms = 82;
  int max_threads_per_sm = 1536;
  int padding_blocks_per_sm = 1;    // occupancy of the 1024-thread padding kernel
};

struct RefEstimate {
  double init_bytes = 0;     // "init Extra Memory size"
  double new_ratio = 0;      // "new ratio" (only printed when init_bytes > 0)
  bool printed_new_ratio = false;
  bool decomposed = false;   // "Final Extra Memory Size" line printed
  double final_bytes = 0;
  double final_ratio = 0;
  int nb = 1;
  double alpha = 0;          // the alpha written to proper_input.txt
};

int estimate_reference(const int64_t* dims, const int* perm1, int rank, size_t elem,
                       double alpha, const RefProfile& prof, RefEstimate& out);

// Native estimate for this build's kernels: scratch bytes of the plan at the
// given alpha, the sub-tensor count, and the resulting extra-memory ratio.
int estimate_native(const int64_t* dims, const int* perm1, int rank, size_t elem, double alpha,
                    const DeviceProfile& prof, int* n_sub, double* alpha_out,
                    size_t* extra_bytes);

// ---- LDS layout of a tile footprint -------------------------------------------
// Tile rows are padded so that column reads hit distinct LDS banks: the row
// stride is an odd multiple of max(4, word) bytes.  The planner sizes tile
// footprints with the largest word a slice allows, which bounds the stride
// the kernel picks for any smaller word.
inline int64_t tile_row_stride(int64_t rowbytes, int64_t wbytes) {
  const int64_t g = wbytes > 4 ? wbytes : 4;
  int64_t rs = (rowbytes + g - 1) / g * g;
  if (((rs / g) & 1) == 0) rs += g;
  return rs;
}
inline int64_t tile_row_stride_bound(int64_t rowbytes, int64_t us) {
  int64_t w = 1;
  while (w < 16 && us % (w * 2) == 0) w *= 2;
  int64_t best = 0;
  for (int64_t g = 1; g <= w; g *= 2) {
    const int64_t r = tile_row_stride(rowbytes, g);
    if (r > best) best = r;
  }
  return best;
}

// ---- integer helpers ----------------------------------------------------------
int64_t gcd64(int64_t a, int64_t b);
int64_t modinv64(int64_t a, int64_t m);  // a^{-1} mod m (0 when m == 1)

}  // namespace eithot
