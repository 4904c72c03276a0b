/*
 * eithot -- in-place high-order tensor transposition on MI355X (gfx950).
 *
 * C-ABI of libeithot.so: plain pointers and sizes, no HIP or torch types.
 * Streams are passed as `void*` (a hipStream_t, or NULL for the null stream).
 *
 * Conventions (ITNT/src/Nd_transpose.cpp:43-58, ITNT/src/tensor_util.cpp:91-104):
 *   dims[0] is the fastest-varying dimension.  Output dim k is input dim
 *   perm[k]; `perm` is 1-based in this ABI (as on the reference CLI,
 *   ITNT/src/test_inplace.cpp:155), except in eithot_inplace_transpose which
 *   keeps the reference library's 0-based signature.
 *   Element bytes 1, 2, 4 or 8 are moved, never interpreted.
 *
 * Every entry point returns EITHOT_OK (0) or a negative EITHOT_ERR_* code and
 * never exits the process; eithot_last_error() returns the message of the
 * calling thread's last failure.  After EITHOT_ERR_DEVICE the tensor's
 * contents are indeterminate; after any other error it is untouched.
 */
#ifndef EITHOT_H
#define EITHOT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EITHOT_OK 0
#define EITHOT_ERR_INVALID (-1)
#define EITHOT_ERR_ELEM (-2)
#define EITHOT_ERR_ALPHA (-3)
#define EITHOT_ERR_DEVICE (-4)
#define EITHOT_ERR_NOMEM (-5)
#define EITHOT_ERR_UNSUPPORTED (-6)
#define EITHOT_ERR_NO_DEVICE (-7)
#define EITHOT_ERR_BUSY (-8)

/* ---- the hot path --------------------------------------------------------- */

/* Transposes the tensor at device pointer `dev` in place and returns when it
 * is done.  n_sub > 1 asks for the reference's sub-tensor partitioning along
 * the largest dim, applied when that dim exceeds floor(32/alpha): the dim is
 * cut into min(n_sub, dim) sub-tensors, the last one ragged when the count
 * does not divide the dim (the reference pads the dim up to a multiple of
 * n_sub instead, growing the tensor; here the buffer is exactly the tensor,
 * with no headroom).  n_sub <= 0 takes the estimator's choice
 * (eithot_estimate: 1, since this build's plans meet alpha without
 * partitioning).  Extra device memory stays <= alpha * tensor bytes, or one
 * pass footprint when that is larger.
 * Replaces inplace::transpose (ITNT/include/transpose.h:8,
 * ITNT/src/transpose.cpp:21-30). */
int eithot_transpose(void* dev, const int64_t* dims, const int* perm, int order,
                     size_t elem_bytes, int n_sub, double alpha);

/* Same, enqueued on `stream` without a host synchronisation.  The scratch
 * arena stays reserved until the work completes; do not free `dev` first.
 * Calls may come from several threads: the launches of calls that use the
 * arena or pass tables are enqueued one call at a time, and a call's first
 * pass waits (on the device) for the last pass of the previous arena user. */
int eithot_transpose_async(void* dev, const int64_t* dims, const int* perm, int order,
                           size_t elem_bytes, int n_sub, double alpha, void* stream);

/* Same as eithot_transpose, but allocates no device memory: no counterpart
 * in the reference; for callers that cannot allocate (tight serving loops,
 * harnesses that forbid hipMalloc).  The plan's scratch and the device
 * copies of its pass tables live in one static device array of 64 MiB per
 * GPU, held until the process ends; a plan that needs more is refused with
 * EITHOT_ERR_NOMEM before anything is launched.  Synchronous on the null
 * stream.  Calls on one GPU share the array: calls that overlap on one GPU
 * race (serialise them).  eithot_last_peak_extra_bytes() afterwards gives
 * the bytes of the array the call used. */
int eithot_transpose_noalloc(void* dev, const int64_t* dims, const int* perm, int order,
                             size_t elem_bytes, int n_sub, double alpha);

/* The reference library's own signature and semantics: 0-based permutation,
 * int dims, element offset `source`, sizeofType 1/2/4/8, num_block and the
 * extra-memory ratio.  Replaces inplace::transpose
 * (ITNT/include/transpose.h:8). */
int eithot_inplace_transpose(void* data, int source, int rank, const int* dim,
                             const int* permutation, size_t sizeofType, int num_block,
                             double large_ratio);

/* Extra device bytes the library held during the calling thread's last
 * transpose, measured rather than planned: its scratch arena as allocated
 * plus every pass table resident on the device for the plans it caches (this
 * call's and earlier ones').  Each call first frees other plans' tables,
 * least recently used first, and shrinks an arena larger than it needs, so
 * that this stays within the largest of alpha * tensor bytes, the call's
 * own need (when one pass footprint exceeds that) and the "hold_floor"
 * tuning value (64 MiB, which spares calls on small tensors from evicting
 * and re-uploading each other's tables).  0 when nothing is held.
 * Server pools and batch scratch belong to their handles and are not
 * counted. */
size_t eithot_last_peak_extra_bytes(void);

/* Message of the calling thread's last failure ("" if none). */
const char* eithot_last_error(void);

/* Frees the scratch arena, the cached plans and their device tables of the
 * current device, after waiting for its work to finish. */
void eithot_release(void);

/* ---- planner / estimator (host only, no GPU needed) ------------------------ */

/* The Nd_estimation planner for this build's kernels: sub-tensor count, the
 * extra-memory ratio the plan needs at most, and its scratch bytes.
 * Replaces the dry-run inplace::transpose of
 * Nd_memory_estimation/include/transpose.h:8 as used by
 * Nd_memory_estimation/src/test_inplace.cpp:18-147. */
int eithot_estimate(const int64_t* dims, const int* perm, int order, size_t elem_bytes,
                    double alpha, int* n_sub_out, double* alpha_out, size_t* extra_bytes_out);

/* The reference estimator's own policy and size model evaluated on a CUDA
 * device profile (shared memory per block, SM count, max threads per SM,
 * occupancy of its padding kernel), reproducing what Nd_estimation prints
 * (Nd_memory_estimation/src/test_inplace.cpp:18-147).  Outputs: init and
 * final extra bytes, nb, alpha written to proper_input.txt, and flags
 * (bit0: "new ratio" printed, bit1: "Final Extra Memory" printed). */
typedef struct {
  int64_t smem_per_block;
  int sms;
  int max_threads_per_sm;
  int padding_blocks_per_sm;
} eithot_ref_profile;

int eithot_estimate_reference(const int64_t* dims, const int* perm, int order, size_t elem_bytes,
                              double alpha, const eithot_ref_profile* prof, double* init_bytes,
                              double* new_ratio, double* final_bytes, double* final_ratio,
                              int* nb, double* alpha_out, int* flags);

/* Decomposer rounds (permutation_util.cpp, Nd_transpose.cpp:146-250):
 * reorder 0 = mix, 1 = pre-order, 2 = post-order.  Per stage: rank, a 1-based
 * low-order perm (4 ints) and merged dims (4 int64).  Returns the stage count. */
int eithot_decompose(const int64_t* dims, const int* perm, int order, int reorder,
                     int* stage_rank, int* stage_perm, int64_t* stage_dims, int max_stages);

/* One device pass of a plan, for introspection and tests. */
typedef struct {
  int kind;      /* 1 tile, 2 column, 3 row, 4 tile by cycle following,
                    5 fused block permutation (B blocks of R cells of U bytes),
                    6 segmented tile (see eithot_plan_segments) */
  int map;       /* 0 none, 1 rot, 2 rotinv, 3 colsh, 4 colshinv, 5 rowg, 6 rowd */
  int staged;    /* 1 when executed through scratch */
  int stage;     /* decomposer stage, -1 for partition/join */
  uint64_t base; /* byte offset */
  int64_t B, R, C, U, us, W;
  int64_t m, n, c, a, b, k, binv;
} eithot_op_desc;

/* Fills up to max_ops pass descriptors of the plan for the gfx950 profile
 * (lds_max bytes of LDS per workgroup; <= 0 selects 160 KiB) and returns
 * the pass count; *scratch_bytes receives the plan's extra device memory
 * (scratch plus the cycle tables of its cycle-following passes). */
int eithot_plan_describe(const int64_t* dims, const int* perm, int order, size_t elem_bytes,
                         int n_sub, double alpha, int64_t lds_max, eithot_op_desc* ops,
                         int max_ops, int64_t* scratch_bytes);

/* Sub-tensor partition of the plan: *n_sub_used sub-tensors along 0-based
 * input dim *dim (-1 and 1 when the plan does not partition), the first
 * n_sub_used - 1 holding *part_len indices of that dim and the last
 * *last_len.  Returns 0 or a negative EITHOT_ERR_*.  The reference prints
 * the count as "number of sub_tensors" (ITNT/src/test_inplace.cpp:34). */
int eithot_plan_partition(const int64_t* dims, const int* perm, int order, size_t elem_bytes,
                          int n_sub, double alpha, int* n_sub_used, int* dim, int64_t* part_len,
                          int64_t* last_len);

/* Gather table of pass op_index when it is a fused block permutation:
 * output cell o of every block takes input cell src[o].  Returns the cell
 * count (copies at most max_cells) or a negative EITHOT_ERR_*. */
int64_t eithot_plan_block_table(const int64_t* dims, const int* perm, int order,
                                size_t elem_bytes, int n_sub, double alpha, int64_t lds_max,
                                int op_index, uint32_t* src, int64_t max_cells);

/* Cell maps of pass op_index when it is a cycle-following pass, in pass
 * order, 4 values each (bc, R, C, m): blocks of bc cells of U bytes, each an
 * R x C matrix of m-cell units, transposed.  Returns the map count (copies
 * at most max_maps) or a negative EITHOT_ERR_*. */
int eithot_plan_cycle_maps(const int64_t* dims, const int* perm, int order, size_t elem_bytes,
                           int n_sub, double alpha, int64_t lds_max, int op_index,
                           int64_t* maps, int max_maps);

/* Segments of pass op_index when it is a segmented tile pass (kind 6):
 * B matrices of R x C units of U bytes, each transposed in place, whose
 * bytes (row-major) lie in segments of *segb bytes; segment s of matrix b
 * starts at byte outer(b) + inner(s) of the tensor, where inner(s) sums
 * digit * stride over the mixed-radix digits of s (fastest first) and
 * outer(b) the same over b.  inner and outer receive 6 (radix, stride)
 * pairs each (unused: radix 1).  Returns n_inner + 16 * n_outer, or a
 * negative EITHOT_ERR_*.  No counterpart in the reference. */
int eithot_plan_segments(const int64_t* dims, const int* perm, int order, size_t elem_bytes,
                         int n_sub, double alpha, int64_t lds_max, int op_index, int64_t* segb,
                         int64_t* inner, int64_t* outer);

/* ---- tuning -------------------------------------------------------------------
 * No counterpart in the reference, whose equivalents are compile-time
 * constants (REORDER_TYPE, Nd_transpose.cpp:28; tier thresholds in col_op.cu
 * and row_op.cu).  Process-wide knobs, applied to plans made after the call:
 *   "reorder"        decomposer order: 0 plans pre and post and keeps the
 *                    cheaper (default), 1 pre, 2 post, 3 mix (the reference's rule)
 *   "regroup"        with reorder 0, also plans a cell pass, one tile pass and
 *                    a cell pass around the moving fastest dims (1; 0 = off)
 *   "lds_target"     LDS bytes a workgroup packs footprints up to (65536)
 *   "tile_fp_target" tile footprints above this are sliced (65536; 0 = off)
 *   "slice_min"      smallest slice such slicing may choose (128)
 *   "stage_chunk"    > 0: bytes of scratch one staged group may fill (0)
 *   "cycle_min_cell" smallest cell (bytes) moved by cycle following (128; 0 = off)
 *   "fuse_max_block" largest block of a fused tile permutation (65536; 0 = off)
 *   "block_digits"   fused blocks compute source cells from digits (1) or read
 *                    a gather table (0)
 *   "block_pad"      fused blocks pad their LDS copy where a bank-conflict
 *                    model says it pays (1) or never (0)
 *   "cycle_variant"  cycle-following kernel shape (0..7; 1..3 only for cells of
 *                    at least 64 words).  0 (default) picks: shifted 16-byte
 *                    chunks for cells that are not whole 16-byte words,
 *                    non-temporal loads and stores (as 5) for cells of at
 *                    least 512 bytes, non-temporal stores (as 4) for
 *                    smaller 16-byte-word cells in passes of at least
 *                    512 MiB, else plain 16-byte words
 *   "cycle_lmax"     > 0: cycle segments of at most this many positions
 *                    (0: chosen per pass)
 *   "cycle_u_depth"  positions in flight per item when cycle following moves
 *                    cells that are not whole 16-byte words (4 default, 6, 8)
 *   "min_groups"     workgroups a pass launched on its own spreads its LDS
 *                    footprints over, at least (1024; 1..2^20); batched
 *                    launches pack footprints to the LDS target instead
 *   "cycle_cell_pref" cycle passes of cells of 256 B and more cost less in
 *                    the planner (0 default: not faster on MI355X so far)
 *   "cycle_word_cost" cycle passes over cells that are not whole 16-byte
 *                    words cost more in the planner (1) or the same (0)
 *   "threads", "threads_big"  workgroup sizes (256, 512; multiples of 64, at most 512)
 *   "segtile_threads" workgroup size of segmented tile passes (0: one round of
 *                    8 loads per thread, at most 512; else a multiple of 64)
 *   "segtile_nt"     segmented tile passes of 16-byte words: 3 non-temporal loads
 *                    and stores (default), 2 non-temporal stores, 0 neither
 *   "lds_threads_auto" tile, column, row and block workgroups sized for one round of
 *                    8 loads per thread (1) or from "threads"/"threads_big" (0)
 *   "xcd_remap"      keep consecutive footprints on one XCD (1); with 1, a
 *                    plan's second and third calls run its segmented-tile
 *                    passes in each block order (XCD-contiguous, then
 *                    round-robin) between timing events and later calls
 *                    keep the faster (not under graph capture); 0 forces
 *                    round-robin
 *   "hold_floor"     device bytes (arena and pass tables) the library may keep
 *                    after a call even when alpha * tensor bytes is smaller
 *                    (64 MiB; 0: alpha alone bounds them)
 *   "lower_mask"     lowering options of each 2-D transpose (31 = all): 1 LDS
 *                    tile, 2 Catanzaro column/row passes, 4 hierarchical split,
 *                    8 tile through scratch, 16 cycle following; a transpose
 *                    no allowed option can lower fails with EITHOT_ERR_ALPHA
 * Returns EITHOT_ERR_INVALID for an unknown key or an out-of-range value. */
int eithot_set_tuning(const char* key, int64_t value);
int eithot_get_tuning(const char* key, int64_t* value);

/* ---- measurement helpers ----------------------------------------------------- */

/* When enabled, the next transposes record per-pass GPU time (HIP events on
 * the launching stream); eithot_last_pass_times copies up to max of them (ms). */
void eithot_set_pass_timing(int enable);
int eithot_last_pass_times(float* ms, int max);

/* Device-side pattern fill and check, for tensors too large to copy to the
 * host: fills v(seed, first+i) (mode 0: v = i, mode 1: 64-bit mix of seed
 * and i) and counts output elements whose value differs from the permuted
 * pattern. */
int eithot_fill_pattern(void* dev, uint64_t n, size_t elem_bytes, uint64_t seed, int mode,
                        uint64_t first, void* stream);
int eithot_verify_pattern(const void* dev, const int64_t* dims, const int* perm, int order,
                          size_t elem_bytes, uint64_t seed, int mode,
                          unsigned long long* mismatches);

/* ---- batched transposition (many small tensors per launch) -------------------- */

typedef struct eithot_batch eithot_batch;

eithot_batch* eithot_batch_create(void);
/* Adds one tensor (device pointer, 1-based perm).  Returns its index. */
int eithot_batch_add(eithot_batch* b, void* dev, const int64_t* dims, const int* perm, int order,
                     size_t elem_bytes);
/* Plans every tensor, groups passes by stage index into persistent launches,
 * runs the autotuner (tune != 0) and captures the launches in a HIP graph. */
int eithot_batch_build(eithot_batch* b, double alpha, int tune);
/* Replays the batch on `stream` (NULL: null stream) and waits when sync. */
int eithot_batch_run(eithot_batch* b, void* stream, int sync);
/* Launch count per replay and the batch's scratch bytes. */
int eithot_batch_info(const eithot_batch* b, int* launches, int64_t* scratch_bytes,
                      int* graph_captured);
void eithot_batch_destroy(eithot_batch* b);

/* ---- serving (continuous batching of transpose requests) ----------------------- */

typedef struct eithot_server eithot_server;

/* scratch_pool_bytes is the shared extra-memory budget all in-flight requests
 * are admitted against (used in whole 256-byte granules); max_inflight caps
 * concurrent requests; nstreams is the number of device streams requests are
 * spread over. */
eithot_server* eithot_server_create(int64_t scratch_pool_bytes, int max_inflight, int nstreams);
/* Enqueues a request on resident tensor `dev`; returns a request id >= 0.
 * Requests are admitted in submission order.  Requests whose bytes overlap
 * run one after the other in that order (a tensor may be transposed and then
 * transposed back); others may run concurrently.  A request whose scratch,
 * rounded up to 256 bytes, exceeds the pool is refused (EITHOT_ERR_ALPHA). */
int64_t eithot_server_submit(eithot_server* s, void* dev, const int64_t* dims, const int* perm,
                             int order, size_t elem_bytes);
/* Admits queued requests while slots and scratch allow and retires finished
 * ones; returns the number of requests completed by this call.  With wait
 * != 0 it blocks until at least one in-flight request completes. */
int eithot_server_poll(eithot_server* s, int wait);
/* Status of a request: 0 queued, 1 running, 2 done, <0 error code.
 * On done, *latency_ms holds submit-to-completion time (host clock). */
int eithot_server_status(eithot_server* s, int64_t id, double* latency_ms);
/* Blocks until every submitted request has completed; returns an error
 * instead of waiting forever if a queued request can never be admitted. */
int eithot_server_drain(eithot_server* s);
void eithot_server_destroy(eithot_server* s);

#ifdef __cplusplus
}
#endif
#endif /* EITHOT_H */
