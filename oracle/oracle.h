/*
 * eithot oracle -- CPU restatement of the EITHOT in-place transposition path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (eithot_amd/, include/)
 * links, loads or calls this code.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg use it, and only as the checker.
 *
 * Conventions (ITNT/src/Nd_transpose.cpp:43-58, ITNT/src/tensor_util.cpp:91-104):
 *   dims[0] is the fastest-varying dimension; output dim k is input dim perm[k].
 *   perm0 arrays are 0-based, perm1 arrays are 1-based.
 *
 * Parity pinning of each function (see the tests that check them):
 *   or_permute / or_permute_mt   closed form (seq_tt semantics); brute force
 *                                over all perms of small tensors; v(i)=i
 *                                known-answer check (verify_inplace closed form)
 *   or_extended_gcd              the reference's gcd.cpp compiled into
 *                                oracle/_ref; closed form (Bezout) on random pairs
 *   or_decompose                 hand-traced reference plans for the benchmark
 *                                configs (tests/golden/plans.json); invariants:
 *                                stage perms compose to the target, every stage
 *                                is 21/213/132/1324, at most rank rounds
 *   or_fn_*  (functors)          each pass map is a bijection; the reference
 *                                lowering built from them equals or_permute on
 *                                every shape with m,n <= 48 and on random larger
 *                                shapes (a wrong sign/index fails that check)
 *   or_simulate_reference        end-to-end equality with or_permute on the 57
 *                                exp2 shapes (scaled) and random perms, rank 2-6
 */
#ifndef EITHOT_ORACLE_H
#define EITHOT_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAX_RANK 16

/* one low-order stage produced by the decomposer */
typedef struct {
  int rank;          /* 2, 3 or 4 */
  int perm[4];       /* 1-based low-order permutation, e.g. {1,3,2,4} */
  int64_t dims[4];   /* merged dims, fastest first */
} or_stage;

/* ---- ground truth ---------------------------------------------------- */
int or_permute(const void* in, void* out, const int64_t* dims, const int* perm0,
               int rank, int elem);
int or_permute_mt(const void* in, void* out, const int64_t* dims, const int* perm0,
                  int rank, int elem, int nthreads);
/* closed-form destination (output linear index) of input linear index */
int64_t or_dest_index(int64_t src, const int64_t* dims, const int* perm0, int rank);

/* counter-based value generator: mode 0 -> v(i) = i, mode 1 -> splitmix64(seed, i) */
uint64_t or_value(uint64_t seed, int64_t i, int mode);
void or_fill(void* buf, int64_t n, int elem, uint64_t seed, int mode, int64_t first);

/* ---- integer math ---------------------------------------------------- */
void or_extended_gcd(int64_t a, int64_t b, int64_t* gcd, int64_t* mmi);

/* ---- decomposer (permutation_util.cpp, Nd_transpose.cpp:104-250) ------ */
/* reorder: 0 mix, 1 pre-order, 2 post-order.  Returns number of stages. */
int or_decompose(const int64_t* dims, const int* perm1, int rank, int reorder,
                 or_stage* out, int max_out);

/* ---- reference functors (ITNT/src/equations.h) -------------------------- */
int64_t or_fn_c2r_rotate(int64_t m, int64_t b, int64_t i, int64_t j);
int64_t or_fn_c2r_row_shuffle(int64_t m, int64_t n, int64_t c, int64_t k, int64_t i, int64_t j);
int64_t or_fn_c2r_col_shuffle(int64_t m, int64_t n, int64_t c, int64_t i, int64_t j);
int64_t or_fn_r2c_col_shuffle(int64_t a, int64_t c, int64_t m, int64_t q, int64_t i, int64_t j);
int64_t or_fn_r2c_rotate(int64_t m, int64_t b, int64_t i, int64_t j);

/* ---- reference lowering executed on a host buffer ----------------------- */
/* Runs the reference's ITNT plan (decompose, low-order lowering, Catanzaro
 * passes, optional partition/join/padding) in place on buf.  buf must hold
 * capacity_elems elements (>= volume; padding needs the reference's headroom).
 * Returns 0 on success, <0 if the plan needs more capacity. */
int or_simulate_reference(void* buf, int64_t capacity_elems, const int64_t* dims,
                          const int* perm1, int rank, int elem, int num_block,
                          double alpha, int reorder);

/* one batched 2-D transpose (B batches of R x C units of U bytes) executed
 * with the reference's Catanzaro passes: method 0 = c2r on the R x C grid
 * (C >= R), 1 = r2c on the C x R grid.  Used to pin a product plan's stage
 * lowering independently of the product's kernels. */
int or_transpose2d_catanzaro(void* buf, int64_t B, int64_t R, int64_t C, int64_t U, int method);
int or_transpose2d_direct(void* buf, int64_t B, int64_t R, int64_t C, int64_t U);

#ifdef __cplusplus
}
#endif
#endif
