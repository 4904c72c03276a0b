/*
 * eithot oracle -- CPU restatement of the EITHOT in-place transposition path.
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker, never by the product.
 *
 * Each function cites the reference file:line it restates.  All index
 * arithmetic is 64-bit (the reference's int32 arithmetic overflows above
 * 2^31 elements; the restatement keeps the mathematics, not the overflow).
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* value generator                                                           */
/* ------------------------------------------------------------------------ */

static inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

/* mode 0: v(i) = i (ITNT/src/tensor_util.cpp:53-56 init_data);
 * mode 1: random bits, a pure function of (seed, i) so any output element
 * can be re-derived without a second copy of the tensor. */
uint64_t or_value(uint64_t seed, int64_t i, int mode) {
  if (mode == 0) return (uint64_t)i;
  return splitmix64(splitmix64(seed) ^ (uint64_t)i * 0xD1B54A32D192ED03ull);
}

void or_fill(void* buf, int64_t n, int elem, uint64_t seed, int mode, int64_t first) {
  unsigned char* p = (unsigned char*)buf;
  for (int64_t i = 0; i < n; ++i) {
    uint64_t v = or_value(seed, first + i, mode);
    memcpy(p + i * elem, &v, (size_t)(elem < 8 ? elem : 8));
    if (elem > 8) memset(p + i * elem + 8, 0, (size_t)(elem - 8));
  }
}

/* ------------------------------------------------------------------------ */
/* out-of-place permutation (ITNT/src/tensor_util.cpp:91-104 seq_tt)         */
/* ------------------------------------------------------------------------ */

/* output stride of every input dim */
static void out_strides_of_input_dims(const int64_t* dims, const int* perm0, int rank,
                                      int64_t* ostride) {
  int64_t s = 1;
  for (int k = 0; k < rank; ++k) {
    ostride[perm0[k]] = s;
    s *= dims[perm0[k]];
  }
}

int64_t or_dest_index(int64_t src, const int64_t* dims, const int* perm0, int rank) {
  int64_t os[OR_MAX_RANK];
  out_strides_of_input_dims(dims, perm0, rank, os);
  int64_t d = 0;
  for (int i = 0; i < rank; ++i) {
    int64_t x = src % dims[i];
    src /= dims[i];
    d += x * os[i];
  }
  return d;
}

static int check_perm(const int* perm0, int rank) {
  int seen[OR_MAX_RANK] = {0};
  if (rank < 1 || rank > OR_MAX_RANK) return -1;
  for (int k = 0; k < rank; ++k) {
    if (perm0[k] < 0 || perm0[k] >= rank || seen[perm0[k]]) return -1;
    seen[perm0[k]] = 1;
  }
  return 0;
}

int or_permute(const void* in, void* out, const int64_t* dims, const int* perm0, int rank,
               int elem) {
  if (check_perm(perm0, rank)) return -1;
  int64_t os[OR_MAX_RANK], idx[OR_MAX_RANK] = {0};
  out_strides_of_input_dims(dims, perm0, rank, os);
  int64_t vol = 1;
  for (int i = 0; i < rank; ++i) vol *= dims[i];
  const unsigned char* src = (const unsigned char*)in;
  unsigned char* dst = (unsigned char*)out;
  int64_t d = 0;
  for (int64_t s = 0; s < vol; ++s) {
    memcpy(dst + d * elem, src + s * elem, (size_t)elem);
    /* odometer over input dims, tracking the output offset */
    for (int i = 0; i < rank; ++i) {
      if (++idx[i] < dims[i]) { d += os[i]; break; }
      d -= (dims[i] - 1) * os[i];
      idx[i] = 0;
    }
  }
  return 0;
}

/* multithreaded, cache-tiled out-of-place permute (the CPU baseline).
 * Input dim 0 (contiguous in the input) and input dim `o` = perm0[0]
 * (contiguous in the output) form a plane that is transposed in 32x32
 * tiles; all other dims are an outer odometer split across threads. */
typedef struct {
  uint64_t lo, hi;
} or_u128;

typedef struct {
  const unsigned char* in;
  unsigned char* out;
  const int64_t* dims;
  int rank, elem, o;
  int64_t is[OR_MAX_RANK], os[OR_MAX_RANK];
  int outer[OR_MAX_RANK];
  int nouter;
  int64_t nwork, wbegin, wend;
  int64_t nblk_o;
} mt_job;

static void* mt_worker(void* arg) {
  mt_job* J = (mt_job*)arg;
  const int64_t T = 32;
  int elem = J->elem;
  int64_t d0 = J->dims[0];
  int64_t dO = J->o == 0 ? 1 : J->dims[J->o];
  int64_t os0 = J->os[0], isO = J->o == 0 ? 0 : J->is[J->o];
  for (int64_t w = J->wbegin; w < J->wend; ++w) {
    int64_t ob = w % J->nblk_o, rem = w / J->nblk_o;
    int64_t ibase = 0, obase = 0;
    for (int t = 0; t < J->nouter; ++t) {
      int dd = J->outer[t];
      int64_t x = rem % J->dims[dd];
      rem /= J->dims[dd];
      ibase += x * J->is[dd];
      obase += x * J->os[dd];
    }
    if (J->o == 0) { /* fastest dim unchanged: contiguous runs */
      memcpy(J->out + obase * elem, J->in + ibase * elem, (size_t)(d0 * elem));
      continue;
    }
    int64_t o0 = ob * T, o1 = o0 + T < dO ? o0 + T : dO;
    for (int64_t i0 = 0; i0 < d0; i0 += T) {
      int64_t i1 = i0 + T < d0 ? i0 + T : d0;
      /* element copies of a width known at compile time for the common sizes */
#define OR_TILE(TYPE)                                                              \
  for (int64_t y = o0; y < o1; ++y)                                                \
    for (int64_t x = i0; x < i1; ++x) {                                            \
      TYPE v;                                                                      \
      memcpy(&v, J->in + (ibase + x + y * isO) * elem, sizeof(TYPE));              \
      memcpy(J->out + (obase + x * os0 + y) * elem, &v, sizeof(TYPE));             \
    }
      if (elem == 1) { OR_TILE(uint8_t) }
      else if (elem == 2) { OR_TILE(uint16_t) }
      else if (elem == 4) { OR_TILE(uint32_t) }
      else if (elem == 8) { OR_TILE(uint64_t) }
      else if (elem == 16) { OR_TILE(or_u128) }
      else
        for (int64_t y = o0; y < o1; ++y)
          for (int64_t x = i0; x < i1; ++x)
            memcpy(J->out + (obase + x * os0 + y) * elem, J->in + (ibase + x + y * isO) * elem,
                   (size_t)elem);
#undef OR_TILE
    }
  }
  return NULL;
}

int or_permute_mt(const void* in, void* out, const int64_t* dims, const int* perm0, int rank,
                  int elem, int nthreads) {
  if (check_perm(perm0, rank)) return -1;
  /* leading dims that stay in place move with the element: fold them into
   * it (up to 4 KiB), so that the tiled plane is formed by dims that move */
  int m = 0;
  int64_t fold = 1;
  while (m < rank - 1 && perm0[m] == m && (int64_t)elem * fold * dims[m] <= 4096) {
    fold *= dims[m];
    ++m;
  }
  if (m > 0) {
    int64_t d2[OR_MAX_RANK];
    int p2[OR_MAX_RANK];
    for (int k = m; k < rank; ++k) {
      d2[k - m] = dims[k];
      p2[k - m] = perm0[k] - m;
    }
    return or_permute_mt(in, out, d2, p2, rank - m, (int)(elem * fold), nthreads);
  }
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  mt_job base;
  memset(&base, 0, sizeof(base));
  base.in = (const unsigned char*)in;
  base.out = (unsigned char*)out;
  base.dims = dims;
  base.rank = rank;
  base.elem = elem;
  base.o = perm0[0];
  int64_t s = 1;
  for (int i = 0; i < rank; ++i) { base.is[i] = s; s *= dims[i]; }
  out_strides_of_input_dims(dims, perm0, rank, base.os);
  int64_t nouter_vol = 1;
  for (int i = 1; i < rank; ++i)
    if (i != base.o) { base.outer[base.nouter++] = i; nouter_vol *= dims[i]; }
  base.nblk_o = base.o == 0 ? 1 : (dims[base.o] + 31) / 32;
  base.nwork = nouter_vol * base.nblk_o;
  pthread_t th[256];
  mt_job jobs[256];
  int n = nthreads;
  if (n > base.nwork) n = (int)base.nwork;
  if (n < 1) n = 1;
  for (int t = 0; t < n; ++t) {
    jobs[t] = base;
    jobs[t].wbegin = base.nwork * t / n;
    jobs[t].wend = base.nwork * (t + 1) / n;
    if (n == 1) { mt_worker(&jobs[t]); return 0; }
    pthread_create(&th[t], NULL, mt_worker, &jobs[t]);
  }
  for (int t = 0; t < n; ++t) pthread_join(th[t], NULL);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* integer math (ITNT/src/gcd.cpp:5-32)                                      */
/* ------------------------------------------------------------------------ */

void or_extended_gcd(int64_t a, int64_t b, int64_t* gcd, int64_t* mmi) {
  int64_t x = 0, lastx = 1, origb = b;
  while (b != 0) {
    int64_t q = a / b, r = a % b;
    a = b;
    b = r;
    int64_t nx = lastx - q * x;
    lastx = x;
    x = nx;
  }
  *gcd = a;
  *mmi = 0;
  if (a == 1) *mmi = lastx < 0 ? lastx + origb : lastx;
}

static int64_t emod(int64_t x, int64_t m) {
  int64_t r = x % m;
  return r < 0 ? r + m : r;
}

/* ------------------------------------------------------------------------ */
/* functors (ITNT/src/equations.h)                                           */
/* ------------------------------------------------------------------------ */

/* equations.h:319-337  i' = (i + floor(j/b)) mod m */
int64_t or_fn_c2r_rotate(int64_t m, int64_t b, int64_t i, int64_t j) { return (i + j / b) % m; }

/* equations.h:339-386  row gather of the C2R row shuffle */
int64_t or_fn_c2r_row_shuffle(int64_t m, int64_t n, int64_t c, int64_t k, int64_t i, int64_t j) {
  int64_t b = n / c;
  int64_t f = j + i * (n - 1);
  if (!(i - (j % c) <= m - c)) f += m;
  return (k * (f / c)) % b + (f % c) * b;
}

/* equations.h:388-405  i' = (i*n - floor(i*c/m) + j) mod m */
int64_t or_fn_c2r_col_shuffle(int64_t m, int64_t n, int64_t c, int64_t i, int64_t j) {
  return emod(i * n - (i * c) / m + j, m);
}

/* equations.h:410-438 */
int64_t or_fn_r2c_col_shuffle(int64_t a, int64_t c, int64_t m, int64_t q, int64_t i, int64_t j) {
  int64_t x = (i + m - (j % m)) % m;
  int64_t cm1 = c - 1;
  int64_t t1 = a * ((cm1 * x) % c);
  int64_t t2 = (((cm1 + x) / c) * q) % a;
  return t1 + t2;
}

/* equations.h:514-534  i' = (i + m - floor(j/b)) mod m */
int64_t or_fn_r2c_rotate(int64_t m, int64_t b, int64_t i, int64_t j) { return (i + m - j / b) % m; }

typedef enum {
  FN_C2R_ROT,
  FN_C2R_ROWSH,
  FN_C2R_COLSH,
  FN_R2C_COLSH,
  FN_R2C_ROT,
  FN_132_ROWPERM,  /* equations.h:261-281 */
  FN_213_ROWSH,    /* equations.h:285-311 */
  FN_1324_ROWPERM, /* equations.h:140-162 */
  FN_1324_ROWSH    /* equations.h:164-191 */
} fn_kind;

typedef struct {
  fn_kind k;
  int64_t p[5];
} fn_t;

static int64_t fn_eval(const fn_t* f, int64_t i, int64_t j) {
  const int64_t* p = f->p;
  switch (f->k) {
    case FN_C2R_ROT: return or_fn_c2r_rotate(p[0], p[1], i, j);
    case FN_C2R_ROWSH: return or_fn_c2r_row_shuffle(p[0], p[1], p[2], p[3], i, j);
    case FN_C2R_COLSH: return or_fn_c2r_col_shuffle(p[0], p[1], p[2], i, j);
    case FN_R2C_COLSH: return or_fn_r2c_col_shuffle(p[0], p[1], p[2], p[3], i, j);
    case FN_R2C_ROT: return or_fn_r2c_rotate(p[0], p[1], i, j);
    case FN_132_ROWPERM: { /* (ik, j) -> i + k*d2 ; p = {d3, d2} */
      int64_t d3 = p[0], d2 = p[1];
      return i / d3 + (i % d3) * d2;
    }
    case FN_213_ROWSH: { /* j -> j/d2 + (j mod d2)*d1 ; p = {d1, d2} */
      int64_t d1 = p[0], d2 = p[1];
      return j / d2 + (j % d2) * d1;
    }
    case FN_1324_ROWPERM: { /* p = {d3, d2} */
      int64_t d3 = p[0], d2 = p[1];
      int64_t l = i / (d2 * d3), ii = (i / d3) % d2, kk = i % d3;
      return ii + kk * d2 + l * d2 * d3;
    }
    case FN_1324_ROWSH: { /* p = {d2, d3, d1} */
      int64_t d2 = p[0], d3 = p[1], d1 = p[2];
      return (j % d1) + (j / d1 / d3) * d1 + ((j / d1) % d3) * d1 * d2;
    }
  }
  return 0;
}

static fn_t mkfn(fn_kind k, int64_t a, int64_t b, int64_t c, int64_t d, int64_t e) {
  fn_t f;
  f.k = k;
  f.p[0] = a; f.p[1] = b; f.p[2] = c; f.p[3] = d; f.p[4] = e;
  return f;
}

/* ------------------------------------------------------------------------ */
/* pass kernels' semantics (ITNT/src/col_op.cu, ITNT/src/row_op.cu)          */
/* ------------------------------------------------------------------------ */

typedef unsigned char u8;

/* col_op.cu:443-529  data[i][j] = old[fn(i,j)][j], i < d2 rows, j < d1 cols */
static void col_2d(u8* d, int es, const fn_t* f, int64_t d1, int64_t d2) {
  u8* tmp = (u8*)malloc((size_t)(d2 * es));
  for (int64_t j = 0; j < d1; ++j) {
    for (int64_t i = 0; i < d2; ++i) memcpy(tmp + i * es, d + (i * d1 + j) * es, (size_t)es);
    for (int64_t i = 0; i < d2; ++i)
      memcpy(d + (i * d1 + j) * es, tmp + fn_eval(f, i, j) * es, (size_t)es);
  }
  free(tmp);
}

/* col_op.cu:335-425  batched over d3 */
static void col_213(u8* d, int es, const fn_t* f, int64_t d1, int64_t d2, int64_t d3) {
  for (int64_t k = 0; k < d3; ++k) col_2d(d + k * d1 * d2 * es, es, f, d1, d2);
}

/* col_op.cu:140-319  data[k][i][j] = old[k][fn(i,k)][j] */
static void col_132(u8* d, int es, const fn_t* f, int64_t d1, int64_t d2, int64_t d3) {
  u8* tmp = (u8*)malloc((size_t)(d1 * d2 * es));
  for (int64_t k = 0; k < d3; ++k) {
    u8* s = d + k * d1 * d2 * es;
    memcpy(tmp, s, (size_t)(d1 * d2 * es));
    for (int64_t i = 0; i < d2; ++i)
      memcpy(s + i * d1 * es, tmp + fn_eval(f, i, k) * d1 * es, (size_t)(d1 * es));
  }
  free(tmp);
}

/* row_op.cu:687-883  row i: gather data[i][j] = old[i][fn_i(j)] / scatter */
static void row_2d(u8* d, int es, const fn_t* f, int64_t d1, int64_t d2, int scatter) {
  u8* tmp = (u8*)malloc((size_t)(d1 * es));
  for (int64_t i = 0; i < d2; ++i) {
    u8* row = d + i * d1 * es;
    memcpy(tmp, row, (size_t)(d1 * es));
    for (int64_t j = 0; j < d1; ++j) {
      int64_t g = fn_eval(f, i, j);
      if (scatter) memcpy(row + g * es, tmp + j * es, (size_t)es);
      else memcpy(row + j * es, tmp + g * es, (size_t)es);
    }
  }
  free(tmp);
}

/* row_op.cu:455-671  batched over d3 with the row index i < d2 */
static void row_213(u8* d, int es, const fn_t* f, int64_t d1, int64_t d2, int64_t d3,
                    int scatter) {
  for (int64_t k = 0; k < d3; ++k) row_2d(d + k * d1 * d2 * es, es, f, d1, d2, scatter);
}

/* row_op.cu:284-394  row i made of d3 d1-vectors at j + i*d1 + k*d1*d2 */
static void row_132(u8* d, int es, const fn_t* f, int64_t d1, int64_t d2, int64_t d3,
                    int scatter) {
  u8* tmp = (u8*)malloc((size_t)(d1 * d3 * es));
  for (int64_t i = 0; i < d2; ++i) {
    for (int64_t k = 0; k < d3; ++k)
      memcpy(tmp + k * d1 * es, d + (i * d1 + k * d1 * d2) * es, (size_t)(d1 * es));
    for (int64_t k = 0; k < d3; ++k) {
      int64_t g = fn_eval(f, i, k);
      if (scatter)
        memcpy(d + (i * d1 + g * d1 * d2) * es, tmp + k * d1 * es, (size_t)(d1 * es));
      else
        memcpy(d + (i * d1 + k * d1 * d2) * es, tmp + g * d1 * es, (size_t)(d1 * es));
    }
  }
  free(tmp);
}

/* ------------------------------------------------------------------------ */
/* Catanzaro C2R / R2C drivers (ITNT/src/2dtranspose.cpp:128-183,            */
/* 3dtranspose.cpp:96-128, 249-297)                                           */
/* ------------------------------------------------------------------------ */

typedef struct {
  int64_t c, k, q, a, b;
} gcd_params;

/* c2r constants for a matrix of d2 rows and d1 columns */
static void c2r_consts(int64_t d1, int64_t d2, gcd_params* g) {
  int64_t c, t, k;
  or_extended_gcd(d2, d1, &c, &t);
  if (c > 1) { int64_t t2; or_extended_gcd(d2 / c, d1 / c, &t2, &k); }
  else k = t;
  g->c = c;
  g->k = k;
  g->a = d2 / c;
  g->b = d1 / c;
  g->q = 0;
}

/* r2c constants for a matrix viewed as d2 rows and d1 columns */
static void r2c_consts(int64_t d1, int64_t d2, gcd_params* g) {
  int64_t c, t, q, k;
  or_extended_gcd(d1, d2, &c, &t);
  if (c > 1) { int64_t t2; or_extended_gcd(d1 / c, d2 / c, &t2, &q); }
  else q = t;
  g->a = d2 / c;
  g->b = d1 / c;
  or_extended_gcd(d2, d1, &c, &t);
  if (c > 1) { int64_t t2; or_extended_gcd(d2 / c, d1 / c, &t2, &k); }
  else k = t;
  g->c = c;
  g->k = k;
  g->q = q;
}

/* shape selector for the generic C2R/R2C pass sequence */
enum { SHAPE_2D, SHAPE_213, SHAPE_132 };

static void do_col(int shape, u8* d, int es, const fn_t* f, int64_t d1, int64_t d2, int64_t d3) {
  if (shape == SHAPE_2D) col_2d(d, es, f, d1, d2);
  else if (shape == SHAPE_213) col_213(d, es, f, d1, d2, d3);
  else col_132(d, es, f, d1, d2, d3);
}

static void do_row(int shape, u8* d, int es, const fn_t* f, int64_t d1, int64_t d2, int64_t d3,
                   int scatter) {
  if (shape == SHAPE_2D) row_2d(d, es, f, d1, d2, scatter);
  else if (shape == SHAPE_213) row_213(d, es, f, d1, d2, d3, scatter);
  else row_132(d, es, f, d1, d2, d3, scatter);
}

/* m = rows of the grid the passes act on, n = its columns.  For SHAPE_2D and
 * SHAPE_213 the grid is (d2 rows, d1 cols); for SHAPE_132 it is (d2 rows,
 * d3 cols) of d1-vectors (3dtranspose.cpp:249-267). */
static void c2r_generic(int shape, u8* d, int es, int64_t d1, int64_t d2, int64_t d3) {
  int64_t m = d2, n = shape == SHAPE_132 ? d3 : d1;
  gcd_params g;
  c2r_consts(n, m, &g);
  if (g.c > 1) {
    fn_t r = mkfn(FN_C2R_ROT, m, g.b, 0, 0, 0);
    do_col(shape, d, es, &r, d1, d2, d3);
  }
  fn_t rs = mkfn(FN_C2R_ROWSH, m, n, g.c, g.k, 0);
  do_row(shape, d, es, &rs, d1, d2, d3, 0);
  fn_t cs = mkfn(FN_C2R_COLSH, m, n, g.c, 0, 0);
  do_col(shape, d, es, &cs, d1, d2, d3);
}

static void r2c_generic(int shape, u8* d, int es, int64_t d1, int64_t d2, int64_t d3) {
  int64_t m = d2, n = shape == SHAPE_132 ? d3 : d1;
  gcd_params g;
  r2c_consts(n, m, &g);
  fn_t cs = mkfn(FN_R2C_COLSH, g.a, g.c, m, g.q, 0);
  do_col(shape, d, es, &cs, d1, d2, d3);
  fn_t rs = mkfn(FN_C2R_ROWSH, m, n, g.c, g.k, 0);
  do_row(shape, d, es, &rs, d1, d2, d3, 1);
  if (g.c > 1) {
    fn_t r = mkfn(FN_R2C_ROT, m, g.b, 0, 0, 0);
    do_col(shape, d, es, &r, d1, d2, d3);
  }
}

/* ------------------------------------------------------------------------ */
/* low-order lowering (ITNT/src/2dtranspose.cpp:38-123, 3dtranspose.cpp,     */
/* 4dtranspose.cpp:69-187)                                                   */
/* ------------------------------------------------------------------------ */

static int msb64(int64_t x) { /* util.cpp:8-10  (int)log2(x) */
  int r = -1;
  while (x > 0) { ++r; x >>= 1; }
  return r;
}

/* 2dtranspose.cpp:38-123 (live branch :119-122) */
static void lo_2d(u8* d, int es, int64_t d1, int64_t d2) {
  if (d1 >= d2) c2r_generic(SHAPE_2D, d, es, d1, d2, 1);
  else r2c_generic(SHAPE_2D, d, es, d2, d1, 1);
}

/* 3dtranspose.cpp:70-92 */
static void lo_231(u8* d, int es, int64_t d1, int64_t d2, int64_t d3) { lo_2d(d, es, d1, d2 * d3); }
static void lo_312(u8* d, int es, int64_t d1, int64_t d2, int64_t d3) { lo_2d(d, es, d1 * d2, d3); }

/* 3dtranspose.cpp:131-242 */
static void lo_213(u8* d, int es, int64_t d1, int64_t d2, int64_t d3) {
  if (d3 >= 8) {
    fn_t f = mkfn(FN_213_ROWSH, d1, d2, 0, 0, 0);
    row_2d(d, es, &f, d1 * d2, d3, 0);
  } else if (d1 > d2) {
    c2r_generic(SHAPE_213, d, es, d1, d2, d3);
  } else {
    r2c_generic(SHAPE_213, d, es, d2, d1, d3);
  }
}

/* 3dtranspose.cpp:300-499 */
static void lo_132(u8* d, int es, int64_t d1, int64_t d2, int64_t d3) {
  if (d1 >= 48) {
    fn_t f = mkfn(FN_132_ROWPERM, d3, d2, 0, 0, 0);
    col_2d(d, es, &f, d1, d2 * d3);
  } else if (d1 > 12) {
    if (d2 > d3) c2r_generic(SHAPE_132, d, es, d1, d3, d2);
    else r2c_generic(SHAPE_132, d, es, d1, d2, d3);
  } else {
    if (d2 >= d3) {
      lo_312(d, es, d1, d2, d3);
      fn_t f = mkfn(FN_213_ROWSH, d3, d1, 0, 0, 0);
      row_2d(d, es, &f, d3 * d1, d2, 0);
    } else {
      fn_t f = mkfn(FN_213_ROWSH, d1, d2, 0, 0, 0);
      row_2d(d, es, &f, d1 * d2, d3, 0);
      lo_231(d, es, d2, d1, d3);
    }
  }
}

/* 4dtranspose.cpp:69-187 (the ITNT variant, which consults alpha * vol) */
static void lo_1324(u8* d, int es, int64_t d1, int64_t d2, int64_t d3, int64_t d4,
                    int64_t tensor_vol, double alpha) {
  int64_t nthr = msb64(d1) <= 18 ? 32 : 256;
  double lim = alpha * (double)tensor_vol;
  fn_t rp = mkfn(FN_1324_ROWPERM, d3, d2, 0, 0, 0);
  fn_t rs = mkfn(FN_1324_ROWSH, d2, d3, d1, 0, 0);
  if (d1 >= 96 && (double)(nthr * d2 * d3 * d4) <= lim) {
    col_2d(d, es, &rp, d1, d2 * d3 * d4);
  } else if (d4 >= 4 && (double)(d1 * d2 * d3) > lim) {
    row_2d(d, es, &rs, d1 * d2 * d3, d4, 0);
  } else if (d4 >= 4) {
    row_2d(d, es, &rs, d1 * d2 * d3, d4, 0);
  } else if (d1 >= 96) {
    col_2d(d, es, &rp, d1, d2 * d3 * d4);
  } else if (d2 >= d3) {
    lo_213(d, es, d1 * d2, d3, d4);
    lo_213(d, es, d3, d1, d2 * d4);
  } else {
    fn_t f = mkfn(FN_213_ROWSH, d1, d2, 0, 0, 0);
    row_2d(d, es, &f, d1 * d2, d3 * d4, 0);
    lo_231(d, es, d2, d1, d3 * d4);
    fn_t g = mkfn(FN_132_ROWPERM, d2, d4, 0, 0, 0);
    col_2d(d, es, &g, d1 * d3, d4 * d2);
  }
}

/* Nd_transpose.cpp:62-99 */
static int lo_dispatch(u8* d, int es, const or_stage* st, int64_t tensor_vol, double alpha) {
  const int64_t* x = st->dims;
  int code = 0;
  for (int i = 0; i < st->rank; ++i) code = code * 10 + st->perm[i];
  switch (code) {
    case 21: lo_2d(d, es, x[0], x[1]); return 0;
    case 231: lo_231(d, es, x[0], x[1], x[2]); return 0;
    case 312: lo_312(d, es, x[0], x[1], x[2]); return 0;
    case 213: lo_213(d, es, x[0], x[1], x[2]); return 0;
    case 132: lo_132(d, es, x[0], x[1], x[2]); return 0;
    case 1324: lo_1324(d, es, x[0], x[1], x[2], x[3], tensor_vol, alpha); return 0;
    default: return -1;
  }
}

/* ------------------------------------------------------------------------ */
/* decomposer (ITNT/src/permutation_util.cpp, Nd_transpose.cpp:104-250)      */
/* ------------------------------------------------------------------------ */

/* permutation_util.cpp:16-38: replace values by their rank, 1-based */
static void rerank(int* p, int r) {
  int tmp[OR_MAX_RANK];
  for (int i = 0; i < r; ++i) {
    int cnt = 1;
    for (int j = 0; j < r; ++j)
      if (p[j] < p[i]) ++cnt;
    tmp[i] = cnt;
  }
  memcpy(p, tmp, sizeof(int) * (size_t)r);
}

/* permutation_util.cpp:40-79: input dims that stay adjacent (a run of
 * consecutive perm values) are multiplied into one dim, in input order */
static int merge_run_dims(const int64_t* dim, int64_t* out, const int* perm, int r) {
  int acc[OR_MAX_RANK] = {0};
  int p = 0;
  while (p < r) {
    int L = 1;
    while (p + L < r && perm[p + L] == perm[p + L - 1] + 1) ++L;
    acc[perm[p] - 1] = L - 1;
    p += L;
  }
  int mr = 0;
  for (int i = 0; i < r; ++i) {
    int64_t v = dim[i];
    for (int c = 1; c <= acc[i]; ++c) v *= dim[i + c];
    out[mr++] = v;
    i += acc[i];
  }
  return mr;
}

/* permutation_util.cpp:81-97: keep the first value of every run, rerank */
static int merge_run_perm(const int* in, int* out, int r) {
  int mc = 0, p = 0;
  while (p < r) {
    out[mc++] = in[p];
    int L = 1;
    while (p + L < r && in[p + L] == in[p + L - 1] + 1) ++L;
    p += L;
  }
  rerank(out, mc);
  return mc;
}

static int index_of(const int* v, int r, int x) {
  for (int i = 0; i < r; ++i)
    if (v[i] == x) return i;
  return r;
}

/* permutation_util.cpp:146-180 */
static void pre_round_perm(const int* target, const int* cur, int* rp, int r, int u) {
  int ti = index_of(cur, r, target[u]);
  int mc = 1;
  while (u + mc < r && ti + mc < r && target[u + mc] == cur[ti + mc]) ++mc;
  for (int i = 0; i < r; ++i) rp[i] = 0;
  for (int s = 0; s < mc; ++s) rp[u + s] = ti + s + 1;
  for (int cnt = 1; cnt <= r; ++cnt) {
    int used = 0;
    for (int s = 0; s < mc; ++s)
      if (rp[u + s] == cnt) used = 1;
    if (used) continue;
    for (int p = 0; p < r; ++p)
      if (rp[p] == 0) { rp[p] = cnt; break; }
  }
}

/* permutation_util.cpp:183-230 (sec_seq_part_size at :196-206 is unused) */
static void post_round_perm(const int* target, const int* cur, int* rp, int r, int u) {
  int ti = index_of(cur, r, target[u]);
  int mc = 1;
  while (u - mc >= 0 && ti - mc >= 0 && target[u - mc] == cur[ti - mc]) ++mc;
  for (int i = 0; i < r; ++i) rp[i] = 0;
  int first = u - mc + 1;
  for (int s = 0; s < mc; ++s) rp[first + s] = ti - mc + 2 + s;
  for (int cnt = r; cnt >= 1; --cnt) {
    int used = 0;
    for (int p = first; p < u; ++p)
      if (rp[p] == cnt) used = 1;
    if (used || cnt == ti + 1) continue;
    for (int p = r - 1; p >= 0; --p)
      if (rp[p] == 0) { rp[p] = cnt; break; }
  }
}

/* permutation_util.cpp:300-337 (pre) and :339-374 (post).
 * Returns the low-order rank, or r+1 when cur already equals target. */
static int gen_round(int post, const int* cur, const int64_t* dims, const int* target,
                     int64_t* lo_dims, int* lo_perm, int* rp, int r) {
  int u;
  if (!post) {
    u = r; /* :99-110 first mismatch */
    for (int i = 0; i < r; ++i)
      if (cur[i] != target[i]) { u = i; break; }
    if (u == r) return r + 1;
  } else {
    u = 0; /* :112-123 last mismatch */
    for (int i = r - 1; i >= 0; --i)
      if (cur[i] != target[i]) { u = i; break; }
    if (u == 0) return r + 1;
  }
  int64_t reordered[OR_MAX_RANK], merged_dim[OR_MAX_RANK];
  int merged_perm[OR_MAX_RANK];
  for (int i = 0; i < r; ++i) reordered[i] = dims[cur[i] - 1];
  int mr;
  if (!post) {
    /* :125-132 merge the matched prefix into one dim */
    if (u != 0) {
      mr = 1 + r - u;
      merged_dim[0] = 1;
      for (int i = 0; i < u; ++i) merged_dim[0] *= reordered[i];
      for (int i = u; i < r; ++i) merged_dim[1 + i - u] = reordered[i];
    } else {
      mr = r;
      for (int i = 0; i < r; ++i) merged_dim[i] = reordered[i];
    }
    pre_round_perm(target, cur, rp, r, u);
    /* :232-243 */
    if (u != 0) {
      merged_perm[0] = 1;
      for (int i = u; i < r; ++i) merged_perm[1 + i - u] = rp[i] - (u - 1);
    } else {
      for (int i = 0; i < r; ++i) merged_perm[i] = rp[i];
    }
  } else {
    /* :134-143 merge the matched suffix into one dim */
    if (u < r - 2) {
      mr = u + 2;
      for (int i = 0; i <= u; ++i) merged_dim[i] = reordered[i];
      merged_dim[u + 1] = 1;
      for (int i = u + 1; i < r; ++i) merged_dim[u + 1] *= reordered[i];
    } else {
      mr = r;
      for (int i = 0; i < r; ++i) merged_dim[i] = reordered[i];
    }
    post_round_perm(target, cur, rp, r, u);
    /* :245-255 */
    if (u < r - 2) {
      for (int i = 0; i <= u; ++i) merged_perm[i] = rp[i];
      merged_perm[u + 1] = u + 2;
    } else {
      for (int i = 0; i < r; ++i) merged_perm[i] = rp[i];
    }
  }
  int lr = merge_run_dims(merged_dim, lo_dims, merged_perm, mr);
  merge_run_perm(merged_perm, lo_perm, mr);
  return lr;
}

/* Nd_transpose.cpp:104-113 */
static int nonlinear(const int64_t* d, const int* p, int r) {
  if (r == 3) {
    if (p[2] == 3 && d[2] >= 8) return 0;
    if (p[0] == 1 && d[0] >= 48) return 0;
  } else if (r == 4) {
    if (p[3] == 4 && d[3] >= 4) return 0;
    if (p[0] == 1 && d[0] >= 96) return 0;
  }
  return 1;
}

static int same(const int* a, const int* b, int r) {
  for (int i = 0; i < r; ++i)
    if (a[i] != b[i]) return 0;
  return 1;
}

/* Nd_transpose.cpp:115-144 */
static int nonlinear_count(const int64_t* dims, const int* target, int r, int post) {
  int cur[OR_MAX_RANK], rp[OR_MAX_RANK], lp[OR_MAX_RANK], nxt[OR_MAX_RANK];
  int64_t ld[OR_MAX_RANK];
  for (int i = 0; i < r; ++i) cur[i] = i + 1;
  int cnt = 0;
  for (int round = 0; round < r; ++round) {
    int lr = gen_round(post, cur, dims, target, ld, lp, rp, r);
    if (lr == r + 1) break;
    cnt += nonlinear(ld, lp, lr);
    for (int i = 0; i < r; ++i) nxt[i] = cur[rp[i] - 1];
    memcpy(cur, nxt, sizeof(int) * (size_t)r);
    if (same(cur, target, r)) break;
  }
  return cnt;
}

int or_decompose(const int64_t* dims, const int* perm1, int rank, int reorder, or_stage* out,
                 int max_out) {
  if (rank < 1 || rank > OR_MAX_RANK) return -1;
  int p0[OR_MAX_RANK];
  for (int i = 0; i < rank; ++i) p0[i] = perm1[i] - 1;
  if (check_perm(p0, rank)) return -1;
  /* Nd_transpose.cpp:146-157: pre iff strictly fewer non-linearizable rounds */
  int post;
  if (reorder == 1) post = 0;
  else if (reorder == 2) post = 1;
  else post = nonlinear_count(dims, perm1, rank, 0) < nonlinear_count(dims, perm1, rank, 1) ? 0 : 1;
  int cur[OR_MAX_RANK], nxt[OR_MAX_RANK];
  for (int i = 0; i < rank; ++i) cur[i] = i + 1;
  int n = 0;
  /* Nd_transpose.cpp:192-249 */
  for (int round = 0; round < rank; ++round) {
    int rp_pre[OR_MAX_RANK], rp_post[OR_MAX_RANK], lp_pre[OR_MAX_RANK], lp_post[OR_MAX_RANK];
    int64_t ld_pre[OR_MAX_RANK], ld_post[OR_MAX_RANK];
    int r_pre = gen_round(0, cur, dims, perm1, ld_pre, lp_pre, rp_pre, rank);
    int r_post = gen_round(1, cur, dims, perm1, ld_post, lp_post, rp_post, rank);
    if (r_pre == rank + 1 || r_post == rank + 1) break;
    int lr = post ? r_post : r_pre;
    const int64_t* ld = post ? ld_post : ld_pre;
    const int* lp = post ? lp_post : lp_pre;
    const int* rp = post ? rp_post : rp_pre;
    if (lr > 4 || n >= max_out) return -2;
    out[n].rank = lr;
    for (int i = 0; i < 4; ++i) { out[n].perm[i] = i < lr ? lp[i] : 0; out[n].dims[i] = i < lr ? ld[i] : 1; }
    ++n;
    for (int i = 0; i < rank; ++i) nxt[i] = cur[rp[i] - 1];
    memcpy(cur, nxt, sizeof(int) * (size_t)rank);
    if (same(cur, perm1, rank)) break;
  }
  return n;
}

/* ------------------------------------------------------------------------ */
/* partition / join / padding (ITNT/src/Nd_decompose.cpp, Nd_padding.cu)     */
/* ------------------------------------------------------------------------ */

static void series(u8* d, int es, const int64_t* dims, const int* perm1, int rank,
                   int64_t tensor_vol, double alpha, int reorder) {
  or_stage st[64];
  int n = or_decompose(dims, perm1, rank, reorder, st, 64);
  for (int s = 0; s < n; ++s) lo_dispatch(d, es, &st[s], tensor_vol, alpha);
}

/* Nd_padding.cu:26-59,102-161: grow dim p to Dp, slab by slab, back to front */
static void pad_pre(u8* d, int es, const int64_t* dims, int rank, int p, int64_t Dp) {
  int64_t lo = 1, hi = 1;
  for (int i = 0; i < p; ++i) lo *= dims[i];
  for (int i = p + 1; i < rank; ++i) hi *= dims[i];
  int64_t D = dims[p];
  for (int64_t s = hi - 1; s >= 1; --s)
    memmove(d + s * lo * Dp * es, d + s * lo * D * es, (size_t)(lo * D * es));
}

/* Nd_padding.cu:61-99,168-223: compact the padded output dim */
static void pad_post(u8* d, int es, const int64_t* dims, const int* perm0, int rank, int p,
                     int64_t D, int64_t Dp) {
  int pp = 0;
  for (int i = 0; i < rank; ++i)
    if (perm0[i] == p) pp = i;
  int64_t lo = 1, hi = 1;
  for (int i = 0; i < pp; ++i) lo *= dims[perm0[i]];
  for (int i = pp + 1; i < rank; ++i) hi *= dims[perm0[i]];
  for (int64_t s = 1; s < hi; ++s)
    memmove(d + s * lo * D * es, d + s * lo * Dp * es, (size_t)(lo * D * es));
}

int or_simulate_reference(void* buf, int64_t capacity_elems, const int64_t* dims_in,
                          const int* perm1, int rank, int elem, int num_block, double alpha,
                          int reorder) {
  u8* d = (u8*)buf;
  int64_t dims[OR_MAX_RANK];
  int perm0[OR_MAX_RANK];
  if (rank < 1 || rank > OR_MAX_RANK) return -1;
  for (int i = 0; i < rank; ++i) { dims[i] = dims_in[i]; perm0[i] = perm1[i] - 1; }
  if (check_perm(perm0, rank)) return -1;
  int64_t vol = 1;
  for (int i = 0; i < rank; ++i) vol *= dims[i];
  /* Nd_transpose.cpp:265-275 */
  int64_t large = (int64_t)(32.0 * (1.0 / alpha));
  int p = 0; /* ITNT/src/util.cpp:73-76 first argmax */
  for (int i = 1; i < rank; ++i)
    if (dims[i] > dims[p]) p = i;
  int nb = num_block;
  int decompose = dims[p] > large && nb > 1;
  if (!decompose) nb = 1;
  int64_t D = dims[p], Dp = D;
  if (decompose && D % nb != 0) {
    Dp = (D / nb + 1) * nb;
    if (vol / D * Dp > capacity_elems) return -2;
    pad_pre(d, elem, dims, rank, p, Dp);
    dims[p] = Dp;
    vol = vol / D * Dp;
  }
  if (decompose) {
    /* Nd_decompose.cpp:48-56: (prod_{<=p}/nb, nb, prod_{>p}) then 132 */
    int64_t ds = 1, nds = 1;
    for (int i = 0; i <= p; ++i) ds *= dims[i];
    for (int i = p + 1; i < rank; ++i) nds *= dims[i];
    lo_132(d, elem, ds / nb, nb, nds);
    dims[p] /= nb;
    int64_t sub = vol / nb;
    for (int b = 0; b < nb; ++b) series(d + b * sub * elem, elem, dims, perm1, rank, vol, alpha, reorder);
    dims[p] *= nb;
    /* Nd_decompose.cpp:65-82: (prod_{<=p'}/nb, prod_{>p'}, nb) then 132 */
    int pp = 0;
    for (int i = 0; i < rank; ++i)
      if (perm0[i] == p) pp = i;
    ds = 1;
    nds = 1;
    for (int i = 0; i <= pp; ++i) ds *= dims[perm0[i]];
    for (int i = pp + 1; i < rank; ++i) nds *= dims[perm0[i]];
    lo_132(d, elem, ds / nb, nds, nb);
  } else {
    series(d, elem, dims, perm1, rank, vol, alpha, reorder);
  }
  if (Dp != D) {
    pad_post(d, elem, dims, perm0, rank, p, D, Dp);
  }
  return 0;
}

int or_transpose2d_catanzaro(void* buf, int64_t B, int64_t R, int64_t C, int64_t U, int method) {
  u8* d = (u8*)buf;
  if (U <= 0 || U > (1 << 30)) return -1;
  for (int64_t b = 0; b < B; ++b) {
    u8* m = d + b * R * C * U;
    if (method == 0) c2r_generic(SHAPE_2D, m, (int)U, C, R, 1);
    else r2c_generic(SHAPE_2D, m, (int)U, R, C, 1);
  }
  return 0;
}

int or_transpose2d_direct(void* buf, int64_t B, int64_t R, int64_t C, int64_t U) {
  u8* d = (u8*)buf;
  u8* tmp = (u8*)malloc((size_t)(R * C * U));
  if (!tmp) return -1;
  for (int64_t b = 0; b < B; ++b) {
    u8* m = d + b * R * C * U;
    for (int64_t i = 0; i < R; ++i)
      for (int64_t j = 0; j < C; ++j) memcpy(tmp + (j * R + i) * U, m + (i * C + j) * U, (size_t)U);
    memcpy(m, tmp, (size_t)(R * C * U));
  }
  free(tmp);
  return 0;
}
