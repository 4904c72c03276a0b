// gfx950 pass kernels for in-place transposition.
//
// Every pass permutes units inside disjoint "footprints":
//   OP_TILE  one B-matrix (R x C units), written back transposed (C x R)
//   OP_COL   one column panel: W adjacent columns x all R rows of the grid
//   OP_ROW   one row of C units
// optionally restricted to one `us`-byte slice of every unit (a "cell";
// units larger than a footprint budget are cut into slices permuted alike).
//
// LDS kernels: a workgroup loads K footprints into LDS (coalesced 16-byte
// loads, UNR of them in flight per thread), barriers, and writes them back
// in gather order with 16-byte stores that pack several small cells.  Tile
// rows in LDS are padded so that column reads hit distinct banks.
// Staged kernels (footprints larger than LDS): the footprints are copied to
// a scratch buffer and written back through the same index maps; tiles go
// through an LDS-tiled out-of-place transpose from scratch.
#include "kernels.hpp"

#include <algorithm>
#include <cstring>
#include <type_traits>

#include "eithot_status.h"
#include "maps.hpp"

namespace eithot {
namespace {

constexpr int UNR = 8;  // global loads in flight per thread

// 16-byte word: a native vector, so arrays of it stay in registers
typedef uint32_t W16 __attribute__((ext_vector_type(4)));

// V words packed into one wider store (the union keeps them in registers)
template <int BYTES>
struct PackOf;
template <>
struct PackOf<1> {
  using T = uint8_t;
};
template <>
struct PackOf<2> {
  using T = uint16_t;
};
template <>
struct PackOf<4> {
  using T = uint32_t;
};
template <>
struct PackOf<8> {
  using T = uint64_t;
};
template <>
struct PackOf<16> {
  using T = W16;
};
template <typename Word, int V>
union Packed {
  using P = typename PackOf<V * sizeof(Word)>::T;
  Word e[V];
  P p;
};

__device__ __forceinline__ uint32_t xcd_logical(uint32_t bid, uint32_t n, int remap) {
  if (!remap || n < 16) return bid;
  const uint32_t per = n / 8;
  if (bid >= per * 8) return bid;
  return (bid % 8) * per + bid / 8;
}

// ---------------------------------------------------------------------------
// common arguments
// ---------------------------------------------------------------------------
// 16-byte LDS store at a 4-byte aligned offset (padded rows may break
// 16-byte alignment)
__device__ __forceinline__ void lds_put16(unsigned char* p, const W16& v) {
  const uintptr_t a = (uintptr_t)p;
  if ((a & 15) == 0) {
    *reinterpret_cast<W16*>(p) = v;
  } else if ((a & 7) == 0) {
    reinterpret_cast<uint64_t*>(p)[0] = (uint64_t)v[0] | ((uint64_t)v[1] << 32);
    reinterpret_cast<uint64_t*>(p)[1] = (uint64_t)v[2] | ((uint64_t)v[3] << 32);
  } else {
    uint32_t* q = reinterpret_cast<uint32_t*>(p);
    q[0] = v[0];
    q[1] = v[1];
    q[2] = v[2];
    q[3] = v[3];
  }
}

template <typename Idx>
struct Args {
  char* data;
  const char* src;   // staged tile: scratch holding the footprints
  char* scratch;
  uint64_t base;     // byte offset of batch 0
  uint64_t U, us, RCU, CU;
  Idx R, C, W, S, K, nfp, P, wu;
  Idx rs;            // LDS bytes per footprint row (padded)
  Idx fpl;           // LDS bytes per footprint
  Idx cpr;           // 16-byte chunks per footprint row (0: load word by word)
  Idx ld_per;        // load-loop items per footprint
  Idx st_per;        // store-loop items per footprint
  Idx vpr;           // store vectors per output row (vector stores)
  Idx cps;           // 16-byte chunks per cell (0: cells are not whole chunks)
  int cps_log2;      // log2(cps) when cps is a power of two, else -1
  uint64_t fp0, G;   // staged: first footprint and count of the group
  int contiguous;    // S == 1 and the K footprints are one contiguous byte range
  // wide path (column and row passes of whole units narrower than 16-byte
  // words): global rows move as aligned 16-byte chunks into an LDS copy
  // that keeps each row's address modulo 16; stores are 16-byte chunks
  // gathered word by word.  hi: end of the pass's bytes (offset from data).
  int wide;
  Idx ncw;           // column pass: 16-byte chunks per LDS row (rs / 16)
  uint64_t hi;
  typename FDsel<Idx>::type fS, fR, fC, fW, fP, fwu, fcpr, fld, fst, fvpr, fcps, fncw, fU, fCU;
  GridMaps<Idx> g;
};

// byte address (offset from data) of cell 0 of footprint f, and its grid role
template <typename Idx, int KIND>
struct Foot {
  uint64_t addr;
  Idx row;   // OP_ROW: row within the grid; OP_COL: first column of the panel
  Idx weff;  // OP_COL: columns present in the panel
};

template <typename Idx, int KIND>
__device__ __forceinline__ Foot<Idx, KIND> decode(const Args<Idx>& a, Idx f) {
  Foot<Idx, KIND> r;
  const Idx o = a.S == 1 ? f : a.fS.div(f);
  const Idx s = f - o * a.S;
  r.row = 0;
  r.weff = 0;
  if (KIND == OP_TILE) {
    r.addr = a.base + (uint64_t)o * a.RCU + (uint64_t)s * a.us;
  } else if (KIND == OP_COL) {
    const Idx b = a.fP.div(o);
    const Idx p = o - b * a.P;
    const Idx j0 = p * a.W;
    r.addr = a.base + (uint64_t)b * a.RCU + (uint64_t)j0 * a.U + (uint64_t)s * a.us;
    r.row = j0;
    r.weff = (a.C - j0) < a.W ? (a.C - j0) : a.W;
  } else {
    r.addr = a.base + (uint64_t)o * a.CU + (uint64_t)s * a.us;
    r.row = a.fR.mod(o);
  }
  return r;
}

// source row of the column pass / source cell of the row pass
template <typename Idx, int MAP>
__device__ __forceinline__ Idx col_src(const Args<Idx>& a, Idx i, Idx j) {
  if (MAP == MAP_ROT) return a.g.rot(i, j);
  if (MAP == MAP_ROTINV) return a.g.rotinv(i, j);
  if (MAP == MAP_COLSH) return a.g.colsh(i, j);
  return a.g.colshinv(i, j);
}
template <typename Idx, int MAP>
__device__ __forceinline__ Idx row_src(const Args<Idx>& a, Idx i, Idx j) {
  if (MAP == MAP_ROWG) return a.g.rowg(i, j);
  return a.g.rowd(i, j);
}

// ---------------------------------------------------------------------------
// tile pass: K matrices (or matrix slices) per workgroup
// LDS: footprint k row i at k*fpl + i*rs (rs padded), cell j at +j*us
// ---------------------------------------------------------------------------
// Loads the contiguous run [g0, g1) (offsets from data) into lds at its
// address modulo 16: lds byte y holds byte a0 + y, a0 = g0 rounded down to
// an aligned 16-byte chunk, so aligned chunks land on aligned LDS chunks.
// A chunk wholly inside the pass's bytes [base, hi) moves with one 16-byte
// load (bytes outside the run are loaded but never stored); a chunk at the
// pass's edge moves word by word.
template <typename Word, typename Idx>
__device__ __forceinline__ void run_to_lds(const Args<Idx>& a, int64_t g0, int64_t g1, int64_t a0,
                                           unsigned char* lds) {
  constexpr Idx w = sizeof(Word);
  const Idx tid = threadIdx.x, nt = blockDim.x;
  const Idx n = (Idx)((g1 - a0 + 15) / 16);
  for (Idx t0 = tid; t0 < n; t0 += nt * UNR) {
    W16 v[UNR];
    bool full[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const Idx t = t0 + (Idx)u * nt;
      full[u] = false;
      if (t >= n) continue;
      const int64_t cs = a0 + (int64_t)t * 16;
      if (cs >= (int64_t)a.base && cs + 16 <= (int64_t)a.hi) {
        v[u] = *reinterpret_cast<const W16*>(a.data + cs);
        full[u] = true;
      } else {
        const int64_t b0 = cs < g0 ? g0 : cs, b1 = cs + 16 < g1 ? cs + 16 : g1;
        for (int64_t b = b0; b < b1; b += w)
          *reinterpret_cast<Word*>(lds + (b - a0)) = *reinterpret_cast<const Word*>(a.data + b);
      }
    }
#pragma unroll
    for (int u = 0; u < UNR; ++u)
      if (full[u]) *reinterpret_cast<W16*>(lds + (t0 + (Idx)u * nt) * 16) = v[u];
  }
}

// Wide tile pass (whole units of 1- and 2-byte words, S == 1): the
// workgroup's nf matrices are one contiguous run, loaded as aligned 16-byte
// chunks (run_to_lds); every cell is then copied word by word to its
// transposed place in a second LDS copy of the run, which goes back to HBM
// as aligned 16-byte stores (word by word in the run's two edge chunks).
// Cells are visited in input order (ord 0: consecutive threads read
// consecutive cells) or output order (ord 1), whichever the host found to
// have fewer LDS bank conflicts.
template <typename Word, typename Idx>
__device__ __forceinline__ void tile_wide(const Args<Idx>& a, const Idx f0, const Idx nf,
                                          unsigned char* lds) {
  constexpr Idx w = sizeof(Word);
  const Idx tid = threadIdx.x, nt = blockDim.x;
  const uintptr_t d0 = (uintptr_t)a.data;
  const int64_t g0 = (int64_t)(a.base + (uint64_t)f0 * a.RCU);
  const int64_t g1 = g0 + (int64_t)nf * (int64_t)a.RCU;
  const int64_t a0 = g0 - (int64_t)((d0 + (uint64_t)g0) & 15);
  const Idx sh = (Idx)(g0 - a0);
  unsigned char* out = lds + (((Idx)a.K * (Idx)a.RCU + 32 + 15) & ~(Idx)15);
  run_to_lds<Word, Idx>(a, g0, g1, a0, lds);
  __syncthreads();
  const Idx RC = a.R * a.C, U = (Idx)a.U, RCU = (Idx)a.RCU;
  const Idx ncell = nf * RC;
  for (Idx q = tid; q < ncell; q += nt) {
    const Idx k = a.fld.div(q);  // ld_per == R*C
    const Idx r = q - k * RC;
    Idx i, j;
    if (a.wide == 1) {  // r = i*C + j: input order
      i = a.fC.div(r);
      j = r - i * a.C;
    } else {            // r = j*R + i: output order
      j = a.fR.div(r);
      i = r - j * a.R;
    }
    const Idx s = sh + k * RCU + (i * a.C + j) * U;
    const Idx d = sh + k * RCU + (j * a.R + i) * U;
    for (Idx x = 0; x < U; x += w)
      *reinterpret_cast<Word*>(out + d + x) = *reinterpret_cast<const Word*>(lds + s + x);
  }
  __syncthreads();
  const Idx n = (Idx)((g1 - a0 + 15) / 16);
  for (Idx t = tid; t < n; t += nt) {
    const int64_t cs = a0 + (int64_t)t * 16;
    if (cs >= g0 && cs + 16 <= g1) {
      *reinterpret_cast<W16*>(a.data + cs) = *reinterpret_cast<const W16*>(out + t * 16);
    } else {
      const int64_t b0 = cs < g0 ? g0 : cs, b1 = cs + 16 < g1 ? cs + 16 : g1;
      for (int64_t b = b0; b < b1; b += w)
        *reinterpret_cast<Word*>(a.data + b) = *reinterpret_cast<const Word*>(out + (b - a0));
    }
  }
}

// Wide tile pass with one LDS copy (wide == 3: matrices too large for two
// copies): the run is loaded as in tile_wide, then every aligned 16-byte
// output chunk gathers its words straight from that copy.  Output byte
// (matrix k, output cell j*R + i, byte xu) is input byte k*RCU + i*CU +
// j*U + xu; consecutive words step a four-digit odometer (xu, i, j, k)
// whose carries are selects, not branches.
template <typename Word, typename Idx>
__device__ __forceinline__ void tile_gather(const Args<Idx>& a, const Idx f0, const Idx nf,
                                            unsigned char* lds) {
  constexpr uint32_t w = sizeof(Word);
  constexpr int V = 16 / (int)sizeof(Word);
  const Idx tid = threadIdx.x, nt = blockDim.x;
  const uintptr_t d0 = (uintptr_t)a.data;
  const int64_t g0 = (int64_t)(a.base + (uint64_t)f0 * a.RCU);
  const int64_t g1 = g0 + (int64_t)nf * (int64_t)a.RCU;
  const int64_t a0 = g0 - (int64_t)((d0 + (uint64_t)g0) & 15);
  const uint32_t sh = (uint32_t)(g0 - a0);
  run_to_lds<Word, Idx>(a, g0, g1, a0, lds);
  __syncthreads();
  const uint32_t U = (uint32_t)a.U, CU = (uint32_t)a.CU, RCU = (uint32_t)a.RCU;
  const uint32_t R = (uint32_t)a.R, C = (uint32_t)a.C;
  const Idx n = (Idx)((g1 - a0 + 15) / 16);
  for (Idx t = tid; t < n; t += nt) {
    const int64_t cs = a0 + 16 * (int64_t)t;
    const int64_t b0 = cs < g0 ? g0 : cs, b1 = cs + 16 < g1 ? cs + 16 : g1;
    // the first word's matrix, output cell and byte
    const Idx x = (Idx)(b0 - g0);
    const Idx c = a.fU.div(x);
    const Idx k = a.fld.div(c);  // ld_per == R*C
    const Idx r = c - k * (Idx)(R * C);
    const Idx jq = a.fR.div(r);
    uint32_t i = (uint32_t)(r - jq * R), j = (uint32_t)jq, xu = (uint32_t)(x - c * U);
    uint32_t src = sh + (uint32_t)k * RCU + i * CU + j * U + xu;
    auto step = [&]() {
      xu += w;
      src += w;
      const bool e1 = xu == U;
      xu = e1 ? 0 : xu;
      i += e1 ? 1 : 0;
      src += e1 ? CU - U : 0;
      const bool e2 = i == R;
      i = e2 ? 0 : i;
      j += e2 ? 1 : 0;
      src += e2 ? U - R * CU : 0;
      const bool e3 = j == C;
      j = e3 ? 0 : j;
      src += e3 ? RCU - C * U : 0;
    };
    if (b0 == cs && b1 == cs + 16) {
      Packed<Word, V> tmp;
#pragma unroll
      for (int v = 0; v < V; ++v) {
        tmp.e[v] = *reinterpret_cast<const Word*>(lds + src);
        step();
      }
      *reinterpret_cast<typename Packed<Word, V>::P*>(a.data + cs) = tmp.p;
    } else {
      for (int64_t b = b0; b < b1; b += w) {
        *reinterpret_cast<Word*>(a.data + b) = *reinterpret_cast<const Word*>(lds + src);
        step();
      }
    }
  }
}

template <typename Word, typename Idx, int VEC>
__device__ __forceinline__ void tile_body(const Args<Idx>& a, const Idx f0, unsigned char* lds) {
  const Idx nf = (a.nfp - f0) < a.K ? (a.nfp - f0) : a.K;
  const Idx tid = threadIdx.x, nt = blockDim.x;
  constexpr Idx w = sizeof(Word);
  if constexpr (sizeof(Word) < 16) {
    if (a.wide == 3) {
      tile_gather<Word, Idx>(a, f0, nf, lds);
      return;
    }
    if (a.wide) {
      tile_wide<Word, Idx>(a, f0, nf, lds);
      return;
    }
  }

  if (a.cpr) {
    // nf whole contiguous matrices: rows of C*us bytes in 16-byte chunks;
    // row r of the group lands at r*rs (fpl == R*rs)
    const W16* g = reinterpret_cast<const W16*>(a.data + a.base + (uint64_t)f0 * a.RCU);
    const Idx n = nf * a.R * a.cpr;
    for (Idx t0 = tid; t0 < n; t0 += nt * UNR) {
      W16 v[UNR];
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        const Idx t = t0 + (Idx)u * nt;
        if (t < n) v[u] = g[t];
      }
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        const Idx t = t0 + (Idx)u * nt;
        if (t < n) {
          const Idx r = a.fcpr.div(t);
          lds_put16(lds + r * a.rs + (t - r * a.cpr) * 16, v[u]);
        }
      }
    }
  } else if (a.cps) {
    // cells of whole 16-byte chunks (sliced units): chunk t -> footprint k,
    // cell q = i*C + j, chunk c; global cells are U bytes apart
    const Idx per = a.R * a.C * a.cps;
    const Idx n = nf * per;
    const Foot<Idx, OP_TILE> f1 = decode<Idx, OP_TILE>(a, f0);
    for (Idx t0 = tid; t0 < n; t0 += nt * UNR) {
      W16 v[UNR];
      Idx off[UNR];
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        const Idx t = t0 + (Idx)u * nt;
        if (t < n) {
          Idx k = 0, rem = t;
          uint64_t addr = f1.addr;
          if (nf > 1) {
            k = a.fld.div(t);
            rem = t - k * per;
            addr = decode<Idx, OP_TILE>(a, f0 + k).addr;
          }
          const Idx q = a.cps_log2 >= 0 ? rem >> a.cps_log2 : a.fcps.div(rem);
          const Idx c = rem - q * a.cps;
          const Idx i = a.fC.div(q);
          const Idx j = q - i * a.C;
          v[u] = *reinterpret_cast<const W16*>(a.data + addr + (uint64_t)q * a.U + c * 16);
          off[u] = k * a.fpl + i * a.rs + j * (Idx)a.us + c * 16;
        }
      }
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        const Idx t = t0 + (Idx)u * nt;
        if (t < n) *reinterpret_cast<W16*>(lds + off[u]) = v[u];
      }
    }
  } else {
    // word granular: footprint k, cell q = i*C + j, word x
    const Idx n = nf * a.ld_per;
    for (Idx t0 = tid; t0 < n; t0 += nt * UNR) {
      Word v[UNR];
      Idx off[UNR];
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        const Idx t = t0 + (Idx)u * nt;
        if (t < n) {
          const Idx k = a.fld.div(t);
          const Idx rem = t - k * a.ld_per;
          const Idx q = a.wu == 1 ? rem : a.fwu.div(rem);
          const Idx x = rem - q * a.wu;
          const Idx i = a.fC.div(q);
          const Idx j = q - i * a.C;
          const Foot<Idx, OP_TILE> f = decode<Idx, OP_TILE>(a, f0 + k);
          v[u] = *reinterpret_cast<const Word*>(a.data + f.addr + (uint64_t)q * a.U +
                                                (uint64_t)x * w);
          off[u] = k * a.fpl + i * a.rs + j * (Idx)a.us + x * w;
        }
      }
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        const Idx t = t0 + (Idx)u * nt;
        if (t < n) *reinterpret_cast<Word*>(lds + off[u]) = v[u];
      }
    }
  }
  __syncthreads();

  if (VEC > 1) {
    // S == 1, one word per cell: VEC cells (consecutive i of output row j)
    // per store; output row j of matrix k starts at (j*R)*us
    const Idx n = nf * a.st_per;
    char* gbase = a.data + a.base + (uint64_t)f0 * a.RCU;
    for (Idx t = tid; t < n; t += nt) {
      const Idx k = a.fst.div(t);
      const Idx rem = t - k * a.st_per;
      const Idx j = a.fvpr.div(rem);
      const Idx i0 = (rem - j * a.vpr) * VEC;
      Packed<Word, VEC> tmp;
      const unsigned char* l = lds + k * a.fpl + i0 * a.rs + j * w;
#pragma unroll
      for (int v = 0; v < VEC; ++v) tmp.e[v] = *reinterpret_cast<const Word*>(l + v * a.rs);
      *reinterpret_cast<typename Packed<Word, VEC>::P*>(gbase + (uint64_t)k * a.RCU +
                                                  ((uint64_t)j * a.R + i0) * w) = tmp.p;
    }
  } else if (a.cps) {
    const Idx per = a.R * a.C * a.cps;
    const Idx n = nf * per;
    const Foot<Idx, OP_TILE> f1 = decode<Idx, OP_TILE>(a, f0);
    for (Idx t = tid; t < n; t += nt) {
      Idx k = 0, rem = t;
      uint64_t addr = f1.addr;
      if (nf > 1) {
        k = a.fld.div(t);
        rem = t - k * per;
        addr = decode<Idx, OP_TILE>(a, f0 + k).addr;
      }
      const Idx q = a.cps_log2 >= 0 ? rem >> a.cps_log2 : a.fcps.div(rem);  // q = j*R + i
      const Idx c = rem - q * a.cps;
      const Idx j = a.fR.div(q);
      const Idx i = q - j * a.R;
      *reinterpret_cast<W16*>(a.data + addr + (uint64_t)q * a.U + c * 16) =
          *reinterpret_cast<const W16*>(lds + k * a.fpl + i * a.rs + j * (Idx)a.us + c * 16);
    }
  } else {
    const Idx n = nf * a.st_per;
    for (Idx t = tid; t < n; t += nt) {
      const Idx k = a.fst.div(t);
      const Idx rem = t - k * a.st_per;
      const Idx q = a.wu == 1 ? rem : a.fwu.div(rem);  // output cell q = j*R + i
      const Idx x = rem - q * a.wu;
      const Idx j = a.fR.div(q);
      const Idx i = q - j * a.R;
      const Foot<Idx, OP_TILE> f = decode<Idx, OP_TILE>(a, f0 + k);
      *reinterpret_cast<Word*>(a.data + f.addr + (uint64_t)q * a.U + (uint64_t)x * w) =
          *reinterpret_cast<const Word*>(lds + k * a.fpl + i * a.rs + j * (Idx)a.us + x * w);
    }
  }
}

// ---------------------------------------------------------------------------
// column pass: panels of W columns x R rows; LDS row i of footprint k holds
// the panel's segment (W*us bytes) at k*fpl + i*rs
// ---------------------------------------------------------------------------
// 16-byte global load / store of an aligned chunk
__device__ __forceinline__ W16 ld16(const char* p) { return *reinterpret_cast<const W16*>(p); }
__device__ __forceinline__ void st16(char* p, const W16& v) { *reinterpret_cast<W16*>(p) = v; }

// Wide column pass.  Row i of panel k (global bytes [g, g + L), g = f.addr +
// i*CU, L = weff*us: W whole units, or one us-byte slice of a unit when
// W == 1) lies in LDS at k*fpl + i*rs + (address of g mod 16): the
// aligned 16-byte chunks that cover the row land on aligned LDS chunks.  A
// chunk wholly inside the pass's bytes moves with one 16-byte load (bytes
// outside the row are loaded but never stored); a chunk at the pass's edge
// moves word by word.  Stores: output chunk c of row i is gathered word by
// word (column jj's words come from LDS row col_src(i, j0 + jj)) and stored
// with one 16-byte store when it lies inside the row, word by word at the
// row's two ends.
template <typename Word, typename Idx, int MAP>
__device__ __forceinline__ void col_wide(const Args<Idx>& a, const Idx f0, const Idx nf,
                                         unsigned char* lds) {
  constexpr Idx w = sizeof(Word);
  constexpr int V = 16 / (int)sizeof(Word);
  const Idx tid = threadIdx.x, nt = blockDim.x;
  const uintptr_t d0 = (uintptr_t)a.data;
  const Idx n = nf * a.ld_per;  // ld_per = R * ncw
  for (Idx t0 = tid; t0 < n; t0 += nt * UNR) {
    W16 v[UNR];
    Idx off[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const Idx t = t0 + (Idx)u * nt;
      off[u] = ~(Idx)0;
      if (t >= n) continue;
      const Idx k = a.fld.div(t);
      const Idx rem = t - k * a.ld_per;
      const Idx i = a.fncw.div(rem);
      const Idx c = rem - i * a.ncw;
      const Foot<Idx, OP_COL> f = decode<Idx, OP_COL>(a, f0 + k);
      const int64_t g = (int64_t)(f.addr + (uint64_t)i * a.CU);
      const int64_t L = (int64_t)f.weff * (int64_t)a.us;
      // chunk start, relative to data: may lie up to 15 bytes before it
      const int64_t cs = g - (int64_t)((d0 + (uint64_t)g) & 15) + (int64_t)c * 16;
      if (cs >= g + L) continue;
      const Idx lo = k * a.fpl + i * a.rs + c * 16;
      if (cs >= (int64_t)a.base && cs + 16 <= (int64_t)a.hi) {
        v[u] = ld16(a.data + cs);
        off[u] = lo;
      } else {
        const int64_t b0 = cs < g ? g : cs, b1 = cs + 16 < g + L ? cs + 16 : g + L;
        for (int64_t b = b0; b < b1; b += w)
          *reinterpret_cast<Word*>(lds + lo + (b - cs)) = *reinterpret_cast<const Word*>(a.data + b);
      }
    }
#pragma unroll
    for (int u = 0; u < UNR; ++u)
      if (off[u] != ~(Idx)0) *reinterpret_cast<W16*>(lds + off[u]) = v[u];
  }
  __syncthreads();
  for (Idx t = tid; t < n; t += nt) {
    const Idx k = a.fld.div(t);
    const Idx rem = t - k * a.ld_per;
    const Idx i = a.fncw.div(rem);
    const Idx c = rem - i * a.ncw;
    const Foot<Idx, OP_COL> f = decode<Idx, OP_COL>(a, f0 + k);
    const int64_t g = (int64_t)(f.addr + (uint64_t)i * a.CU);
    const int64_t L = (int64_t)f.weff * (int64_t)a.us;
    const int64_t cs = g - (int64_t)((d0 + (uint64_t)g) & 15) + (int64_t)c * 16;
    if (cs >= g + L) continue;
    const int64_t b0 = cs < g ? g : cs, b1 = cs + 16 < g + L ? cs + 16 : g + L;
    // word b0 of the row: column jj, byte xu of its unit (or slice)
    const Idx x = (Idx)(b0 - g);
    Idx jj = a.fU.div(x);
    Idx xu = x - jj * (Idx)a.us;
    // LDS start of source row r's segment: k*fpl + r*rs plus the segment's
    // address modulo 16, (f15 + r*cu15) mod 16 in 32 bits.  Byte y of the
    // output row (column jj, byte xu) is byte y of the source row's segment,
    // so only the row base changes, once per column.
    const uint32_t f15 = (uint32_t)((d0 + f.addr) & 15), cu15 = (uint32_t)(a.CU & 15);
    auto rowb = [&](Idx r) -> const unsigned char* {
      return lds + k * a.fpl + r * a.rs + ((f15 + (uint32_t)r * cu15) & 15);
    };
    const unsigned char* rb = rowb(col_src<Idx, MAP>(a, i, f.row + jj));
    if (b0 == cs && b1 == cs + 16) {
      Packed<Word, V> tmp;
#pragma unroll
      for (int q = 0; q < V; ++q) {
        if (xu == (Idx)a.us) {
          xu = 0;
          ++jj;
          rb = rowb(col_src<Idx, MAP>(a, i, f.row + jj));
        }
        tmp.e[q] = *reinterpret_cast<const Word*>(rb + x + (Idx)q * w);
        xu += w;
      }
      *reinterpret_cast<typename Packed<Word, V>::P*>(a.data + cs) = tmp.p;
    } else {
      Idx y = x;
      for (int64_t b = b0; b < b1; b += w, y += w) {
        if (xu == (Idx)a.us) {
          xu = 0;
          ++jj;
          rb = rowb(col_src<Idx, MAP>(a, i, f.row + jj));
        }
        *reinterpret_cast<Word*>(a.data + b) = *reinterpret_cast<const Word*>(rb + y);
        xu += w;
      }
    }
  }
}

template <typename Word, typename Idx, int MAP, int VEC>
__device__ __forceinline__ void col_body(const Args<Idx>& a, const Idx f0, unsigned char* lds) {
  const Idx nf = (a.nfp - f0) < a.K ? (a.nfp - f0) : a.K;
  const Idx tid = threadIdx.x, nt = blockDim.x;
  constexpr Idx w = sizeof(Word);
  if constexpr (sizeof(Word) < 16) {
    if (a.wide) {
      col_wide<Word, Idx, MAP>(a, f0, nf, lds);
      return;
    }
  }

  {
    // a.cpr != 0: every panel is full and its segment is whole 16-byte chunks
    const Idx n = nf * a.ld_per;
    for (Idx t0 = tid; t0 < n; t0 += nt * UNR) {
      W16 v16[UNR];
      Word vw[UNR];
      Idx off[UNR];
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        const Idx t = t0 + (Idx)u * nt;
        off[u] = ~(Idx)0;
        if (t < n) {
          const Idx k = a.fld.div(t);
          const Idx rem = t - k * a.ld_per;
          const Foot<Idx, OP_COL> f = decode<Idx, OP_COL>(a, f0 + k);
          if (a.cpr) {
            const Idx i = a.fcpr.div(rem);
            const Idx c = rem - i * a.cpr;
            v16[u] = *reinterpret_cast<const W16*>(a.data + f.addr + (uint64_t)i * a.CU + c * 16);
            off[u] = k * a.fpl + i * a.rs + c * 16;
          } else {
            const Idx q = a.wu == 1 ? rem : a.fwu.div(rem);
            const Idx x = rem - q * a.wu;
            const Idx i = a.fW.div(q);
            const Idx jj = q - i * a.W;
            if (jj < f.weff) {
              vw[u] = *reinterpret_cast<const Word*>(a.data + f.addr + (uint64_t)i * a.CU +
                                                     (uint64_t)jj * a.U + x * w);
              off[u] = k * a.fpl + i * a.rs + jj * (Idx)a.us + x * w;
            }
          }
        }
      }
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        if (off[u] == ~(Idx)0) continue;
        if (a.cpr) *reinterpret_cast<W16*>(lds + off[u]) = v16[u];
        else *reinterpret_cast<Word*>(lds + off[u]) = vw[u];
      }
    }
  }
  __syncthreads();

  if (VEC > 1) {
    // full panels, one word per cell: VEC adjacent columns of row i per store
    const Idx n = nf * a.st_per;
    for (Idx t = tid; t < n; t += nt) {
      const Idx k = a.fst.div(t);
      const Idx rem = t - k * a.st_per;
      const Idx i = a.fvpr.div(rem);
      const Idx jj0 = (rem - i * a.vpr) * VEC;
      const Foot<Idx, OP_COL> f = decode<Idx, OP_COL>(a, f0 + k);
      Packed<Word, VEC> tmp;
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        const Idx s = col_src<Idx, MAP>(a, i, f.row + jj0 + v);
        tmp.e[v] = *reinterpret_cast<const Word*>(lds + k * a.fpl + s * a.rs + (jj0 + v) * w);
      }
      *reinterpret_cast<typename Packed<Word, VEC>::P*>(a.data + f.addr + (uint64_t)i * a.CU +
                                                  (uint64_t)jj0 * w) = tmp.p;
    }
  } else {
    const Idx n = nf * a.st_per;
    for (Idx t = tid; t < n; t += nt) {
      const Idx k = a.fst.div(t);
      const Idx rem = t - k * a.st_per;
      const Idx q = a.wu == 1 ? rem : a.fwu.div(rem);
      const Idx x = rem - q * a.wu;
      const Idx i = a.fW.div(q);
      const Idx jj = q - i * a.W;
      const Foot<Idx, OP_COL> f = decode<Idx, OP_COL>(a, f0 + k);
      if (jj >= f.weff) continue;
      const Idx s = col_src<Idx, MAP>(a, i, f.row + jj);
      *reinterpret_cast<Word*>(a.data + f.addr + (uint64_t)i * a.CU + (uint64_t)jj * a.U + x * w) =
          *reinterpret_cast<const Word*>(lds + k * a.fpl + s * a.rs + jj * (Idx)a.us + x * w);
    }
  }
}

// ---------------------------------------------------------------------------
// row pass: K rows of C cells per workgroup; LDS row k at k*rs
// ---------------------------------------------------------------------------
// Wide row pass (whole units, S == 1): the workgroup's nf rows are one
// contiguous run [g0, g0 + nf*CU), kept in LDS at its address modulo 16, so
// that the run's aligned 16-byte chunks land on aligned LDS chunks.  Loads
// and stores as in the wide column pass; output word b of the run is byte
// xu of cell j of row k, taken from cell row_src(i_k, j) of the same row.
template <typename Word, typename Idx, int MAP>
__device__ __forceinline__ void row_wide(const Args<Idx>& a, const Idx f0, const Idx nf,
                                         unsigned char* lds) {
  constexpr Idx w = sizeof(Word);
  constexpr int V = 16 / (int)sizeof(Word);
  const Idx tid = threadIdx.x, nt = blockDim.x;
  const uintptr_t d0 = (uintptr_t)a.data;
  const int64_t g0 = (int64_t)(a.base + (uint64_t)f0 * a.CU);
  const int64_t g1 = g0 + (int64_t)nf * (int64_t)a.CU;
  // first chunk, relative to data: may lie up to 15 bytes before it
  const int64_t a0 = g0 - (int64_t)((d0 + (uint64_t)g0) & 15);
  const Idx sh = (Idx)(g0 - a0);
  const Idx n = (Idx)((g1 - a0 + 15) / 16);
  for (Idx t0 = tid; t0 < n; t0 += nt * UNR) {
    W16 v[UNR];
    bool full[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const Idx t = t0 + (Idx)u * nt;
      full[u] = false;
      if (t >= n) continue;
      const int64_t cs = a0 + (int64_t)t * 16;
      if (cs >= (int64_t)a.base && cs + 16 <= (int64_t)a.hi) {
        v[u] = ld16(a.data + cs);
        full[u] = true;
      } else {
        const int64_t b0 = cs < g0 ? g0 : cs, b1 = cs + 16 < g1 ? cs + 16 : g1;
        for (int64_t b = b0; b < b1; b += w)
          *reinterpret_cast<Word*>(lds + (b - a0)) = *reinterpret_cast<const Word*>(a.data + b);
      }
    }
#pragma unroll
    for (int u = 0; u < UNR; ++u)
      if (full[u]) *reinterpret_cast<W16*>(lds + (t0 + (Idx)u * nt) * 16) = v[u];
  }
  __syncthreads();
  for (Idx t = tid; t < n; t += nt) {
    const int64_t cs = a0 + (int64_t)t * 16;
    const int64_t b0 = cs < g0 ? g0 : cs, b1 = cs + 16 < g1 ? cs + 16 : g1;
    const Idx x = (Idx)(b0 - g0);
    Idx k = a.fCU.div(x);
    const Idx xr = x - k * (Idx)a.CU;
    Idx j = a.fU.div(xr);
    Idx xu = xr - j * (Idx)a.U;
    Idx i = a.fR.mod(f0 + k);
    Idx src = row_src<Idx, MAP>(a, i, j);
    auto next = [&]() {
      xu += w;
      if (xu == (Idx)a.U) {
        xu = 0;
        if (++j == a.C) {
          j = 0;
          ++k;
          i = a.fR.mod(f0 + k);
        }
        if (k < nf) src = row_src<Idx, MAP>(a, i, j);
      }
    };
    if (b0 == cs && b1 == cs + 16) {
      Packed<Word, V> tmp;
#pragma unroll
      for (int q = 0; q < V; ++q) {
        tmp.e[q] = *reinterpret_cast<const Word*>(lds + sh + k * (Idx)a.CU + src * (Idx)a.U + xu);
        next();
      }
      *reinterpret_cast<typename Packed<Word, V>::P*>(a.data + cs) = tmp.p;
    } else {
      for (int64_t b = b0; b < b1; b += w) {
        *reinterpret_cast<Word*>(a.data + b) =
            *reinterpret_cast<const Word*>(lds + sh + k * (Idx)a.CU + src * (Idx)a.U + xu);
        next();
      }
    }
  }
}

template <typename Word, typename Idx, int MAP, int VEC>
__device__ __forceinline__ void row_body(const Args<Idx>& a, const Idx f0, unsigned char* lds) {
  const Idx nf = (a.nfp - f0) < a.K ? (a.nfp - f0) : a.K;
  const Idx tid = threadIdx.x, nt = blockDim.x;
  constexpr Idx w = sizeof(Word);
  if constexpr (sizeof(Word) < 16) {
    if (a.wide) {
      row_wide<Word, Idx, MAP>(a, f0, nf, lds);
      return;
    }
  }

  if (a.cpr) {
    // S == 1: the nf rows are one contiguous byte range (rs = C*U, no padding)
    const W16* g = reinterpret_cast<const W16*>(a.data + a.base + (uint64_t)f0 * a.CU);
    const Idx n = nf * a.cpr;
    for (Idx t0 = tid; t0 < n; t0 += nt * UNR) {
      W16 v[UNR];
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        const Idx t = t0 + (Idx)u * nt;
        if (t < n) v[u] = g[t];
      }
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        const Idx t = t0 + (Idx)u * nt;
        if (t < n) reinterpret_cast<W16*>(lds)[t] = v[u];
      }
    }
  } else {
    const Idx n = nf * a.ld_per;
    for (Idx t0 = tid; t0 < n; t0 += nt * UNR) {
      Word v[UNR];
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        const Idx t = t0 + (Idx)u * nt;
        if (t < n) {
          const Idx k = a.fld.div(t);
          const Idx rem = t - k * a.ld_per;
          const Idx q = a.wu == 1 ? rem : a.fwu.div(rem);
          const Idx x = rem - q * a.wu;
          const Foot<Idx, OP_ROW> f = decode<Idx, OP_ROW>(a, f0 + k);
          v[u] = *reinterpret_cast<const Word*>(a.data + f.addr + (uint64_t)q * a.U + x * w);
        }
      }
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        const Idx t = t0 + (Idx)u * nt;
        if (t < n) reinterpret_cast<Word*>(lds)[t] = v[u];
      }
    }
  }
  __syncthreads();

  if (VEC > 1) {
    // S == 1, one word per cell: VEC consecutive destination cells per store
    const Idx n = nf * a.st_per;
    char* gbase = a.data + a.base + (uint64_t)f0 * a.CU;
    for (Idx t = tid; t < n; t += nt) {
      const Idx k = a.fst.div(t);
      const Idx j0 = (t - k * a.st_per) * VEC;
      const Idx i = a.fR.mod(f0 + k);
      Packed<Word, VEC> tmp;
#pragma unroll
      for (int v = 0; v < VEC; ++v)
        tmp.e[v] = reinterpret_cast<const Word*>(lds)[k * a.C + row_src<Idx, MAP>(a, i, j0 + v)];
      *reinterpret_cast<typename Packed<Word, VEC>::P*>(gbase + (uint64_t)k * a.CU + (uint64_t)j0 * w) = tmp.p;
    }
  } else {
    const Idx n = nf * a.st_per;
    for (Idx t = tid; t < n; t += nt) {
      const Idx k = a.fst.div(t);
      const Idx rem = t - k * a.st_per;
      const Idx q = a.wu == 1 ? rem : a.fwu.div(rem);
      const Idx x = rem - q * a.wu;
      const Foot<Idx, OP_ROW> f = decode<Idx, OP_ROW>(a, f0 + k);
      const Idx s = row_src<Idx, MAP>(a, f.row, q);
      *reinterpret_cast<Word*>(a.data + f.addr + (uint64_t)q * a.U + x * w) =
          reinterpret_cast<const Word*>(lds)[k * a.st_per + s * a.wu + x];
    }
  }
}

// one workgroup per group of K footprints
template <typename Word, typename Idx, int VEC>
__global__ __launch_bounds__(512) void k_tile(Args<Idx> a, int remap) {
  extern __shared__ __align__(16) unsigned char lds[];
  tile_body<Word, Idx, VEC>(a, (Idx)xcd_logical(blockIdx.x, gridDim.x, remap) * a.K, lds);
}
template <typename Word, typename Idx, int MAP, int VEC>
__global__ __launch_bounds__(512) void k_col(Args<Idx> a, int remap) {
  extern __shared__ __align__(16) unsigned char lds[];
  col_body<Word, Idx, MAP, VEC>(a, (Idx)xcd_logical(blockIdx.x, gridDim.x, remap) * a.K, lds);
}
template <typename Word, typename Idx, int MAP, int VEC>
__global__ __launch_bounds__(512) void k_row(Args<Idx> a, int remap) {
  extern __shared__ __align__(16) unsigned char lds[];
  row_body<Word, Idx, MAP, VEC>(a, (Idx)xcd_logical(blockIdx.x, gridDim.x, remap) * a.K, lds);
}

// ---------------------------------------------------------------------------
// staged passes
// ---------------------------------------------------------------------------

// copy footprints [fp0, fp0+G) to scratch, footprint-major (cells in
// footprint order); for OP_TILE with S == 1 this is a contiguous copy
template <typename Word, typename Idx, int KIND>
__global__ __launch_bounds__(256) void k_stage_out(Args<Idx> a) {
  constexpr Idx w = sizeof(Word);
  const Idx NU = KIND == OP_TILE ? a.R * a.C : (KIND == OP_COL ? a.R * a.W : a.C);
  const uint64_t per = (uint64_t)NU * a.wu;
  const uint64_t total = a.G * per;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  Word* sc = reinterpret_cast<Word*>(a.scratch);
  for (uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t0 < total;
       t0 += stride * UNR) {
    Word v[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const uint64_t t = t0 + u * stride;
      if (t >= total) continue;
      const uint64_t fl = t / per;
      const Idx rem = (Idx)(t - fl * per);
      const Idx q = a.wu == 1 ? rem : a.fwu.div(rem);
      const Idx x = rem - q * a.wu;
      const Foot<Idx, KIND> f = decode<Idx, KIND>(a, (Idx)(a.fp0 + fl));
      uint64_t off;
      if (KIND == OP_COL) {
        const Idx i = a.fW.div(q);
        const Idx jj = q - i * a.W;
        if (jj >= f.weff) { v[u] = Word(); continue; }
        off = (uint64_t)i * a.CU + (uint64_t)jj * a.U;
      } else {
        off = (uint64_t)q * a.U;
      }
      v[u] = *reinterpret_cast<const Word*>(a.data + f.addr + off + x * w);
    }
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const uint64_t t = t0 + u * stride;
      if (t < total) sc[t] = v[u];
    }
  }
}

// gather the footprints back from scratch through the column/row maps
template <typename Word, typename Idx, int KIND, int MAP>
__global__ __launch_bounds__(256) void k_stage_in(Args<Idx> a) {
  constexpr Idx w = sizeof(Word);
  const Idx NU = KIND == OP_COL ? a.R * a.W : a.C;
  const uint64_t per = (uint64_t)NU * a.wu;
  const uint64_t total = a.G * per;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const Word* sc = reinterpret_cast<const Word*>(a.scratch);
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
    const uint64_t fl = t / per;
    const Idx rem = (Idx)(t - fl * per);
    const Idx q = a.wu == 1 ? rem : a.fwu.div(rem);
    const Idx x = rem - q * a.wu;
    const Foot<Idx, KIND> f = decode<Idx, KIND>(a, (Idx)(a.fp0 + fl));
    uint64_t off;
    Idx s;
    if (KIND == OP_COL) {
      const Idx i = a.fW.div(q);
      const Idx jj = q - i * a.W;
      if (jj >= f.weff) continue;
      off = (uint64_t)i * a.CU + (uint64_t)jj * a.U;
      s = col_src<Idx, MAP>(a, i, f.row + jj) * a.W + jj;
    } else {
      off = (uint64_t)q * a.U;
      s = row_src<Idx, MAP>(a, f.row, q);
    }
    *reinterpret_cast<Word*>(a.data + f.addr + off + x * w) = sc[fl * per + (uint64_t)s * a.wu + x];
  }
}

// Staged row pass of whole units (S == 1), part 2: the G rows of the group
// lie in scratch as one copy of their bytes; row fp0 + fl is gathered back
// through the row map, one aligned 16-byte output chunk per thread (V words,
// the units stepped one by one, not divided per word) and one 16-byte store;
// the row's two edge chunks are stored word by word.  grid.y strides rows.
template <typename Word, typename Idx, int MAP>
__global__ __launch_bounds__(256) void k_stage_in_row(Args<Idx> a) {
  constexpr Idx w = sizeof(Word);
  constexpr int V = 16 / sizeof(Word) > 1 ? 16 / sizeof(Word) : 1;
  const uintptr_t d0 = (uintptr_t)a.data;
  const int64_t CU = (int64_t)a.CU;
  const Idx U = (Idx)a.U;
  for (uint64_t fl = blockIdx.y; fl < a.G; fl += gridDim.y) {
    const Foot<Idx, OP_ROW> f = decode<Idx, OP_ROW>(a, (Idx)(a.fp0 + fl));
    const int64_t r0 = (int64_t)f.addr, r1 = r0 + CU;
    const int64_t a0 = r0 - (int64_t)((d0 + (uint64_t)r0) & 15);
    const uint64_t nch = (uint64_t)((r1 - a0 + 15) / 16);
    const char* sc = a.scratch + fl * (uint64_t)CU;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nch;
         t += (uint64_t)gridDim.x * blockDim.x) {
      const int64_t cs = a0 + 16 * (int64_t)t;
      const int64_t b0 = cs < r0 ? r0 : cs, b1 = cs + 16 < r1 ? cs + 16 : r1;
      const Idx x = (Idx)(b0 - r0);
      Idx q = a.fU.div(x);
      Idx xu = x - q * U;
      Idx src = row_src<Idx, MAP>(a, f.row, q);
      auto next = [&]() {
        xu += w;
        if (xu == U) {
          xu = 0;
          if (++q < a.C) src = row_src<Idx, MAP>(a, f.row, q);
        }
      };
      if (V > 1 && b0 == cs && b1 == cs + 16) {
        Packed<Word, V> tmp;
#pragma unroll
        for (int v = 0; v < V; ++v) {
          tmp.e[v] = *reinterpret_cast<const Word*>(sc + (uint64_t)src * U + xu);
          next();
        }
        *reinterpret_cast<typename Packed<Word, V>::P*>(a.data + cs) = tmp.p;
      } else {
        for (int64_t b = b0; b < b1; b += w) {
          *reinterpret_cast<Word*>(a.data + b) =
              *reinterpret_cast<const Word*>(sc + (uint64_t)src * U + xu);
          next();
        }
      }
    }
  }
}

// Staged row pass of whole units (S == 1), part 2, tiled through LDS.  The
// row-gather map has rowg(i, x + m) = rowg(i, x) + 1: units m apart map to
// consecutive units.  A tile is the index range [t0, t0 + T), T = m * L;
// rowg sends x = t0 + v + m*u (u < L) to L consecutive units for each v, so
// lanes ordered (v, u, word) move runs of L units on the rowg side:
//  - MAP_ROWG (unit x of the row <- scratch unit rowg(i, x)): word loads
//    from scratch in runs, LDS in row order, aligned 16-byte stores of the
//    tile's contiguous bytes;
//  - MAP_ROWD (unit x <- scratch unit rowd(i, x); rowd inverts rowg, so
//    scratch unit x goes to unit rowg(i, x)): aligned 16-byte loads of the
//    tile's contiguous scratch bytes, word stores in runs.
// Each unit's rowg is evaluated once, into LDS, so the result never depends
// on the runs being unbroken; grid.y strides the group's rows.
template <typename Word, typename Idx, int MAP>
__global__ __launch_bounds__(256) void k_stage_row_tile(Args<Idx> a, Idx L,
                                                        typename FDsel<Idx>::type fL) {
  extern __shared__ __align__(16) unsigned char lds[];
  constexpr Idx w = sizeof(Word);
  const Idx tid = threadIdx.x, nt = blockDim.x;
  const Idx U = (Idx)a.U, m = a.g.m, n = a.C, T = m * L, wpu = a.wu;
  Idx* gi = reinterpret_cast<Idx*>(lds);
  unsigned char* buf = lds + ((T * (Idx)sizeof(Idx) + 15) & ~(Idx)15);
  const uint64_t ntile = ((uint64_t)n + T - 1) / T;
  const uintptr_t d0 = (uintptr_t)a.data, s0 = (uintptr_t)a.scratch;
  const int64_t sc_end = (int64_t)(a.G * a.CU);
  for (uint64_t fl = blockIdx.y; fl < a.G; fl += gridDim.y) {
    const Foot<Idx, OP_ROW> f = decode<Idx, OP_ROW>(a, (Idx)(a.fp0 + fl));
    const int64_t sro = (int64_t)(fl * a.CU);
    for (uint64_t tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
      const Idx t0 = (Idx)(tile * T);
      const Idx tn = n - t0 < T ? n - t0 : T;
      const Idx nw = T * wpu;  // (v, u, word) items, those with v + m*u >= tn idle
      __syncthreads();         // the previous tile's LDS reads are done
      for (Idx e = tid; e < tn; e += nt) gi[e] = a.g.rowg(f.row, t0 + e);
      if (MAP == MAP_ROWG) {
        const int64_t g0 = (int64_t)f.addr + (int64_t)t0 * U, g1 = g0 + (int64_t)tn * U;
        const Idx sh = (Idx)((d0 + (uint64_t)g0) & 15);
        __syncthreads();
        for (Idx q0 = tid; q0 < nw; q0 += nt * UNR) {
          Word v[UNR];
          Idx e[UNR], xw[UNR];
#pragma unroll
          for (int k = 0; k < UNR; ++k) {
            const Idx q = q0 + (Idx)k * nt;
            e[k] = tn;
            if (q >= nw) continue;
            const Idx r = a.fwu.div(q);
            xw[k] = q - r * wpu;
            const Idx vv = fL.div(r);
            e[k] = vv + m * (r - vv * L);
            if (e[k] < tn)
              v[k] = *reinterpret_cast<const Word*>(a.scratch + sro + (int64_t)gi[e[k]] * U + xw[k] * w);
          }
#pragma unroll
          for (int k = 0; k < UNR; ++k)
            if (e[k] < tn) *reinterpret_cast<Word*>(buf + sh + e[k] * U + xw[k] * w) = v[k];
        }
        __syncthreads();
        const int64_t a0 = g0 - sh;
        const Idx nch = (Idx)((g1 - a0 + 15) / 16);
        for (Idx t = tid; t < nch; t += nt) {
          const int64_t cs = a0 + 16 * (int64_t)t;
          if (cs >= g0 && cs + 16 <= g1) {
            *reinterpret_cast<W16*>(a.data + cs) = *reinterpret_cast<const W16*>(buf + t * 16);
          } else {
            const int64_t b0 = cs < g0 ? g0 : cs, b1 = cs + 16 < g1 ? cs + 16 : g1;
            for (int64_t b = b0; b < b1; b += w)
              *reinterpret_cast<Word*>(a.data + b) = *reinterpret_cast<const Word*>(buf + (b - a0));
          }
        }
      } else {
        // scratch bytes of the tile, [g0, g1) from scratch, at their address mod 16
        const int64_t g0 = sro + (int64_t)t0 * U, g1 = g0 + (int64_t)tn * U;
        const Idx sh = (Idx)((s0 + (uint64_t)g0) & 15);
        const int64_t a0 = g0 - sh;
        const Idx nch = (Idx)((g1 - a0 + 15) / 16);
        for (Idx t = tid; t < nch; t += nt) {
          const int64_t cs = a0 + 16 * (int64_t)t;
          if (cs >= 0 && cs + 16 <= sc_end) {
            *reinterpret_cast<W16*>(buf + t * 16) = *reinterpret_cast<const W16*>(a.scratch + cs);
          } else {
            const int64_t b0 = cs < g0 ? g0 : cs, b1 = cs + 16 < g1 ? cs + 16 : g1;
            for (int64_t b = b0; b < b1; b += w)
              *reinterpret_cast<Word*>(buf + (b - a0)) = *reinterpret_cast<const Word*>(a.scratch + b);
          }
        }
        __syncthreads();
        for (Idx q = tid; q < nw; q += nt) {
          const Idx r = a.fwu.div(q);
          const Idx xw = q - r * wpu;
          const Idx vv = fL.div(r);
          const Idx e = vv + m * (r - vv * L);
          if (e < tn)
            *reinterpret_cast<Word*>(a.data + f.addr + (uint64_t)gi[e] * U + xw * w) =
                *reinterpret_cast<const Word*>(buf + sh + e * U + xw * w);
        }
      }
    }
  }
}

// LDS bytes of k_stage_row_tile for runs of L units: the indices, then the
// tile's bytes from their address mod 16 through the end of the last 16-byte
// chunk that holds any of them (whole chunks are moved)
inline int64_t stage_row_tile_lds(int64_t m, int64_t L, int64_t U, int64_t idx_bytes) {
  const int64_t T = m * L;
  return ((T * idx_bytes + 15) & ~(int64_t)15) + ((T * U + 30) & ~(int64_t)15);
}

// Run length of k_stage_row_tile for a staged row pass: tiles of about
// 16 KiB of LDS (units and their indices) and runs of at least 96 bytes,
// within 48 KiB; 0 when no such tile fits (the pass then gathers unit by
// unit).  Longer runs at the cost of larger tiles measured slower: 270-byte
// runs in 43 KiB tiles took 4.49 ms where 120-byte runs in 19 KiB took 3.13.
inline int64_t stage_row_tile_len(const Op& op, int64_t idx_bytes) {
  if (op.kind != OP_ROW || op.m < 1 || op.U < 1) return 0;
  int64_t L = std::max<int64_t>(16384 / (op.m * (op.U + idx_bytes)), (96 + op.U - 1) / op.U);
  if (L > op.C) L = op.C;
  return stage_row_tile_lds(op.m, L, op.U, idx_bytes) <= 48 * 1024 ? L : 0;
}

// staged tile, part 2: out-of-place LDS-tiled transpose of the footprints
// held in scratch (each an R x C row-major matrix of us-byte cells) back into
// the tensor, where output cell (j, i) of footprint f lives at
// addr(f) + (j*R + i)*U.  One workgroup per (footprint, TR x TC tile); TR and
// TC are powers of two.  Loads move 16-byte chunks of tile rows; stores pack
// VEC small cells (or move 16-byte chunks of large cells) along output rows.
struct TileGeom {
  uint32_t tr_log2, tc_log2;  // tile edge logs
  uint32_t tiles_r, tiles_c;  // tiles per footprint
  uint32_t rs;                // padded LDS row stride (bytes)
  uint32_t cpr;               // 16-byte chunks per full tile row (0: word loads)
};

template <typename Word, typename Idx, int VEC>
__global__ __launch_bounds__(256) void k_tile_oop(Args<Idx> a, TileGeom tg) {
  extern __shared__ __align__(16) unsigned char lds[];
  constexpr Idx w = sizeof(Word);
  const Idx tid = threadIdx.x, nt = blockDim.x;
  const uint64_t per_fp = (uint64_t)tg.tiles_r * tg.tiles_c;
  const uint64_t tile = blockIdx.x + (uint64_t)blockIdx.y * gridDim.x;
  if (tile >= a.G * per_fp) return;
  const uint64_t fl = tile / per_fp;
  const uint32_t tt = (uint32_t)(tile - fl * per_fp);
  const uint32_t ti = tt / tg.tiles_c, tj = tt - ti * tg.tiles_c;
  const Idx TR = (Idx)1 << tg.tr_log2, TC = (Idx)1 << tg.tc_log2;
  const Idx i0 = ti * TR, j0 = tj * TC;
  const Idx nr = (a.R - i0) < TR ? (a.R - i0) : TR;
  const Idx nc = (a.C - j0) < TC ? (a.C - j0) : TC;
  const unsigned char* src =
      reinterpret_cast<const unsigned char*>(a.src) + fl * (uint64_t)a.R * a.C * a.us;
  // ---- load tile rows i0..i0+nr, cells j0..j0+nc (nc*us contiguous bytes each)
  if (tg.cpr && nc == TC) {
    const Idx n = nr * tg.cpr;
    for (Idx t0 = tid; t0 < n; t0 += nt * UNR) {
      W16 v[UNR];
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        const Idx t = t0 + (Idx)u * nt;
        if (t < n) {
          const Idx r = t / tg.cpr, c = t - (t / tg.cpr) * tg.cpr;
          v[u] = *reinterpret_cast<const W16*>(src + ((uint64_t)(i0 + r) * a.C + j0) * a.us + c * 16);
        }
      }
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        const Idx t = t0 + (Idx)u * nt;
        if (t < n) {
          const Idx r = t / tg.cpr, c = t - (t / tg.cpr) * tg.cpr;
          lds_put16(lds + r * tg.rs + c * 16, v[u]);
        }
      }
    }
  } else {
    const Idx wpr = nc * a.wu;  // words per tile row
    const Idx n = nr * wpr;
    for (Idx t = tid; t < n; t += nt) {
      const Idx r = t / wpr, c = t - (t / wpr) * wpr;
      *reinterpret_cast<Word*>(lds + r * tg.rs + c * w) =
          *reinterpret_cast<const Word*>(src + ((uint64_t)(i0 + r) * a.C + j0) * a.us + c * w);
    }
  }
  __syncthreads();
  // ---- store output rows j0..j0+nc, cells i0..i0+nr
  const Foot<Idx, OP_TILE> f = decode<Idx, OP_TILE>(a, (Idx)(a.fp0 + fl));
  if (VEC > 1 && nr == TR) {
    // cells of one word, contiguous output rows (S == 1): VEC cells per store
    const Idx vpr = TR / VEC;
    const Idx n = nc * vpr;
    for (Idx t = tid; t < n; t += nt) {
      const Idx jr = t / vpr, iv = (t - (t / vpr) * vpr) * VEC;
      Packed<Word, VEC> tmp;
#pragma unroll
      for (int v = 0; v < VEC; ++v)
        tmp.e[v] = *reinterpret_cast<const Word*>(lds + (iv + v) * tg.rs + jr * w);
      *reinterpret_cast<typename Packed<Word, VEC>::P*>(
          a.data + f.addr + ((uint64_t)(j0 + jr) * a.R + i0 + iv) * a.U) = tmp.p;
    }
  } else {
    const Idx cw = a.wu;
    const Idx n = nc * nr * cw;
    const Idx per_row = nr * cw;
    for (Idx t = tid; t < n; t += nt) {
      const Idx jr = t / per_row, rem = t - (t / per_row) * per_row;
      const Idx ir = rem / cw, x = rem - (rem / cw) * cw;
      *reinterpret_cast<Word*>(a.data + f.addr + ((uint64_t)(j0 + jr) * a.R + i0 + ir) * a.U +
                               x * w) =
          *reinterpret_cast<const Word*>(lds + ir * tg.rs + jr * a.us + x * w);
    }
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

template <typename Idx>
Args<Idx> make_args(const Op& op, char* data, int64_t wbytes) {
  Args<Idx> a{};
  a.data = data;
  a.base = op.base;
  a.U = op.U;
  a.us = op.us;
  a.RCU = (uint64_t)op.R * op.C * op.U;
  a.CU = (uint64_t)op.C * op.U;
  a.R = (Idx)op.R;
  a.C = (Idx)op.C;
  a.W = (Idx)(op.kind == OP_COL ? op.W : 1);
  a.S = (Idx)(op.U / op.us);
  a.nfp = (Idx)op.num_footprints();
  a.P = (Idx)((op.C + a.W - 1) / a.W);
  a.wu = (Idx)(op.us / wbytes);
  a.K = 1;
  a.fS.init(a.S);
  a.fR.init(op.R);
  a.fC.init(op.C);
  a.fW.init(a.W);
  a.fP.init(a.P);
  a.fwu.init(a.wu);
  a.fU.init(op.kind == OP_COL ? op.us : op.U);  // wide and staged-row paths: slice or unit
  if (op.kind != OP_TILE) a.g.init(op.m, op.n, op.c, op.a, op.b, op.k, op.binv);
  return a;
}

bool needs64(const Op& op) {
  const unsigned __int128 lim = (unsigned __int128)1 << 31;
  auto big = [&](unsigned __int128 v) { return v >= lim; };
  if (big((unsigned __int128)op.num_footprints() * 2)) return true;
  if (big((unsigned __int128)op.footprint_units() * op.us * 2)) return true;
  if (op.kind == OP_TILE) return big((unsigned __int128)op.R * op.C * 2);
  const unsigned __int128 m = op.m, n = op.n, a = op.a;
  if (big(n + 2 * m) || big(m * 3)) return true;
  // the products each map forms (maps.hpp): rot/rotinv i + j/b; colsh
  // i * (n mod m) + j; colshinv a * a and a * c; rowg q * b < n and
  // (j/c) * k < b * b, rowd c * rho + t < n + m and r * a < a * b, the two
  // products reduced mod b in 64 bits whatever the index width
  if (op.kind == OP_COL) return big(m * m + n) || big(a * a);
  return false;
}

int64_t word_bytes(const Op& op, uintptr_t data) {
  int64_t w = 16;
  while (w > 1 && ((op.us % w) || (op.U % w) || ((data + op.base) % w))) w /= 2;
  return w;
}

// Footprints per workgroup: as many as fit the LDS target, but no fewer
// than min_groups workgroups per pass (a pass launched alone wants at least
// this many footprints in flight across the GPU)
int64_t pick_K(int64_t nfp, int64_t fpb, const LaunchCfg& cfg) {
  int64_t K = std::max<int64_t>(1, cfg.lds_target / std::max<int64_t>(fpb, 1));
  const int64_t mg = std::max<int64_t>(1, cfg.min_groups);
  if (nfp / K < mg) K = std::max<int64_t>(1, nfp / mg);
  return K;
}

// Kernel variant of an LDS pass: kind/map (0 tile, 1..4 column maps, 5..6
// row maps), log2 of the word size, and whether cells are packed into
// 16-byte vector stores.
inline int kindmap_of(const Op& op) {
  if (op.kind == OP_TILE) return 0;
  if (op.kind == OP_COL) return op.map;  // MAP_ROT..MAP_COLSHINV = 1..4
  return op.map;                         // MAP_ROWG, MAP_ROWD = 5, 6
}

struct Prepared {
  int variant = -1;  // kindmap * 10 + log2(word) * 2 + vec
  int idx64 = 0;
  int64_t K = 1;
  int64_t lds = 0;   // LDS bytes per workgroup
  int threads = 256;
  int64_t groups = 0;
};

// Column and row passes over whole units narrower than 16-byte words take
// the wide path, unless their rows are whole aligned 16-byte chunks of
// one-word units (then 16-byte loads and packed stores apply as they are).
// The choice depends on the op only (its offsets as if the tensor were
// 16-byte aligned), so the LDS size is the same for any tensor address.
constexpr int64_t kWideLdsMax = 160 * 1024;
inline bool wide_ok(const Op& op) {
  if (op.staged) return false;
  if (op.kind != OP_COL && op.kind != OP_ROW) return false;
  // rows of a column pass may be one slice of a unit (one column per panel)
  if (op.us != op.U && !(op.kind == OP_COL && op.W == 1)) return false;
  const int64_t w = word_bytes(op, 0);
  if (w >= 16) return false;
  const int64_t V = 16 / w;
  const bool aligned = op.base % 16 == 0 && (op.C * op.U) % 16 == 0 && op.U == w && op.us == op.U;
  if (op.kind == OP_ROW) return !(aligned && op.C % V == 0);
  return !(aligned && op.W % V == 0 && op.C % op.W == 0);
}

// Tile passes of whole units in 1- or 2-byte words take the wide path
// (tile_wide), unless one-word cells already pack into 16-byte stores over
// rows of whole 16-byte chunks.  As for column and row passes the choice
// depends on the op only.
inline bool tile_wide_ok(const Op& op) {
  if (op.staged || op.kind != OP_TILE || op.us != op.U) return false;
  const int64_t w = word_bytes(op, 0);
  if (w > 2) return false;
  const bool packs = op.U == w && op.R % (16 / w) == 0 && op.base % 16 == 0 && (op.C * op.U) % 16 == 0;
  return !packs;
}

// LDS bytes of one copy of a wide tile's run of K matrices of F bytes (the
// run plus the partial chunks at its two ends), and of both copies
inline int64_t tile_wide_copy(int64_t K, int64_t F) { return (K * F + 32 + 15) / 16 * 16; }

// Distinct 4-byte LDS banks (of 64) that 64 lanes hit with accesses `stride`
// bytes apart.
inline int lds_banks_hit(int64_t stride) {
  uint64_t seen = 0;
  for (int64_t l = 0; l < 64; ++l) seen |= 1ull << ((l * stride / 4) % 64);
  return __builtin_popcountll(seen);
}

// Fills every LDS-kernel field of `a` for op (not staged) under cfg.
template <typename Idx>
void prepare_lds(const Op& op, char* data, const LaunchCfg& cfg, Args<Idx>& a, Prepared& p) {
  const int64_t w = word_bytes(op, (uintptr_t)data);
  a = make_args<Idx>(op, data, w);
  int wl = 0;
  while ((1 << wl) < w) ++wl;
  const int64_t V = w < 16 ? 16 / w : 1;
  const bool one_word = a.wu == 1 && a.S == 1;
  const uintptr_t base = (uintptr_t)(data + op.base);
  const int km = kindmap_of(op);
  bool vec = false;
  if (op.kind == OP_TILE) {
    vec = V > 1 && one_word && op.R % V == 0 && base % 16 == 0;
    const int64_t rowbytes = op.C * op.us;
    a.rs = (Idx)tile_row_stride(rowbytes, w);
    a.fpl = (Idx)(op.R * a.rs);
    const bool chunk16 = a.S == 1 && rowbytes % 16 == 0 && base % 16 == 0 && (a.rs % 4) == 0;
    a.cpr = chunk16 ? (Idx)(rowbytes / 16) : 0;
    const bool cells16 = !chunk16 && !vec && w == 16 && op.us % 16 == 0 && op.U % 16 == 0 &&
                         base % 16 == 0 && a.rs % 16 == 0;
    a.cps = cells16 ? (Idx)(op.us / 16) : 0;
    a.ld_per = a.cps ? a.R * a.C * a.cps : a.R * a.C * a.wu;
    a.vpr = vec ? (Idx)(op.R / V) : 1;
    a.st_per = vec ? (Idx)(op.C * a.vpr) : a.R * a.C * a.wu;
    const int64_t F = op.R * op.C * op.U;
    if (tile_wide_ok(op) && tile_wide_copy(1, F) <= kWideLdsMax) {
      // two LDS copies when they fit: input order unless output order
      // spreads the copies over more banks (input order: consecutive threads
      // write R*U bytes apart; output order: they read C*U bytes apart);
      // else one copy, gathered from directly (wide 3)
      if (2 * tile_wide_copy(1, F) <= kWideLdsMax)
        a.wide = lds_banks_hit(op.C * op.U) > lds_banks_hit(op.R * op.U) ? 2 : 1;
      else
        a.wide = 3;
      a.cpr = 0;
      a.cps = 0;
      vec = false;
      a.fpl = (Idx)F;
      a.ld_per = a.st_per = a.R * a.C;
    }
  } else if (op.kind == OP_COL) {
    vec = V > 1 && one_word && op.W % V == 0 && op.C % op.W == 0 && base % 16 == 0 &&
          (op.C * op.U) % 16 == 0;
    const int64_t seg = op.W * op.us;
    a.rs = (Idx)seg;
    a.fpl = (Idx)(op.R * seg);
    const bool chunk16 = op.C % op.W == 0 && seg % 16 == 0 && base % 16 == 0 && (a.CU % 16) == 0 &&
                         (op.us == op.U ? true : (op.W == 1 && op.us % 16 == 0));
    a.cpr = chunk16 ? (Idx)(seg / 16) : 0;
    a.cps = 0;
    a.ld_per = chunk16 ? a.R * a.cpr : a.R * a.W * a.wu;
    a.vpr = vec ? (Idx)(op.W / V) : 1;
    a.st_per = vec ? a.R * a.vpr : a.R * a.W * a.wu;
    // whole units that do not move as 16-byte chunks take the wide path
    // when its LDS rows (16 bytes of slack each) fit; the choice depends on
    // the op only, so the LDS size is the same for any tensor address
    const int64_t rsw = (seg + 15 + 15) / 16 * 16;
    if (wide_ok(op) && op.R * rsw <= kWideLdsMax) {
      a.wide = 1;
      a.rs = (Idx)rsw;
      a.fpl = (Idx)(op.R * rsw);
      a.ncw = (Idx)(rsw / 16);
      a.ld_per = a.st_per = a.R * a.ncw;
      a.cpr = 0;
      vec = false;
    }
  } else {
    vec = V > 1 && one_word && op.C % V == 0 && base % 16 == 0;
    a.rs = (Idx)(op.C * op.us);
    a.fpl = a.rs;
    const bool chunk16 = a.S == 1 && (a.rs % 16) == 0 && base % 16 == 0;
    a.cpr = chunk16 ? (Idx)(a.rs / 16) : 0;
    a.cps = 0;
    a.ld_per = a.C * a.wu;
    a.vpr = vec ? (Idx)(op.C / V) : 1;
    a.st_per = vec ? a.vpr : a.C * a.wu;
    if (wide_ok(op) && op.C * op.U + 32 <= kWideLdsMax) {
      a.wide = 1;
      a.cpr = 0;
      vec = false;
    }
  }
  a.hi = op.base + op.bytes();
  a.fU.init(op.kind == OP_COL ? op.us : op.U);  // the wide paths' unit: slice or unit
  a.fCU.init(op.C * op.U);
  a.fncw.init(a.ncw ? a.ncw : 1);
  a.cps_log2 = -1;
  if (a.cps && (a.cps & (a.cps - 1)) == 0) {
    a.cps_log2 = 0;
    while (((Idx)1 << a.cps_log2) < a.cps) ++a.cps_log2;
  }
  a.fcpr.init(a.cpr ? a.cpr : 1);
  a.fcps.init(a.cps ? a.cps : 1);
  a.fld.init(a.ld_per);
  a.fst.init(a.st_per);
  a.fvpr.init(a.vpr);
  // wide tiles: copies of the run of K matrices (two, or one for wide 3)
  const int copies = (a.wide && op.kind == OP_TILE) ? (a.wide == 3 ? 1 : 2) : 0;
  p.K = pick_K((int64_t)a.nfp, (copies ? copies : 1) * (int64_t)a.fpl, cfg);
  if (a.wide && op.kind == OP_ROW)  // the run of K rows and its two partial chunks
    while (p.K > 1 && p.K * (int64_t)a.fpl + 32 > kWideLdsMax) --p.K;
  if (copies)
    while (p.K > 1 && copies * tile_wide_copy(p.K, a.fpl) > kWideLdsMax) --p.K;
  a.K = (Idx)p.K;
  p.lds = p.K * (int64_t)a.fpl + (a.wide && op.kind == OP_ROW ? 32 : 0);
  if (copies) p.lds = copies * tile_wide_copy(p.K, a.fpl);
  p.threads = p.lds > 96 * 1024 ? cfg.threads_big : cfg.threads;
  p.groups = ((int64_t)a.nfp + p.K - 1) / p.K;
  p.variant = km * 10 + wl * 2 + (vec ? 1 : 0);
  p.idx64 = sizeof(Idx) == 8;
}

// calls F::template run<Word, KINDMAP, VEC>() for a variant id
template <int KM, int WL, int VF, typename F>
hipError_t with_variant3(F& f) {
  using Word = typename std::conditional<
      WL == 4, W16,
      typename std::conditional<
          WL == 3, uint64_t,
          typename std::conditional<WL == 2, uint32_t,
                                    typename std::conditional<WL == 1, uint16_t, uint8_t>::type>::
              type>::type>::type;
  constexpr int V = (16 >> WL) > 1 ? (16 >> WL) : 1;
  return f.template run<Word, KM, VF ? V : 1>();
}
template <int KM, int WL, typename F>
hipError_t with_variant2(int vf, F& f) {
  return vf ? with_variant3<KM, WL, (WL < 4 ? 1 : 0)>(f) : with_variant3<KM, WL, 0>(f);
}
template <int KM, typename F>
hipError_t with_variant1(int wl, int vf, F& f) {
  switch (wl) {
    case 0: return with_variant2<KM, 0>(vf, f);
    case 1: return with_variant2<KM, 1>(vf, f);
    case 2: return with_variant2<KM, 2>(vf, f);
    case 3: return with_variant2<KM, 3>(vf, f);
    default: return with_variant2<KM, 4>(vf, f);
  }
}
template <typename F>
hipError_t with_variant(int variant, F& f) {
  const int km = variant / 10, wl = (variant % 10) / 2, vf = variant % 2;
  switch (km) {
    case 0: return with_variant1<0>(wl, vf, f);
    case 1: return with_variant1<1>(wl, vf, f);
    case 2: return with_variant1<2>(wl, vf, f);
    case 3: return with_variant1<3>(wl, vf, f);
    case 4: return with_variant1<4>(wl, vf, f);
    case 5: return with_variant1<5>(wl, vf, f);
    default: return with_variant1<6>(wl, vf, f);
  }
}

template <typename Idx>
struct LdsLauncher {
  const Args<Idx>& a;
  const Prepared& p;
  const LaunchCfg& cfg;
  hipStream_t st;
  template <typename Word, int KM, int V>
  hipError_t run() {
    // standalone launches: enough threads for one round of UNR 16-byte loads
    // of the workgroup's footprints when the knob asks for it
    int threads = p.threads;
    if (cfg.lds_threads_auto) {
      const int64_t words = (p.lds + 15) / 16;
      threads = (int)std::min<int64_t>(512, std::max<int64_t>(64, ((words + UNR - 1) / UNR + 63) / 64 * 64));
    }
    const dim3 grid((uint32_t)p.groups), block(threads);
    if (KM == 0)
      hipLaunchKernelGGL((k_tile<Word, Idx, V>), grid, block, (size_t)p.lds, st, a, cfg.xcd_remap);
    else if (KM <= 4)
      hipLaunchKernelGGL((k_col<Word, Idx, (KM <= 4 ? KM : 1), V>), grid, block, (size_t)p.lds, st, a,
                         cfg.xcd_remap);
    else
      hipLaunchKernelGGL((k_row<Word, Idx, (KM >= 5 ? KM : 5), V>), grid, block, (size_t)p.lds, st, a,
                         cfg.xcd_remap);
    return hipGetLastError();
  }
};

hipError_t launch_copy16(char* dst, const char* src, uint64_t len, const char* slo, const char* shi,
                         hipStream_t st);

template <typename Word, typename Idx, int KIND, int MAP>
hipError_t launch_staged(Args<Idx> a, const Op& op, char* scratch, int64_t scratch_bytes,
                         hipStream_t st) {
  const int64_t fpb = op.footprint_units() * op.us;
  if (!scratch || scratch_bytes < fpb) return hipErrorInvalidValue;
  a.scratch = scratch;
  a.src = scratch;
  const uint64_t nfp = (uint64_t)a.nfp;
  const uint64_t G = std::min<uint64_t>(nfp, (uint64_t)(scratch_bytes / fpb));
  for (uint64_t f0 = 0; f0 < nfp; f0 += G) {
    a.fp0 = f0;
    a.G = std::min<uint64_t>(G, nfp - f0);
    const uint64_t words = a.G * (uint64_t)op.footprint_units() * a.wu;
    const uint64_t blocks = std::min<uint64_t>((words + 255) / 256, 256ull * 64);
    if ((KIND == OP_TILE || KIND == OP_ROW) && a.S == 1) {
      // whole matrices, or whole rows: one contiguous range
      const uint64_t fpb_all = KIND == OP_TILE ? a.RCU : a.CU;
      hipError_t e = launch_copy16(scratch, a.data + a.base + f0 * fpb_all, a.G * fpb_all,
                                   a.data + op.base, a.data + op.base + op.bytes(), st);
      if (e != hipSuccess) return e;
    } else {
      hipLaunchKernelGGL((k_stage_out<Word, Idx, KIND>), dim3((uint32_t)blocks), dim3(256), 0, st, a);
    }
    if (KIND == OP_TILE) {
      // tile edges: about 16 KiB of cells, 16-byte rows at least
      uint32_t trl = 6, tcl = 6;
      while (trl > 0 && ((int64_t)1 << (trl + tcl)) * op.us > 16384) {
        --trl;
        --tcl;
      }
      while (trl > 0 && ((int64_t)1 << (trl - 1)) >= op.R) --trl;
      while (tcl > 0 && ((int64_t)1 << (tcl - 1)) >= op.C) --tcl;
      // a side cut to a short matrix gives its share back to the other one
      while (((int64_t)1 << (trl + tcl + 1)) * op.us <= 16384 && ((int64_t)1 << tcl) < op.C) ++tcl;
      while (((int64_t)1 << (trl + tcl + 1)) * op.us <= 16384 && ((int64_t)1 << trl) < op.R) ++trl;
      TileGeom tg;
      tg.tr_log2 = trl;
      tg.tc_log2 = tcl;
      tg.tiles_r = (uint32_t)((op.R + (1 << trl) - 1) >> trl);
      tg.tiles_c = (uint32_t)((op.C + (1 << tcl) - 1) >> tcl);
      const int64_t rowb = ((int64_t)1 << tcl) * op.us;
      tg.rs = (uint32_t)tile_row_stride(rowb, sizeof(Word));
      tg.cpr = (rowb % 16 == 0 && (op.C * op.us) % 16 == 0 && tg.rs % 4 == 0) ? (uint32_t)(rowb / 16) : 0;
      const uint64_t ntiles = a.G * (uint64_t)tg.tiles_r * tg.tiles_c;
      const size_t lds = (size_t)((int64_t)1 << trl) * tg.rs;
      const uint32_t gx = (uint32_t)std::min<uint64_t>(ntiles, 65535);
      const uint32_t gy = (uint32_t)((ntiles + gx - 1) / gx);
      constexpr int V = 16 / sizeof(Word) > 1 ? 16 / sizeof(Word) : 1;
      const bool vec = V > 1 && a.wu == 1 && a.S == 1 && ((int64_t)1 << trl) % V == 0 &&
                       op.R % V == 0 && ((uintptr_t)(a.data + a.base) % 16) == 0;
      if (vec)
        hipLaunchKernelGGL((k_tile_oop<Word, Idx, V>), dim3(gx, gy), dim3(256), lds, st, a, tg);
      else
        hipLaunchKernelGGL((k_tile_oop<Word, Idx, 1>), dim3(gx, gy), dim3(256), lds, st, a, tg);
    } else if (KIND == OP_ROW && a.S == 1 && stage_row_tile_len(op, sizeof(Idx)) > 0) {
      const int64_t L = stage_row_tile_len(op, sizeof(Idx));
      const uint64_t ntile = ((uint64_t)op.C + op.m * L - 1) / (uint64_t)(op.m * L);
      const uint32_t gx = (uint32_t)std::min<uint64_t>(ntile, 1u << 20);
      const uint32_t gy = (uint32_t)std::min<uint64_t>(a.G, 65535);
      typename FDsel<Idx>::type fL;
      fL.init((uint64_t)L);
      hipLaunchKernelGGL((k_stage_row_tile<Word, Idx, (KIND == OP_ROW ? MAP : MAP_ROWG)>), dim3(gx, gy),
                         dim3(256), (size_t)stage_row_tile_lds(op.m, L, op.U, sizeof(Idx)), st, a,
                         (Idx)L, fL);
    } else if (KIND == OP_ROW && a.S == 1) {
      const uint64_t nch = (a.CU + 31) / 16;  // a row's 16-byte chunks, edges included
      const uint32_t gx = (uint32_t)std::min<uint64_t>((nch + 255) / 256, 1024);
      const uint32_t gy = (uint32_t)std::min<uint64_t>(a.G, 65535);
      hipLaunchKernelGGL((k_stage_in_row<Word, Idx, (KIND == OP_ROW ? MAP : MAP_ROWG)>), dim3(gx, gy),
                         dim3(256), 0, st, a);
    } else {
      hipLaunchKernelGGL((k_stage_in<Word, Idx, KIND, MAP>), dim3((uint32_t)blocks), dim3(256), 0, st,
                         a);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

template <typename Idx>
struct StagedLauncher {
  const Op& op;
  Args<Idx>& a;
  char* s;
  int64_t sb;
  hipStream_t st;
  template <typename Word, int KM, int V>
  hipError_t run() {
    if (KM == 0) return launch_staged<Word, Idx, OP_TILE, MAP_NONE>(a, op, s, sb, st);
    if (KM <= 4) return launch_staged<Word, Idx, OP_COL, (KM <= 4 ? KM : 1)>(a, op, s, sb, st);
    return launch_staged<Word, Idx, OP_ROW, (KM >= 5 ? KM : 5)>(a, op, s, sb, st);
  }
};

template <typename Idx>
hipError_t run_op(const Op& op, char* d, char* s, int64_t sb, const LaunchCfg& c, hipStream_t st) {
  if (op.staged) {
    const int64_t w = word_bytes(op, (uintptr_t)d);
    Args<Idx> a = make_args<Idx>(op, d, w);
    StagedLauncher<Idx> L{op, a, s, sb, st};
    int wl = 0;
    while ((1 << wl) < w) ++wl;
    return with_variant(kindmap_of(op) * 10 + wl * 2, L);
  }
  Args<Idx> a;
  Prepared p;
  prepare_lds<Idx>(op, d, c, a, p);
  LdsLauncher<Idx> L{a, p, c, st};
  return with_variant(p.variant, L);
}


// ---------------------------------------------------------------------------
// persistent batched pass: one launch runs one pass of many tensors; each
// workgroup walks a table of (op, first footprint) items
// ---------------------------------------------------------------------------

// ---- pattern fill / verify ----------------------------------------------------

__device__ __forceinline__ uint64_t sm64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

uint64_t sm64_host(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__device__ __forceinline__ uint64_t pattern(uint64_t seed_mixed, uint64_t i, int mode) {
  if (mode == 0) return i;
  return sm64(seed_mixed ^ (i * 0xD1B54A32D192ED03ull));
}

template <typename T>
__global__ void k_fill(T* p, uint64_t n, uint64_t seed_mixed, int mode, uint64_t first) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    p[i] = (T)pattern(seed_mixed, first + i, mode);
}

struct VerifyArgs {
  int rank;
  FD64 fd[16];
  uint64_t ostride[16];
  uint64_t vol;
};

template <typename T>
__global__ void k_verify(const T* p, VerifyArgs v, uint64_t seed_mixed, int mode,
                         unsigned long long* out) {
  __shared__ unsigned int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  unsigned int mine = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < v.vol; i += stride) {
    uint64_t rem = i, d = 0;
    for (int k = 0; k < v.rank; ++k) {
      const uint64_t q = v.fd[k].div(rem);
      d += (rem - q * v.fd[k].d) * v.ostride[k];
      rem = q;
    }
    if (p[d] != (T)pattern(seed_mixed, i, mode)) ++mine;
  }
  if (mine) atomicAdd(&cnt, mine);
  __syncthreads();
  if (threadIdx.x == 0 && cnt) atomicAdd(out, (unsigned long long)cnt);
}

}  // namespace

int64_t op_lds_bytes(const Op& op, const LaunchCfg& cfg) {
  if (op.staged || op.kind == 0 || op.kind == OP_CYCLE) return 0;
  if (op.kind == OP_BLOCK) return op.blk ? (op.blk->lds_bytes() + 15) / 16 * 16 : op.R * op.U;
  if (op.kind == OP_SEGTILE) {
    int64_t uw = 1;
    while (uw < 16 && op.U % (uw * 2) == 0) uw *= 2;
    return (op.R * tile_row_stride(op.C * op.U, uw) + 15) / 16 * 16 +
           (op.segb > 0 ? op.R * op.C * op.U / op.segb * 8 : 0);
  }
  Args<uint64_t> a;
  Prepared p;
  prepare_lds<uint64_t>(op, nullptr, cfg, a, p);
  return p.lds;
}

namespace {

// ---------------------------------------------------------------------------
// cycle-following pass (OP_CYCLE): whole cells move along the cycles of the
// R x C transpose.  One wave per (matrix, segment, chunk of 64 words of the
// cell); every lane owns one word of the cell at every position, so lanes
// never share data and no LDS or barrier is needed.  A wave walks its
// segment p -> src(p) writing cell(p) <- old cell(src(p)), with D loads in
// flight: each store to a position waits for the load of the cell that
// follows it, and loads return in order, so the position's own old cell has
// been read before it is overwritten.  The last write of a segment takes the
// cell saved in a boundary slot by k_cycle_save (cut cycles) or the cell the
// wave read from its first position before writing it (whole cycles).
// ---------------------------------------------------------------------------
struct CycMapDev {
  uint32_t bc, R, C, m;
  FD32 fbc, fm, fR;
};
struct CycArgs {
  char* data;
  uint64_t base, RCU, U;
  const CycleSeg* segs;
  const uint32_t* slots;
  char* boundary;
  uint32_t nseg, nslots, nmaps, chunk_bytes;
  uint32_t lane_log2;  // log2 of the lanes per item (6: one item per wave)
  uint64_t B, items, chunks;
  CycMapDev maps[kMaxCycleMaps];
  FD64 fchunks, fnseg;
};

// source position of cell position p: the pass's maps composed in gather
// order (NM bounds nmaps at compile time: kernels instantiated for fewer maps
// keep the others' constants out of their scalar registers)
template <int NM = kMaxCycleMaps>
__device__ __forceinline__ uint32_t cyc_src(const CycArgs& a, uint32_t p) {
#pragma unroll
  for (int k = NM - 1; k >= 0; --k) {
    if (k >= (int)a.nmaps) continue;
    const CycMapDev& M = a.maps[k];
    const uint32_t b = M.fbc.div(p), q = p - b * M.bc;
    const uint32_t c = M.fm.div(q), x = q - c * M.m;
    const uint32_t j = M.fR.div(c), i = c - j * M.R;
    p = b * M.bc + (i * M.C + j) * M.m + x;
  }
  return p;
}

template <typename Word>
__global__ __launch_bounds__(256) void k_cycle_save(CycArgs a) {
  constexpr uint64_t w = sizeof(Word);
  const uint64_t wpc = a.U / w;
  const uint64_t total = a.B * a.nslots * wpc;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
    const uint64_t cell = t / wpc;  // boundary cell = b * nslots + slot
    const uint64_t x = t - cell * wpc;
    const uint64_t b = cell / a.nslots;
    const uint64_t s = cell - b * a.nslots;
    *reinterpret_cast<Word*>(a.boundary + cell * a.U + x * w) = *reinterpret_cast<const Word*>(
        a.data + a.base + b * a.RCU + (uint64_t)a.slots[s] * a.U + x * w);
  }
}

// global accesses with optional non-temporal hints (NT bit 0: loads, bit 1: stores)
typedef uint32_t U32x4 __attribute__((ext_vector_type(4)));
template <int NT, typename Word>
__device__ __forceinline__ Word ld_nt(const char* p) {
  if constexpr ((NT & 1) && sizeof(Word) == 16) {
    return __builtin_nontemporal_load(reinterpret_cast<const U32x4*>(p));
  } else if constexpr ((NT & 1) && sizeof(Word) < 16) {
    return __builtin_nontemporal_load(reinterpret_cast<const Word*>(p));
  } else {
    return *reinterpret_cast<const Word*>(p);
  }
}
template <int NT, typename Word>
__device__ __forceinline__ void st_nt(char* p, const Word& w) {
  if constexpr ((NT & 2) && sizeof(Word) == 16) {
    __builtin_nontemporal_store(w, reinterpret_cast<U32x4*>(p));
  } else if constexpr ((NT & 2) && sizeof(Word) < 16) {
    __builtin_nontemporal_store(w, reinterpret_cast<Word*>(p));
  } else {
    *reinterpret_cast<Word*>(p) = w;
  }
}

// gtid: the thread's index among all threads of the pass
template <typename Word, int D, int NT>
__device__ __forceinline__ void cycle_body(const CycArgs& a, const uint64_t gtid) {
  constexpr uint64_t w = sizeof(Word);
  // an item is 2^lane_log2 consecutive lanes: a whole wave, or a group of
  // lanes when cells are smaller than 64 words (several items per wave)
  const uint32_t lane = (uint32_t)(gtid & ((1u << a.lane_log2) - 1));
  const uint64_t item = gtid >> a.lane_log2;
  if (item >= a.items) return;
  const uint64_t rest = a.fchunks.div(item);
  const uint64_t chunk = item - rest * a.chunks;
  const uint64_t b = a.fnseg.div(rest);
  const uint64_t s = rest - b * a.nseg;
  const uint64_t off = chunk * a.chunk_bytes + lane * w;
  if (off >= a.U) return;
  char* m = a.data + a.base + b * a.RCU + off;
  const CycleSeg sg = a.segs[s];
  uint32_t p = sg.start;
  const Word temp =
      sg.src_slot < 0
          ? *reinterpret_cast<const Word*>(m + (uint64_t)p * a.U)
          : *reinterpret_cast<const Word*>(a.boundary + (b * a.nslots + (uint64_t)sg.src_slot) * a.U +
                                           off);
  for (uint32_t t = 0; t < sg.len; t += D) {
    Word v[D];
    uint32_t q[D];
    uint32_t pp = p;
#pragma unroll
    for (int k = 0; k < D; ++k) q[k] = pp = cyc_src(a, pp);
#pragma unroll
    for (int k = 0; k < D; ++k)
      if (t + k < sg.len)
        v[k] = t + k + 1 == sg.len ? temp : ld_nt<NT, Word>(m + (uint64_t)q[k] * a.U);
    uint32_t dst = p;
#pragma unroll
    for (int k = 0; k < D; ++k)
      if (t + k < sg.len) {
        st_nt<NT, Word>(m + (uint64_t)dst * a.U, v[k]);
        dst = q[k];
      }
    p = q[D - 1];
  }
}

template <typename Word, int D, int NT = 0>
__global__ __launch_bounds__(256) void k_cycle(CycArgs a) {
  cycle_body<Word, D, NT>(a, (uint64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// Boundary save of a cycle pass, one word per thread: gtid indexes the
// words of all B * nslots saved cells.
template <typename Word>
__device__ __forceinline__ void cycle_save_word(const CycArgs& a, const uint64_t t) {
  constexpr uint64_t w = sizeof(Word);
  const uint64_t wpc = a.U / w;
  if (t >= a.B * a.nslots * wpc) return;
  const uint64_t cell = t / wpc;  // boundary cell = b * nslots + slot
  const uint64_t x = t - cell * wpc;
  const uint64_t b = cell / a.nslots;
  const uint64_t sl = cell - b * a.nslots;
  *reinterpret_cast<Word*>(a.boundary + cell * a.U + x * w) = *reinterpret_cast<const Word*>(
      a.data + a.base + b * a.RCU + (uint64_t)a.slots[sl] * a.U + x * w);
}

// Pipelined variant: every lane moves V words per position (a wave covers
// V*64 words of the cell, each load one coalesced row of the wave), and the
// loads of the next D positions are issued before the stores of the current
// D, so up to 2*D*V loads per lane are in flight.  Hazards as above: a
// store to a position follows the completed load of that position's old
// cell, and loads of batch t+1 never touch positions stored in batch t.
template <typename Word, int D, int V>
__global__ __launch_bounds__(256) void k_cycle_pipe(CycArgs a) {
  constexpr uint64_t w = sizeof(Word);
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t item = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (item >= a.items) return;
  const uint64_t rest = a.fchunks.div(item);
  const uint64_t chunk = item - rest * a.chunks;
  const uint64_t b = a.fnseg.div(rest);
  const uint64_t s = rest - b * a.nseg;
  const uint64_t off0 = chunk * a.chunk_bytes + lane * w;
  if (off0 >= a.U) return;
  bool on[V];
#pragma unroll
  for (int v = 0; v < V; ++v) on[v] = off0 + (uint64_t)v * 64 * w < a.U;
  char* m = a.data + a.base + b * a.RCU + off0;
  const CycleSeg sg = a.segs[s];
  const char* tsrc = sg.src_slot < 0
                         ? m + (uint64_t)sg.start * a.U
                         : a.boundary + (b * a.nslots + (uint64_t)sg.src_slot) * a.U + off0;
  Word temp[V];
#pragma unroll
  for (int v = 0; v < V; ++v)
    if (on[v]) temp[v] = *reinterpret_cast<const Word*>(tsrc + v * 64 * w);
  auto next = [&](uint32_t pp) { return cyc_src(a, pp); };
  Word cur[D][V], nxt[D][V];
  uint32_t qc[D], qn[D];
  uint32_t p = sg.start;
  // batch 0: old cells of positions 1..D
  {
    uint32_t pp = p;
#pragma unroll
    for (int k = 0; k < D; ++k) qc[k] = pp = next(pp);
#pragma unroll
    for (int k = 0; k < D; ++k)
#pragma unroll
      for (int v = 0; v < V; ++v)
        if ((uint32_t)k < sg.len && on[v])
          cur[k][v] = (uint32_t)k + 1 == sg.len
                          ? temp[v]
                          : *reinterpret_cast<const Word*>(m + (uint64_t)qc[k] * a.U + v * 64 * w);
  }
  for (uint32_t t = 0; t < sg.len; t += D) {
    // loads of batch t + D
    uint32_t pp = qc[D - 1];
#pragma unroll
    for (int k = 0; k < D; ++k) qn[k] = pp = next(pp);
#pragma unroll
    for (int k = 0; k < D; ++k)
#pragma unroll
      for (int v = 0; v < V; ++v)
        if (t + D + k < sg.len && on[v])
          nxt[k][v] = t + D + k + 1 == sg.len
                          ? temp[v]
                          : *reinterpret_cast<const Word*>(m + (uint64_t)qn[k] * a.U + v * 64 * w);
    // stores of batch t
    uint32_t dst = p;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      if (t + k < sg.len) {
#pragma unroll
        for (int v = 0; v < V; ++v)
          if (on[v]) *reinterpret_cast<Word*>(m + (uint64_t)dst * a.U + v * 64 * w) = cur[k][v];
      }
      dst = qc[k];
    }
    p = qc[D - 1];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      qc[k] = qn[k];
#pragma unroll
      for (int v = 0; v < V; ++v) cur[k][v] = nxt[k][v];
    }
  }
}

// The counter that table uploads on this thread are charged to.
thread_local std::atomic<int64_t>* t_account = nullptr;

// Owner of a device allocation of `bytes` on device `dev`: its deleter
// waits for that device to go idle before freeing, so a pass enqueued
// earlier never reads freed memory.  The caller's current device is
// restored.  The bytes are charged to this thread's table account, if any,
// until freed.
std::shared_ptr<void> device_owned(void* p, int dev, int64_t bytes) {
  std::atomic<int64_t>* acct = t_account;
  if (acct) acct->fetch_add(bytes, std::memory_order_relaxed);
  return std::shared_ptr<void>(p, [dev, bytes, acct](void* q) {
    int cur = 0;
    const bool have = hipGetDevice(&cur) == hipSuccess;
    if (have && cur != dev) (void)hipSetDevice(dev);
    (void)hipDeviceSynchronize();
    (void)hipFree(q);
    if (have && cur != dev) (void)hipSetDevice(cur);
    if (acct) acct->fetch_sub(bytes, std::memory_order_relaxed);
  });
}

const void* cycle_table_device(const CycleTable& t, hipError_t* err) {
  *err = hipSuccess;
  int dev = 0;
  if ((*err = hipGetDevice(&dev)) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(t.dev.mu);
  auto it = t.dev.by_dev.find(dev);
  if (it != t.dev.by_dev.end()) return it->second.get();
  const size_t sb = t.segs.size() * sizeof(CycleSeg), lb = t.slots.size() * sizeof(uint32_t);
  void* p = nullptr;
  if ((*err = hipMalloc(&p, std::max<size_t>(sb + lb, 16))) != hipSuccess) return nullptr;
  std::shared_ptr<void> own = device_owned(p, dev, (int64_t)std::max<size_t>(sb + lb, 16));
  if (sb && (*err = hipMemcpy(p, t.segs.data(), sb, hipMemcpyHostToDevice)) != hipSuccess)
    return nullptr;
  if (lb && (*err = hipMemcpy((char*)p + sb, t.slots.data(), lb, hipMemcpyHostToDevice)) !=
                hipSuccess)
    return nullptr;
  t.dev.by_dev[dev] = own;
  return p;
}

template <typename Word>
hipError_t launch_cycle_w(CycArgs a, int variant, hipStream_t stream) {
  if (a.nslots) {
    const uint64_t words = a.B * a.nslots * (a.U / sizeof(Word));
    const uint64_t blocks = std::min<uint64_t>((words + 255) / 256, 256ull * 64);
    hipLaunchKernelGGL((k_cycle_save<Word>), dim3((uint32_t)blocks), dim3(256), 0, stream, a);
  }
  const uint64_t blocks = ((a.items << a.lane_log2) + 255) / 256;
  if (a.lane_log2 < 6 && variant >= 1 && variant <= 3)  // one item per wave only
    hipLaunchKernelGGL((k_cycle<Word, 8>), dim3((uint32_t)blocks), dim3(256), 0, stream, a);
  else if (variant == 1)
    hipLaunchKernelGGL((k_cycle_pipe<Word, 4, 2>), dim3((uint32_t)blocks), dim3(256), 0, stream, a);
  else if (variant == 2)
    hipLaunchKernelGGL((k_cycle_pipe<Word, 4, 1>), dim3((uint32_t)blocks), dim3(256), 0, stream, a);
  else if (variant == 3)
    hipLaunchKernelGGL((k_cycle_pipe<Word, 2, 4>), dim3((uint32_t)blocks), dim3(256), 0, stream, a);
  else if (variant == 4)
    hipLaunchKernelGGL((k_cycle<Word, 8, 2>), dim3((uint32_t)blocks), dim3(256), 0, stream, a);
  else if (variant == 5)
    hipLaunchKernelGGL((k_cycle<Word, 8, 3>), dim3((uint32_t)blocks), dim3(256), 0, stream, a);
  else if (variant == 6)
    hipLaunchKernelGGL((k_cycle<Word, 8, 1>), dim3((uint32_t)blocks), dim3(256), 0, stream, a);
  else if (variant == 7)
    hipLaunchKernelGGL((k_cycle<Word, 16, 0>), dim3((uint32_t)blocks), dim3(256), 0, stream, a);
  else
    hipLaunchKernelGGL((k_cycle<Word, 8>), dim3((uint32_t)blocks), dim3(256), 0, stream, a);
  return hipGetLastError();
}

// lanes' words per position of each cycle kernel variant
inline int cycle_vwords(int variant) { return variant == 1 ? 2 : variant == 3 ? 4 : 1; }
// the variants that take several items per wave (cells under 64 words)
inline bool cycle_small_ok(int variant) { return variant == 0 || variant >= 4; }

// Arguments of a cycle pass over the tensor at `data` with its boundary
// cells in `scratch`; *w receives the word size, *variant the kernel shape
// it can use.  *empty is set when the pass moves nothing (fixed points only).
hipError_t prepare_cycle_args(const Op& op, char* data, char* scratch, int64_t scratch_bytes,
                              int* variant, CycArgs& a, int64_t* w_out, bool* empty,
                              const void* table = nullptr) {
  *empty = false;
  if (!op.cyc || (uint64_t)op.R * (uint64_t)op.C >= (1ull << 32)) return hipErrorInvalidValue;
  const CycleTable& t = *op.cyc;
  if (t.segs.empty()) {
    *empty = true;
    return hipSuccess;
  }
  const uint64_t need = (uint64_t)op.B * t.slots.size() * (uint64_t)op.U;
  if (need && (!scratch || (uint64_t)scratch_bytes < need)) return hipErrorInvalidValue;
  hipError_t e = hipSuccess;
  const char* d = table ? (const char*)table : (const char*)cycle_table_device(t, &e);
  if (!d) return e;
  const int64_t w = word_bytes(op, (uintptr_t)data);
  *w_out = w;
  a.data = data;
  a.base = op.base;
  a.RCU = (uint64_t)op.R * op.C * op.U;
  a.U = op.U;
  a.segs = reinterpret_cast<const CycleSeg*>(d);
  a.slots = reinterpret_cast<const uint32_t*>(d + t.segs.size() * sizeof(CycleSeg));
  a.boundary = scratch;
  a.nseg = (uint32_t)t.segs.size();
  a.nslots = (uint32_t)t.slots.size();
  if (t.maps.empty() || t.maps.size() > (size_t)kMaxCycleMaps) return hipErrorInvalidValue;
  a.nmaps = (uint32_t)t.maps.size();
  for (size_t k = 0; k < (size_t)kMaxCycleMaps; ++k) {
    const CycleMap M = k < t.maps.size() ? t.maps[k] : CycleMap{1, 1, 1, 1};
    a.maps[k].bc = (uint32_t)M.bc;
    a.maps[k].R = (uint32_t)M.R;
    a.maps[k].C = (uint32_t)M.C;
    a.maps[k].m = (uint32_t)M.m;
    a.maps[k].fbc.init((uint32_t)M.bc);
    a.maps[k].fm.init((uint32_t)M.m);
    a.maps[k].fR.init((uint32_t)M.R);
  }
  // cells smaller than a wave's row of words: 2^lane_log2 lanes per cell
  a.lane_log2 = 6;
  while (a.lane_log2 > 0 && ((uint64_t)w << (a.lane_log2 - 1)) >= (uint64_t)op.U) --a.lane_log2;
  if (a.lane_log2 < 6 && !cycle_small_ok(*variant)) *variant = 0;
  a.chunk_bytes = (uint32_t)((w << a.lane_log2) * cycle_vwords(*variant));
  a.B = op.B;
  a.chunks = (op.U + a.chunk_bytes - 1) / a.chunk_bytes;
  a.items = a.B * a.nseg * a.chunks;
  a.fchunks.init(a.chunks);
  a.fnseg.init(a.nseg);
  if (((a.items << a.lane_log2) + 255) / 256 >= (1ull << 31)) return hipErrorInvalidValue;
  return hipSuccess;
}

// ---------------------------------------------------------------------------
// Cycle following of cells that are not whole aligned 16-byte words
// (k_cycle_u).  An item owns the cell-relative bytes [r0, r0 + n) of every
// cell of its segment, r0 = g*G with G at most (lanes - 2) * 16 bytes and a
// whole number of words, so items that share cells never touch each other's
// bytes and every range starts and ends on a word.  At each position the item's
// lanes cover the aligned 16-byte chunks of the destination's range (at
// most lanes - 1 of them): lane l takes chunk X = (first chunk) + 16*l.
// Its bytes come from the source range at the same cell-relative offsets:
// the aligned source chunk at or below X + (src - dst) and the one after
// it, funnel-shifted by their misalignment (the same for the whole item).
// Every lane loads one source chunk and takes the next one from the lane
// above it with a wave shuffle (ds_bpermute), so each source byte is
// requested once.  Bytes read outside the range are never stored.  A
// destination chunk wholly inside the range is stored with one 16-byte
// store, the range's end chunks as at most six aligned pieces of 8, 4, 2 and
// 1 bytes, none narrower than a word.  Offsets are 32-bit when the pass
// spans less than 2 GB.
// Source chunks that would reach outside the pass's bytes (or the boundary
// area) are gathered byte by byte.  The step order and hazards are those
// of k_cycle: within a wave, every load of a batch completes before any of
// its stores.
// Items are lanes consecutive lanes of one wave, per_wave of them to a wave
// (any lane count: an item's last lane only hands its chunk down, so the
// chunk it receives from the next item's first lane is never used).
// ---------------------------------------------------------------------------
struct CycUArgs {
  uint64_t G;         // cell bytes per item
  uint64_t groups;    // items per (matrix, segment)
  uint32_t lanes;     // lanes per item, ceil(G / 16) + 2 (at most 64)
  uint32_t per_wave;  // items per wave, 64 / lanes
  FD64 fgroups;
  FD32 flanes;
};

// bytes sig .. sig+15 of the 32 bytes a:b: dwords q .. q+4 (q = sig / 4) of
// a:b are picked by a shift of 2 dwords and then of 1 (two levels of
// selects), then funnel-shifted by sig % 4 bytes
__device__ __forceinline__ W16 funnel16(const W16& a, const W16& b, uint32_t sig) {
  const uint32_t r = sig & 3;
  const bool h = (sig & 8) != 0, o = (sig & 4) != 0;
  const uint32_t e0 = h ? a.z : a.x, e1 = h ? a.w : a.y, e2 = h ? b.x : a.z;
  const uint32_t e3 = h ? b.y : a.w, e4 = h ? b.z : b.x, e5 = h ? b.w : b.y;
  const uint32_t f0 = o ? e1 : e0, f1 = o ? e2 : e1, f2 = o ? e3 : e2;
  const uint32_t f3 = o ? e4 : e3, f4 = o ? e5 : e4;
  W16 out;
  out.x = __builtin_amdgcn_alignbyte(f1, f0, r);
  out.y = __builtin_amdgcn_alignbyte(f2, f1, r);
  out.z = __builtin_amdgcn_alignbyte(f3, f2, r);
  out.w = __builtin_amdgcn_alignbyte(f4, f3, r);
  return out;
}

// Device copy of len bytes from src to dst (ranges apart), in aligned
// 16-byte stores: output chunk t holds dst bytes [D0 + 16t, D0 + 16t + 16),
// D0 = dst rounded down to 16, taken as bytes [sig, sig + 16) of the aligned
// source chunks at S0 + 16t and S0 + 16t + 16, where S0 + sig is the source
// byte of output byte D0.  Source chunks are loaded whole only within
// [slo, shi), bytes known to exist; output chunks at the range's ends, and
// any whose source chunks leave [slo, shi), move byte by byte.
__global__ __launch_bounds__(256) void k_copy16(char* dst, const char* src, uint64_t len,
                                                const char* slo, const char* shi) {
  constexpr int CU = 4;  // chunks per thread per step, loads issued together
  const int64_t D = (int64_t)(uintptr_t)dst, end = D + (int64_t)len;
  const int64_t D0 = D & ~(int64_t)15;
  const int64_t S = (int64_t)(uintptr_t)src - (D - D0);
  const int64_t S0 = S & ~(int64_t)15;
  const uint32_t sig = (uint32_t)(S - S0);
  const int64_t lo = (int64_t)(uintptr_t)slo, hi = (int64_t)(uintptr_t)shi;
  const uint64_t nch = (uint64_t)((end - D0 + 15) / 16);
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t0 < nch; t0 += stride * CU) {
    W16 c[CU], nx[CU];
    bool whole[CU];
#pragma unroll
    for (int u = 0; u < CU; ++u) {
      const uint64_t t = t0 + (uint64_t)u * stride;
      const int64_t oc = D0 + 16 * (int64_t)t, sc = S0 + 16 * (int64_t)t;
      whole[u] = t < nch && oc >= D && oc + 16 <= end && sc >= lo && sc + 32 <= hi;
      if (whole[u]) {
        c[u] = *reinterpret_cast<const W16*>(sc);
        if (sig) nx[u] = *reinterpret_cast<const W16*>(sc + 16);
      }
    }
#pragma unroll
    for (int u = 0; u < CU; ++u) {
      const uint64_t t = t0 + (uint64_t)u * stride;
      if (t >= nch) continue;
      const int64_t oc = D0 + 16 * (int64_t)t;
      if (whole[u]) {
        *reinterpret_cast<W16*>(oc) = sig ? funnel16(c[u], nx[u], sig) : c[u];
      } else {
        const int64_t b0 = oc < D ? D : oc, b1 = oc + 16 < end ? oc + 16 : end;
        for (int64_t b = b0; b < b1; ++b)
          *reinterpret_cast<char*>(b) = *reinterpret_cast<const char*>(b - D + (int64_t)(uintptr_t)src);
      }
    }
  }
}

hipError_t launch_copy16(char* dst, const char* src, uint64_t len, const char* slo, const char* shi,
                         hipStream_t st) {
  if (!len) return hipSuccess;
  const uint64_t nch = (len + 31) / 16;
  const uint64_t blocks = std::min<uint64_t>((nch + 1023) / 1024, 256ull * 64);
  hipLaunchKernelGGL(k_copy16, dim3((uint32_t)blocks), dim3(256), 0, st, dst, src, len, slo, shi);
  return hipGetLastError();
}

// 16 bytes from offset x of base, only those within [lo, hi) read (the
// others are zero)
__device__ __forceinline__ W16 gather16(const char* base, int64_t x, int64_t lo, int64_t hi) {
  uint32_t d[4] = {0, 0, 0, 0};
#pragma unroll
  for (int j = 0; j < 16; ++j)
    if (x + j >= lo && x + j < hi)
      d[j >> 2] |= (uint32_t)(unsigned char)base[x + j] << (8 * (j & 3));
  W16 o;
  o.x = d[0];
  o.y = d[1];
  o.z = d[2];
  o.w = d[3];
  return o;
}

// the next lane's chunk within groups of 2^ll lanes (the group's last lane
// gets its own back)
__device__ __forceinline__ W16 shfl_next(const W16& v, int width) {
  W16 o;
  o.x = __shfl_down(v.x, 1, width);
  o.y = __shfl_down(v.y, 1, width);
  o.z = __shfl_down(v.z, 1, width);
  o.w = __shfl_down(v.w, 1, width);
  return o;
}

// Stores the bytes of chunk v (at offset X from data, 16-byte aligned) that
// lie in [x0, x1), x0 < x1 within the chunk and both multiples of the word:
// 16 bytes at once when that is the whole chunk, else as aligned pieces of
// 8, 4, 2 and 1 bytes (at most six; none narrower than the word).
template <typename Word, typename Off>
__device__ __forceinline__ void cycu_store(char* data, Off X, Off x0, Off x1, const W16& v) {
  constexpr uint32_t w = sizeof(Word);
  if (x0 == X && x1 == X + 16) {
    *reinterpret_cast<W16*>(data + X) = v;
    return;
  }
  uint32_t o = (uint32_t)(x0 - X);
  const uint32_t e = (uint32_t)(x1 - X);
  char* c = data + X;
  auto dw = [&](uint32_t k) -> uint32_t { return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w; };
  auto st1 = [&](uint32_t q) { c[q] = (char)(dw(q >> 2) >> (8 * (q & 3))); };
  auto st2 = [&](uint32_t q) { *reinterpret_cast<uint16_t*>(c + q) = (uint16_t)(dw(q >> 2) >> (8 * (q & 3))); };
  auto st4 = [&](uint32_t q) { *reinterpret_cast<uint32_t*>(c + q) = dw(q >> 2); };
  auto st8 = [&](uint32_t q) {
    *reinterpret_cast<uint64_t*>(c + q) =
        q ? ((uint64_t)v.w << 32 | v.z) : ((uint64_t)v.y << 32 | v.x);
  };
  if (w <= 1 && (o & 1) && o + 1 <= e) { st1(o); o += 1; }
  if (w <= 2 && (o & 2) && o + 2 <= e) { st2(o); o += 2; }
  if (w <= 4 && (o & 4) && o + 4 <= e) { st4(o); o += 4; }
  if ((o & 8) && o + 8 <= e) { st8(o); o += 8; }
  if (o + 8 <= e) { st8(o); o += 8; }
  if (w <= 4 && o + 4 <= e) { st4(o); o += 4; }
  if (w <= 2 && o + 2 <= e) { st2(o); o += 2; }
  if (w <= 1 && o + 1 <= e) { st1(o); }
}

// Word: the widest word dividing the cell and the tensor's address (below
// 16 bytes); Off: int32_t when every offset of the pass (and of the
// boundary area) stays below 2^31, else int64_t.
// gtid: the thread's index among all threads of the pass (the lanes of an
// item are consecutive threads of one wave)
template <typename Word, typename Off, int D, int NM = kMaxCycleMaps>
__device__ __forceinline__ void cycle_u_body(const CycArgs& a, const CycUArgs& u, const uint64_t gtid) {
  const uint32_t lw = (uint32_t)(gtid & 63);
  const uint32_t sub = u.flanes.div(lw);
  const uint32_t lane = lw - sub * u.lanes;
  const uint64_t item = (gtid >> 6) * u.per_wave + sub;
  // every lane of a live item takes part in the shuffles: an item's lanes
  // are consecutive lanes of one wave, and they exit together (lanes past
  // the wave's last item and items past the end feed no live lane's store)
  if (sub >= u.per_wave || item >= a.items) return;
  const uint64_t rest = u.fgroups.div(item);
  const uint64_t g = item - rest * u.groups;
  const uint64_t b = a.fnseg.div(rest);
  const uint64_t sidx = rest - b * a.nseg;
  const Off U = (Off)a.U;
  const Off r0 = (Off)(g * u.G);
  const Off n = U - r0 < (Off)u.G ? U - r0 : (Off)u.G;
  const Off m = (Off)(a.base + b * a.RCU) + r0;  // byte r0 of cell 0 of the matrix
  const Off lo = (Off)a.base, hi = (Off)(a.base + a.B * a.RCU);
  const Off bhi = (Off)(a.B * a.nslots * a.U);
  const uint32_t dmod = (uint32_t)((uintptr_t)a.data & 15), bmod = (uint32_t)((uintptr_t)a.boundary & 15);
  const Off l16 = (Off)(16 * lane);
  // lane's destination chunk of the range that starts at offset dr (from data)
  auto chunk = [&](Off dr) -> Off { return dr - (Off)((dmod + (uint32_t)dr) & 15) + l16; };
  // the source chunk this lane loads for destination range dr taking the
  // range at sr of sb (address mod 16: smod; valid bytes [slo, shi)); *s0:
  // its offset, *sig: the misalignment of the lane's destination chunk in
  // it.  A lane's chunk feeds its own destination chunk and the one of the
  // lane below it, so only lanes whose chunk starts before dr + n + 16 load.
  auto load = [&](Off dr, const char* sb, uint32_t smod, Off sr, Off slo, Off shi, Off* s0,
                  uint32_t* sig) -> W16 {
    const Off X = chunk(dr);
    const Off xs = X - dr + sr;
    *sig = (smod + (uint32_t)xs) & 15;
    *s0 = xs - (Off)*sig;
    W16 v = {0, 0, 0, 0};
    if (X < dr + n + 16 && *s0 >= slo && *s0 + 16 <= shi) v = *reinterpret_cast<const W16*>(sb + *s0);
    return v;
  };
  // the 16 bytes of the lane's destination chunk from its loaded chunk c
  // (all lanes of the item call this together: it shuffles); a lane whose
  // destination chunk starts at or past dr + n stores nothing and gathers
  // nothing
  auto assemble = [&](const W16& c, Off dr, Off s0, uint32_t sig, const char* sb, Off slo,
                      Off shi) -> W16 {
    const W16 nx = shfl_next(c, 64);
    if (s0 >= slo && s0 + 32 <= shi) return sig ? funnel16(c, nx, sig) : c;
    if (chunk(dr) >= dr + n) return c;
    return gather16(sb, (int64_t)(s0 + (Off)sig), (int64_t)slo, (int64_t)shi);
  };
  const CycleSeg sg = a.segs[sidx];
  uint32_t p = sg.start;
  // a whole cycle's last write takes the old start cell: read it first
  W16 fin = {0, 0, 0, 0};
  if (sg.src_slot < 0) {
    const Off dl = m + (Off)sg.last * U;
    Off s0;
    uint32_t sig;
    const W16 c = load(dl, a.data, dmod, m + (Off)p * U, lo, hi, &s0, &sig);
    fin = assemble(c, dl, s0, sig, a.data, lo, hi);
  }
  const char* bnd = a.boundary;
  const Off bsrc = sg.src_slot < 0 ? 0 : (Off)((b * a.nslots + (uint64_t)sg.src_slot) * a.U) + r0;
  for (uint32_t t = 0; t < sg.len; t += D) {
    uint32_t q[D];
    uint32_t pp = p;
#pragma unroll
    for (int k = 0; k < D; ++k) q[k] = pp = cyc_src<NM>(a, pp);
    W16 v[D];
    Off s0[D];
    uint32_t sig[D];
    // the batch's loads (the last write of a cut segment reads its saved cell)
    uint32_t dst = p;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      if (t + k < sg.len && !(t + k + 1 == sg.len && sg.src_slot < 0)) {
        const Off dr = m + (Off)dst * U;
        if (t + k + 1 < sg.len)
          v[k] = load(dr, a.data, dmod, m + (Off)q[k] * U, lo, hi, &s0[k], &sig[k]);
        else
          v[k] = load(dr, bnd, bmod, bsrc, 0, bhi, &s0[k], &sig[k]);
      }
      dst = q[k];
    }
    dst = p;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      if (t + k < sg.len) {
        const Off dr = m + (Off)dst * U;
        if (t + k + 1 < sg.len)
          v[k] = assemble(v[k], dr, s0[k], sig[k], a.data, lo, hi);
        else if (sg.src_slot < 0)
          v[k] = fin;
        else
          v[k] = assemble(v[k], dr, s0[k], sig[k], bnd, 0, bhi);
      }
      dst = q[k];
    }
    dst = p;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      if (t + k < sg.len) {
        const Off dr = m + (Off)dst * U;
        const Off X = chunk(dr);
        const Off x0 = X < dr ? dr : X, x1 = X + 16 < dr + n ? X + 16 : dr + n;
        if (x0 < x1) cycu_store<Word, Off>(a.data, X, x0, x1, v[k]);
      }
      dst = q[k];
    }
    p = q[D - 1];
  }
}

template <typename Word, typename Off, int D, int NM = kMaxCycleMaps>
__global__ __launch_bounds__(256) void k_cycle_u(CycArgs a, CycUArgs u) {
  cycle_u_body<Word, Off, D, NM>(a, u, (uint64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// Item geometry of the shifted cycle kernel for cells of w-byte words: fills
// u and a.items.  Lanes per item: G <= (lanes - 2) * 16 cell bytes need at
// most lanes - 1 destination chunks, so the last lane only hands its chunk
// down; cells larger than that are cut into equal groups of whole words (the
// last one may be shorter; gmax is a multiple of 16, so G stays within it).
// Returns whether every offset the kernel forms (within [-16, the end of the
// pass or of the boundary area + 64 lanes' chunks + 32)) fits 32 bits.
bool cycle_u_geometry(CycArgs& a, CycUArgs& u, uint64_t w) {
  const uint64_t gmax = 62 * 16;
  u.groups = (a.U + gmax - 1) / gmax;
  u.G = ((a.U + u.groups - 1) / u.groups + w - 1) / w * w;
  u.lanes = (uint32_t)((u.G + 15) / 16 + 2);
  u.per_wave = 64 / u.lanes;
  u.flanes.init(u.lanes);
  u.fgroups.init(u.groups);
  a.items = a.B * a.nseg * u.groups;
  const uint64_t span = std::max<uint64_t>(a.base + a.B * a.RCU, a.B * a.nslots * a.U) + 2048;
  return span < (1ull << 31);
}

// threads of a shifted cycle pass: whole waves of per_wave items each
inline uint64_t cycu_threads(const CycArgs& a, const CycUArgs& u) {
  return (a.items + u.per_wave - 1) / u.per_wave * 64;
}

template <typename Word>
hipError_t launch_cycle_u(CycArgs a, int depth, hipStream_t stream) {
  if (a.nslots) {
    const uint64_t words = a.B * a.nslots * (a.U / sizeof(Word));
    const uint64_t blocks = std::min<uint64_t>((words + 255) / 256, 256ull * 64);
    hipLaunchKernelGGL((k_cycle_save<Word>), dim3((uint32_t)blocks), dim3(256), 0, stream, a);
  }
  CycUArgs u;
  const bool off32 = cycle_u_geometry(a, u, sizeof(Word));
  const uint64_t blocks = (cycu_threads(a, u) + 255) / 256;
  if (blocks >= (1ull << 31)) return hipErrorInvalidValue;
  const dim3 grid((uint32_t)blocks), block(256);
  if (off32) {
    // passes of one or two composed maps (most) at the default depth: the
    // kernel instantiated for two maps
    if (depth == 4 && a.nmaps <= 2)
      hipLaunchKernelGGL((k_cycle_u<Word, int32_t, 4, 2>), grid, block, 0, stream, a, u);
    else if (depth == 4) hipLaunchKernelGGL((k_cycle_u<Word, int32_t, 4>), grid, block, 0, stream, a, u);
    else if (depth == 6) hipLaunchKernelGGL((k_cycle_u<Word, int32_t, 6>), grid, block, 0, stream, a, u);
    else hipLaunchKernelGGL((k_cycle_u<Word, int32_t, 8>), grid, block, 0, stream, a, u);
  } else {
    if (depth == 4) hipLaunchKernelGGL((k_cycle_u<Word, int64_t, 4>), grid, block, 0, stream, a, u);
    else if (depth == 6) hipLaunchKernelGGL((k_cycle_u<Word, int64_t, 6>), grid, block, 0, stream, a, u);
    else hipLaunchKernelGGL((k_cycle_u<Word, int64_t, 8>), grid, block, 0, stream, a, u);
  }
  return hipGetLastError();
}

hipError_t launch_cycle(const Op& op, char* data, char* scratch, int64_t scratch_bytes,
                        int variant, int depth, hipStream_t stream, const void* table) {
  CycArgs a;
  int64_t w = 1;
  bool empty = false;
  hipError_t e =
      prepare_cycle_args(op, data, scratch, scratch_bytes, &variant, a, &w, &empty, table);
  if (e != hipSuccess || empty) return e;
  if (w < 16 && variant == 0) {  // cells that are not whole 16-byte words
    switch (w) {
      // cells of 8-byte words run the 4-byte kernel (95 instead of 112
      // VGPRs: 5 waves per SIMD instead of 4); its edge pieces still include
      // aligned 8-byte stores
      case 8:
      case 4: return launch_cycle_u<uint32_t>(a, depth, stream);
      case 2: return launch_cycle_u<uint16_t>(a, depth, stream);
      default: return launch_cycle_u<uint8_t>(a, depth, stream);
    }
  }
  // cells of at least 512 bytes of 16-byte words: non-temporal loads and
  // stores (variant 5), measured 1.7 % faster on 512-byte cells and 2-7 %
  // on larger ones; smaller cells lose up to 70 % with them (neighbouring
  // cells share cache lines that the hints evict).  Smaller cells in passes
  // of at least 512 MiB: non-temporal stores only (variant 4), 2-6 % faster
  // on 128-464-byte cells of 0.8-4 GB passes; a 231 MB pass, which the
  // 256 MB last-level cache could hold, was 30 % slower with them
  if (variant == 0 && w == 16 && a.U >= 512) variant = 5;
  else if (variant == 0 && w == 16 && a.B * a.RCU >= (512ull << 20)) variant = 4;
  switch (w) {
    case 16: return launch_cycle_w<W16>(a, variant, stream);
    case 8: return launch_cycle_w<uint64_t>(a, variant, stream);
    case 4: return launch_cycle_w<uint32_t>(a, variant, stream);
    case 2: return launch_cycle_w<uint16_t>(a, variant, stream);
    default: return launch_cycle_w<uint8_t>(a, variant, stream);
  }
}

// ---------------------------------------------------------------------------
// fused block permutation (OP_BLOCK): K contiguous blocks of n cells per
// workgroup are loaded into LDS with 16-byte loads, then every output cell
// o takes LDS cell T[o] of its block (T: uint16 indices when n <= 65536).
// Small cells (one word each) are packed V to a 16-byte store.
// ---------------------------------------------------------------------------
struct BlkArgs {
  char* data;
  uint64_t base, F, B;  // F: bytes per block
  const void* table;
  uint32_t n, wpc, K, vec16, idx16, chunks;  // chunks: n / V
  uint32_t pad, Fl;                           // pad: log2 of the pad interval (0: none); Fl: LDS bytes per block
  FD32 fnw, fwpc, fch, fF16, fFw;             // n * wpc, wpc, chunks, F / 16, F / w
  uint32_t ndig, stride[8];                   // digit form of the table (ndig = 0: table)
  FD32 frad[8];
  uint32_t dense;  // 1: the K blocks lie in LDS back to back, padded as one run
  FD32 fn;         // n
  // pattern stores (one word per cell): the run's 16-byte chunks repeat
  // their source offsets every P chunks (P * 16 = lcm(F, 16) bytes, a
  // multiple of the pad interval); 0: off
  uint32_t P;
};

// source cell of output cell o: digit form (at most ND digits), or the
// gather table
template <int ND>
__device__ __forceinline__ uint32_t blk_src(const BlkArgs& a, uint32_t o) {
  if (a.ndig) {
    uint32_t q = o, v = 0;
#pragma unroll
    for (int k = 0; k < ND; ++k) {
      if (k >= (int)a.ndig) break;
      const uint32_t nq = a.frad[k].div(q);
      v += (q - nq * a.frad[k].d) * a.stride[k];
      q = nq;
    }
    return v;
  }
  return a.idx16 ? (uint32_t) reinterpret_cast<const uint16_t*>(a.table)[o]
                 : reinterpret_cast<const uint32_t*>(a.table)[o];
}

// LDS byte offset of byte x of a block (4 pad bytes after every 2^pad bytes)
__device__ __forceinline__ uint32_t lds_off(uint32_t x, uint32_t pad) {
  return pad ? x + ((x >> pad) << 2) : x;
}

// LDS byte offset of byte x of block k of the workgroup when blocks are
// padded one by one at Fl bytes apart (not dense)
__device__ __forceinline__ uint32_t blk_lds(const BlkArgs& a, uint32_t k, uint32_t x) {
  return k * a.Fl + lds_off(x, a.pad);
}

// Dense layout (blocks whose size is not a multiple of 16 bytes, or a
// tensor address that is not 16-byte aligned): the workgroup's nf blocks are
// one run [g0, g1) (offsets from data) kept in LDS at its address modulo 16
// and padded as one byte range: LDS byte lds_off(y) holds byte a0 + y, a0 =
// g0 rounded down to an aligned 16-byte chunk, sh = g0 - a0.  Chunks wholly
// inside the pass's bytes load with one 16-byte load, those at its edge word
// by word.  Output chunks wholly inside the run are gathered V words at a
// time and stored with one 16-byte store (with pattern offsets when P
// threads divide the workgroup, else through the digit odometer); the run's
// two edge chunks are stored word by word.
template <typename Word, int V, int ND>
__device__ __forceinline__ void block_run(const BlkArgs& a, const uint64_t f0, const uint32_t nf,
                                          unsigned char* lds) {
  constexpr uint32_t w = sizeof(Word);
  const uint32_t tid = threadIdx.x, nt = blockDim.x;
  const uintptr_t d0 = (uintptr_t)a.data;
  const int64_t g0 = (int64_t)(a.base + f0 * a.F), g1 = g0 + (int64_t)nf * (int64_t)a.F;
  const int64_t a0 = g0 - (int64_t)((d0 + (uint64_t)g0) & 15);
  const uint32_t sh = (uint32_t)(g0 - a0);
  const int64_t lo = (int64_t)a.base, hi = (int64_t)(a.base + a.B * a.F);
  const uint32_t n16 = (uint32_t)((g1 - a0 + 15) / 16);
  for (uint32_t t0 = tid; t0 < n16; t0 += nt * UNR) {
    W16 v[UNR];
    bool full[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const uint32_t t = t0 + u * nt;
      full[u] = false;
      if (t >= n16) continue;
      const int64_t cs = a0 + 16 * (int64_t)t;
      if (cs >= lo && cs + 16 <= hi) {
        v[u] = *reinterpret_cast<const W16*>(a.data + cs);
        full[u] = true;
      } else {
        const int64_t b0 = cs < g0 ? g0 : cs, b1 = cs + 16 < g1 ? cs + 16 : g1;
        for (int64_t b = b0; b < b1; b += w)
          *reinterpret_cast<Word*>(lds + lds_off((uint32_t)(b - a0), a.pad)) =
              *reinterpret_cast<const Word*>(a.data + b);
      }
    }
    // a 16-byte chunk never straddles a pad (intervals of at least 64 bytes)
#pragma unroll
    for (int u = 0; u < UNR; ++u)
      if (full[u]) lds_put16(lds + lds_off((t0 + u * nt) * 16, a.pad), v[u]);
  }
  __syncthreads();
  const uint32_t per = a.n * a.wpc;  // words per block
  // the word of run byte b (b in [g0, g1), a multiple of w)
  auto word_at = [&](int64_t b) -> Word {
    const uint32_t wi = (uint32_t)((b - g0) / w);
    const uint32_t k = a.fnw.div(wi);
    const uint32_t r = wi - k * per;
    const uint32_t o = a.wpc == 1 ? r : a.fwpc.div(r);
    const uint32_t x = r - o * a.wpc;
    const uint32_t src = blk_src<ND>(a, o);
    return *reinterpret_cast<const Word*>(
        lds + lds_off(sh + k * (uint32_t)a.F + (src * a.wpc + x) * w, a.pad));
  };
  auto edge = [&](int64_t cs) {
    const int64_t b0 = cs < g0 ? g0 : cs, b1 = cs + 16 < g1 ? cs + 16 : g1;
    for (int64_t b = b0; b < b1; b += w) *reinterpret_cast<Word*>(a.data + b) = word_at(b);
  };
  if (V > 1 && a.P && nt % a.P == 0) {
    // one word per cell: run byte 16t - sh of chunk t lies at rel within
    // its period of P16 bytes, the same rel for every chunk of the thread
    const uint32_t P16 = a.P * 16;
    const uint32_t pos = tid % a.P;
    const uint32_t rel = (16 * pos + P16 - sh) % P16;
    uint32_t off[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const uint32_t c = rel / w + (uint32_t)v;  // cell from the period's start
      const uint32_t kb = a.fn.div(c);
      off[v] = lds_off(sh + kb * (uint32_t)a.F + blk_src<ND>(a, c - kb * a.n) * w, a.pad);
    }
    // period of chunk t (-1 for the run's first chunk when it is partial)
    int32_t pi = (int32_t)(tid / a.P) - ((pos == 0 && sh > 0) ? 1 : 0);
    const int32_t q = (int32_t)(nt / a.P);
    for (uint32_t t = tid; t < n16; t += nt, pi += q) {
      const int64_t cs = a0 + 16 * (int64_t)t;
      if (cs >= g0 && cs + 16 <= g1) {
        const uint32_t base = lds_off((uint32_t)pi * P16, a.pad);
        Packed<Word, V> tmp;
#pragma unroll
        for (int v = 0; v < V; ++v) tmp.e[v] = *reinterpret_cast<const Word*>(lds + base + off[v]);
        *reinterpret_cast<typename Packed<Word, V>::P*>(a.data + cs) = tmp.p;
      } else {
        edge(cs);
      }
    }
    return;
  }
  for (uint32_t t = tid; t < n16; t += nt) {
    const int64_t cs = a0 + 16 * (int64_t)t;
    if (V == 1 || cs < g0 || cs + 16 > g1) {
      edge(cs);
      continue;
    }
    // V words from word w0 of the run: the first one's block, cell and
    // source from divisions, each next one's by one step of the digit
    // odometer (or the next table entry)
    const uint32_t w0 = (uint32_t)((cs - g0) / w);
    uint32_t k = a.fnw.div(w0);
    const uint32_t r = w0 - k * per;
    uint32_t o = a.wpc == 1 ? r : a.fwpc.div(r);
    uint32_t x = r - o * a.wpc;
    DigitOdometerN<ND> od;
    if (a.ndig) od.init(a.frad, a.stride, a.ndig, o);
    else od.src = blk_src<ND>(a, o);
    Packed<Word, V> tmp;
    if (a.wpc == 1 && a.ndig && a.frad[0].d >= (uint32_t)V) {
      // one-word cells whose fastest output digit spans at least V cells:
      // the chunk's V words lie in at most two runs of that digit, each an
      // arithmetic sequence in LDS; the second run starts one odometer step
      // past the first one's last cell (the next block's first cell when the
      // block ends there)
      const uint32_t S = a.stride[0] * w;
      const uint32_t k1 = a.frad[0].d - od.dg[0];  // cells left in the first run
      const uint32_t base0 = sh + k * (uint32_t)a.F + od.src * w;
      uint32_t base1 = base0;
      if (k1 < (uint32_t)V) {
        if (o + k1 == a.n) {
          base1 = sh + (k + 1) * (uint32_t)a.F;
        } else {
          od.src += (k1 - 1) * a.stride[0];
          od.dg[0] = a.frad[0].d - 1;
          od.step(a.frad, a.stride, a.ndig);
          base1 = sh + k * (uint32_t)a.F + od.src * w;
        }
      }
      const uint32_t b1 = base1 - k1 * S;  // word v of the second run: b1 + v*S
#pragma unroll
      for (int v = 0; v < V; ++v)
        tmp.e[v] = *reinterpret_cast<const Word*>(
            lds + lds_off(((uint32_t)v < k1 ? base0 : b1) + (uint32_t)v * S, a.pad));
    } else {
#pragma unroll
      for (int v = 0; v < V; ++v) {
        tmp.e[v] = *reinterpret_cast<const Word*>(
            lds + lds_off(sh + k * (uint32_t)a.F + (od.src * a.wpc + x) * w, a.pad));
        if (++x == a.wpc) {
          x = 0;
          if (++o == a.n) {  // the next block's first cell
            o = 0;
            ++k;
            if (a.ndig) od.init(a.frad, a.stride, a.ndig, 0);
            else od.src = blk_src<ND>(a, 0);
          } else if (a.ndig) {
            od.step(a.frad, a.stride, a.ndig);
          } else {
            od.src = blk_src<ND>(a, o);
          }
        }
      }
    }
    *reinterpret_cast<typename Packed<Word, V>::P*>(a.data + cs) = tmp.p;
  }
}

// f0: first block of the workgroup's K blocks
template <typename Word, int V, int ND>
__device__ __forceinline__ void block_body(const BlkArgs& a, const uint64_t f0, unsigned char* lds) {
  constexpr uint32_t w = sizeof(Word);
  const uint32_t nf = (uint32_t)((a.B - f0) < a.K ? (a.B - f0) : a.K);
  const uint32_t tid = threadIdx.x, nt = blockDim.x;
  if (a.dense) {
    block_run<Word, V, ND>(a, f0, nf, lds);
    return;
  }
  // blocks of whole 16-byte chunks at an aligned address: the workgroup's nf
  // blocks are one run of 16-byte chunks, each block padded on its own
  const char* g = a.data + a.base + f0 * a.F;
  const uint32_t c16 = (uint32_t)(a.F / 16);
  if (a.vec16 && ((uint64_t)nf * a.F) % 16 == 0 && ((uintptr_t)g & 15) == 0) {
    const uint32_t n16 = (uint32_t)((uint64_t)nf * a.F / 16);
    for (uint32_t t0 = tid; t0 < n16; t0 += nt * UNR) {
      W16 v[UNR];
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        const uint32_t t = t0 + u * nt;
        if (t < n16) v[u] = reinterpret_cast<const W16*>(g)[t];
      }
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        const uint32_t t = t0 + u * nt;
        if (t < n16) {
          const uint32_t k = a.fF16.div(t);
          lds_put16(lds + k * a.Fl + lds_off((t - k * c16) * 16, a.pad), v[u]);
        }
      }
    }
  } else {
    const uint32_t fw = (uint32_t)(a.F / w);
    const uint32_t nw = nf * fw;
    for (uint32_t t = tid; t < nw; t += nt) {
      const uint32_t k = a.fFw.div(t);
      *reinterpret_cast<Word*>(lds + blk_lds(a, k, (t - k * fw) * w)) =
          reinterpret_cast<const Word*>(g)[t];
    }
  }
  __syncthreads();
  char* out = a.data + a.base + f0 * a.F;
  if (V > 1 && a.P && nt % a.P == 0 && ((uint64_t)nf * a.F) % 16 == 0 && ((uintptr_t)out & 15) == 0) {
    // one word per cell and a whole number of periods per P threads: thread
    // tid always takes chunk tid % P of a period, so the LDS offsets of its
    // V words, relative to the period, are computed once
    const uint32_t total = (uint32_t)((uint64_t)nf * a.F / 16);
    const uint32_t pos = tid % a.P, q = nt / a.P;
    uint32_t off[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const uint32_t c = pos * V + (uint32_t)v;  // cell of the period
      const uint32_t kb = a.fn.div(c);
      const uint32_t src = blk_src<ND>(a, c - kb * a.n);
      off[v] = kb * a.Fl + lds_off(src * w, a.pad);
    }
    uint32_t pi = tid / a.P;  // period of chunk t
    for (uint32_t t = tid; t < total; t += nt, pi += q) {
      const uint32_t base = pi * a.Fl;
      Packed<Word, V> tmp;
#pragma unroll
      for (int v = 0; v < V; ++v) tmp.e[v] = *reinterpret_cast<const Word*>(lds + base + off[v]);
      *reinterpret_cast<typename Packed<Word, V>::P*>(out + (uint64_t)t * 16) = tmp.p;
    }
  } else if (V > 1 && ((uint64_t)nf * a.F) % 16 == 0 && ((uintptr_t)out & 15) == 0) {
    // 16-byte stores: chunk t of the run is words t*V .. t*V+V-1 (cells of
    // one or several words, possibly of two blocks).  The first word's
    // block, cell and source come from divisions; each following word's from
    // the previous one's: the next word of the cell, or the next cell, whose
    // source is one step of the mixed-radix odometer over the cell's output
    // digits (or the next table entry).
    const uint32_t per = a.n * a.wpc;  // words per block
    const uint32_t total = (uint32_t)((uint64_t)nf * a.F / 16);
    for (uint32_t t = tid; t < total; t += nt) {
      const uint32_t w0 = t * V;
      uint32_t k = a.fnw.div(w0);
      const uint32_t r = w0 - k * per;
      uint32_t o = a.wpc == 1 ? r : a.fwpc.div(r);
      uint32_t x = r - o * a.wpc;
      DigitOdometerN<ND> od;
      if (a.ndig) od.init(a.frad, a.stride, a.ndig, o);
      else od.src = blk_src<ND>(a, o);
      Packed<Word, V> tmp;
      if (a.wpc == 1 && a.ndig && a.frad[0].d >= (uint32_t)V) {
        // one-word cells whose fastest output digit spans at least V cells:
        // at most two runs of that digit (see block_run)
        const uint32_t S = a.stride[0] * w;
        const uint32_t k1 = a.frad[0].d - od.dg[0];
        const uint32_t y0 = od.src * w, blk0 = k * a.Fl;
        uint32_t y1 = y0, blk1 = blk0;
        if (k1 < (uint32_t)V) {
          if (o + k1 == a.n) {
            y1 = 0;
            blk1 = (k + 1) * a.Fl;
          } else {
            od.src += (k1 - 1) * a.stride[0];
            od.dg[0] = a.frad[0].d - 1;
            od.step(a.frad, a.stride, a.ndig);
            y1 = od.src * w;
          }
        }
        const uint32_t z1 = y1 - k1 * S;
#pragma unroll
        for (int v = 0; v < V; ++v) {
          const bool first = (uint32_t)v < k1;
          tmp.e[v] = *reinterpret_cast<const Word*>(
              lds + (first ? blk0 : blk1) + lds_off((first ? y0 : z1) + (uint32_t)v * S, a.pad));
        }
        *reinterpret_cast<typename Packed<Word, V>::P*>(out + (uint64_t)t * 16) = tmp.p;
        continue;
      }
#pragma unroll
      for (int v = 0; v < V; ++v) {
        tmp.e[v] = *reinterpret_cast<const Word*>(lds + blk_lds(a, k, (od.src * a.wpc + x) * w));
        if (++x == a.wpc) {
          x = 0;
          if (++o == a.n) {  // the next block's first cell
            o = 0;
            ++k;
            if (a.ndig) od.init(a.frad, a.stride, a.ndig, 0);
            else od.src = blk_src<ND>(a, 0);
          } else if (a.ndig) {
            od.step(a.frad, a.stride, a.ndig);
          } else {
            od.src = blk_src<ND>(a, o);
          }
        }
      }
      *reinterpret_cast<typename Packed<Word, V>::P*>(out + (uint64_t)t * 16) = tmp.p;
    }
  } else {
    const uint32_t per = a.n * a.wpc;
    const uint32_t total = nf * per;
    for (uint32_t t = tid; t < total; t += nt) {
      const uint32_t k = a.fnw.div(t);
      const uint32_t r = t - k * per;
      const uint32_t o = a.wpc == 1 ? r : a.fwpc.div(r);
      const uint32_t x = r - o * a.wpc;
      const uint32_t src = blk_src<ND>(a, o);
      reinterpret_cast<Word*>(out)[(uint64_t)k * per + r] =
          *reinterpret_cast<const Word*>(lds + blk_lds(a, k, (src * a.wpc + x) * w));
    }
  }
}

template <typename Word, int V, int ND>
__global__ __launch_bounds__(512) void k_block(BlkArgs a, int remap) {
  extern __shared__ __align__(16) unsigned char lds[];
  block_body<Word, V, ND>(a, (uint64_t)xcd_logical(blockIdx.x, gridDim.x, remap) * a.K, lds);
}

const void* block_table_device(const BlockTable& t, hipError_t* err) {
  *err = hipSuccess;
  int dev = 0;
  if ((*err = hipGetDevice(&dev)) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(t.dev.mu);
  auto it = t.dev.by_dev.find(dev);
  if (it != t.dev.by_dev.end()) return it->second.get();
  const bool i16 = t.cells <= 65536;
  std::vector<uint16_t> h16;
  if (i16) h16.assign(t.src.begin(), t.src.end());
  const size_t nb = t.src.size() * (i16 ? 2 : 4);
  void* p = nullptr;
  if ((*err = hipMalloc(&p, std::max<size_t>(nb, 16))) != hipSuccess) return nullptr;
  std::shared_ptr<void> own = device_owned(p, dev, (int64_t)std::max<size_t>(nb, 16));
  if ((*err = hipMemcpy(p, i16 ? (const void*)h16.data() : (const void*)t.src.data(), nb,
                        hipMemcpyHostToDevice)) != hipSuccess)
    return nullptr;
  t.dev.by_dev[dev] = own;
  return p;
}

template <typename Word>
hipError_t launch_block_w(BlkArgs a, int vec, int64_t lds, int64_t groups, const LaunchCfg& cfg,
                          hipStream_t stream) {
  constexpr int V = 16 / sizeof(Word) > 1 ? 16 / sizeof(Word) : 1;
  int threads = lds > 96 * 1024 ? cfg.threads_big : cfg.threads;
  if (cfg.lds_threads_auto) {
    const int64_t words = (lds + 15) / 16;
    threads = (int)std::min<int64_t>(512, std::max<int64_t>(64, ((words + UNR - 1) / UNR + 63) / 64 * 64));
  }
  // pattern stores want a whole number of periods per workgroup's threads
  if (vec && V > 1 && a.P) threads = std::max<int>((int)a.P, threads / (int)a.P * (int)a.P);
  // digit tables of at most 4 digits (nearly all) run the 4-digit kernels
  const bool d4 = a.ndig <= 4;
  if (vec && V > 1 && d4)
    hipLaunchKernelGGL((k_block<Word, V, 4>), dim3((uint32_t)groups), dim3(threads), (size_t)lds, stream,
                       a, cfg.xcd_remap);
  else if (vec && V > 1)
    hipLaunchKernelGGL((k_block<Word, V, 8>), dim3((uint32_t)groups), dim3(threads), (size_t)lds, stream,
                       a, cfg.xcd_remap);
  else if (d4)
    hipLaunchKernelGGL((k_block<Word, 1, 4>), dim3((uint32_t)groups), dim3(threads), (size_t)lds, stream,
                       a, cfg.xcd_remap);
  else
    hipLaunchKernelGGL((k_block<Word, 1, 8>), dim3((uint32_t)groups), dim3(threads), (size_t)lds, stream,
                       a, cfg.xcd_remap);
  return hipGetLastError();
}

// Arguments of a fused block pass; *w: word size, *vec: packed stores,
// *lds: LDS bytes per workgroup, *groups: workgroups of K blocks.
hipError_t prepare_block_args(const Op& op, char* data, const LaunchCfg& cfg, BlkArgs& a,
                              int64_t* w_out, int* vec_out, int64_t* lds_out, int64_t* groups_out,
                              const void* table = nullptr) {
  if (!op.blk || op.blk->cells != op.R || op.blk->cell_bytes != op.U) return hipErrorInvalidValue;
  const BlockTable& t = *op.blk;
  const int64_t F = op.R * op.U;
  if (F > 160 * 1024 || op.R >= (1ll << 31) || op.B <= 0) return hipErrorInvalidValue;
  hipError_t e = hipSuccess;
  const void* d = nullptr;
  if (!t.ndig && !(d = table ? table : block_table_device(t, &e))) return e;
  const uintptr_t base = (uintptr_t)(data + op.base);
  int64_t w = 16;
  while (w > 1 && ((op.U % w) || (base % w) || (F % w))) w /= 2;
  a.data = data;
  a.base = op.base;
  a.F = (uint64_t)F;
  a.B = (uint64_t)op.B;
  a.table = d;
  a.n = (uint32_t)op.R;
  a.wpc = (uint32_t)(op.U / w);
  a.vec16 = base % 16 == 0 ? 1 : 0;
  a.idx16 = t.cells <= 65536 ? 1 : 0;
  const int64_t V = w < 16 ? 16 / w : 1;
  const int vec = V > 1;
  a.chunks = (uint32_t)op.R;
  // blocks whose size is not a multiple of 16 bytes, or that do not start
  // 16-byte aligned, lie in LDS back to back as one run kept at its address
  // modulo 16 (dense), padded as one byte range; otherwise each block is
  // padded on its own.  Padding keeps cells of up to 16 bytes whole (pad
  // intervals are >= 64 bytes and cells divide 64 bytes) and LDS words
  // aligned (only 1-, 2- and 4-byte words are padded)
  a.dense = (F % 16 || base % 16) ? 1 : 0;
  a.pad = (t.pad_log2 && 64 % op.U == 0 && w <= 4) ? (uint32_t)t.pad_log2 : 0;
  const int64_t Fl = a.pad ? F + (F >> a.pad) * 4 + 4 : F;
  a.Fl = (uint32_t)(a.dense ? F : (Fl + 15) / 16 * 16);
  // LDS bytes of K blocks (dense: the run, the partial chunks at its two
  // ends and the pads)
  auto lds_of = [&](int64_t K) -> int64_t {
    if (!a.dense) return K * (int64_t)a.Fl;
    const int64_t run = K * F + 32;
    return (run + (a.pad ? (run >> a.pad) * 4 + 4 : 0) + 15) / 16 * 16;
  };
  if (lds_of(1) > 160 * 1024) return hipErrorInvalidValue;
  {
    const int64_t Pc = F / gcd64(F, 16);  // 16-byte chunks per period of lcm(F, 16) bytes
    a.P = (a.wpc == 1 && V > 1 && Pc <= 512 &&
           (!a.dense || !a.pad || (Pc * 16) % (1ll << a.pad) == 0))
              ? (uint32_t)Pc
              : 0;
  }
  int64_t K = pick_K(op.B, a.dense ? Fl : (int64_t)a.Fl, cfg);
  while (K > 1 && lds_of(K) > 160 * 1024) --K;
  // a multiple of the blocks that make 16-byte runs and whole 16-byte stores
  // (past the LDS target when that many blocks fit the workgroup's LDS)
  int64_t mK = a.dense ? 1 : 16 / gcd64(F, 16);
  if (K < mK && mK <= op.B && lds_of(mK) <= 160 * 1024) K = mK;
  if (K >= mK) K -= K % mK;
  a.K = (uint32_t)K;
  a.fnw.init(a.n * a.wpc);
  a.fwpc.init(a.wpc);
  a.fch.init(a.chunks);
  a.fF16.init((uint32_t)std::max<int64_t>(F / 16, 1));
  a.fFw.init((uint32_t)(F / w));
  a.fn.init(a.n);
  a.ndig = (uint32_t)t.ndig;
  for (int k = 0; k < 8; ++k) {
    a.frad[k].init(k < t.ndig ? t.radix[k] : 1);
    a.stride[k] = k < t.ndig ? t.stride[k] : 0;
  }
  const int64_t groups = (op.B + K - 1) / K;
  if (groups >= (1ll << 31)) return hipErrorInvalidValue;
  *w_out = w;
  *vec_out = vec;
  *lds_out = lds_of(K);
  *groups_out = groups;
  return hipSuccess;
}

hipError_t launch_block(const Op& op, char* data, const LaunchCfg& cfg, hipStream_t stream,
                        const void* table) {
  BlkArgs a;
  int64_t w = 1, lds = 0, groups = 0;
  int vec = 0;
  hipError_t e = prepare_block_args(op, data, cfg, a, &w, &vec, &lds, &groups, table);
  if (e != hipSuccess) return e;
  switch (w) {
    case 16: return launch_block_w<W16>(a, vec, lds, groups, cfg, stream);
    case 8: return launch_block_w<uint64_t>(a, vec, lds, groups, cfg, stream);
    case 4: return launch_block_w<uint32_t>(a, vec, lds, groups, cfg, stream);
    case 2: return launch_block_w<uint16_t>(a, vec, lds, groups, cfg, stream);
    default: return launch_block_w<uint8_t>(a, vec, lds, groups, cfg, stream);
  }
}

// ---------------------------------------------------------------------------
// segmented tile pass (OP_SEGTILE): K matrices per workgroup.  A matrix of
// R x C units of E bytes (C fastest) lies in nseg segments of segb bytes at
// addresses given by mixed-radix digits.  The workgroup writes the K*nseg
// segment addresses to LDS, loads its matrices row by row into LDS (rows
// padded to rs bytes), and stores the transposed matrices, in row-major
// order, over the same segments.  Words of w bytes never straddle a
// segment, an LDS row or (when V > 1) an output row: w divides segb, C*E
// and R*E.  V > 1: one word packs V units (E = w / V) of an output row;
// V = 1: a unit is E / w words.
// ---------------------------------------------------------------------------
struct SegArgs {
  char* data;
  uint32_t B, R, C, E;   // matrices; rows, columns and unit bytes of each
  uint32_t nseg;         // segments per matrix
  uint32_t K;            // matrices per workgroup
  uint32_t rs, fpl;      // LDS bytes per padded row, per matrix
  uint32_t tbl;          // LDS byte offset of the segment addresses
  uint32_t wpm;          // words per matrix
  uint32_t nin, nout;    // digits of the segment index and of the matrix index
  uint64_t sin[6], sout[6];
  FD32 fin[6], fout[6];  // their radices
  FD32 fnseg, fwpm, fwps, fwpr, fwpo, fwpe;  // nseg; words per matrix, segment,
                                             // LDS row, output row, unit
};

template <int BYTES>
struct UIntOf;
template <>
struct UIntOf<1> {
  using T = uint8_t;
};
template <>
struct UIntOf<2> {
  using T = uint16_t;
};
template <>
struct UIntOf<4> {
  using T = uint32_t;
};
template <>
struct UIntOf<8> {
  using T = uint64_t;
};

// LDS word store / load at an offset aligned to min(4, word) bytes
template <typename Word>
__device__ __forceinline__ void lds_putw(unsigned char* p, const Word& v) {
  if constexpr (sizeof(Word) == 16) {
    lds_put16(p, v);
  } else if constexpr (sizeof(Word) == 8) {
    if (((uintptr_t)p & 7) == 0) {
      *reinterpret_cast<Word*>(p) = v;
    } else {
      reinterpret_cast<uint32_t*>(p)[0] = (uint32_t)v;
      reinterpret_cast<uint32_t*>(p)[1] = (uint32_t)(v >> 32);
    }
  } else {
    *reinterpret_cast<Word*>(p) = v;
  }
}
template <typename Word>
__device__ __forceinline__ Word lds_getw(const unsigned char* p) {
  if (sizeof(Word) < 8 || ((uintptr_t)p & (sizeof(Word) - 1)) == 0)
    return *reinterpret_cast<const Word*>(p);
  const uint32_t* q = reinterpret_cast<const uint32_t*>(p);
  if constexpr (sizeof(Word) == 16) {
    return W16{q[0], q[1], q[2], q[3]};
  } else if constexpr (sizeof(Word) == 8) {
    return (Word)q[0] | ((Word)q[1] << 32);
  } else {
    return *reinterpret_cast<const Word*>(p);
  }
}

__device__ __forceinline__ uint64_t seg_addr(const SegArgs& a, uint32_t s, uint32_t b) {
  uint64_t addr = 0;
#pragma unroll
  for (int d = 0; d < 6; ++d) {
    if (d < (int)a.nin) {
      const uint32_t q = a.fin[d].div(s);
      addr += (uint64_t)(s - q * a.fin[d].d) * a.sin[d];
      s = q;
    }
  }
#pragma unroll
  for (int d = 0; d < 6; ++d) {
    if (d < (int)a.nout) {
      const uint32_t q = a.fout[d].div(b);
      addr += (uint64_t)(b - q * a.fout[d].d) * a.sout[d];
      b = q;
    }
  }
  return addr;
}

// f0: first matrix of the workgroup's K; NT: non-temporal global loads (bit
// 0) and stores (bit 1)
template <typename Word, int V, int NT = 0>
__device__ __forceinline__ void seg_body(const SegArgs& a, const uint32_t f0, unsigned char* lds) {
  constexpr uint32_t w = sizeof(Word);
  const uint32_t nf = (a.B - f0) < a.K ? (a.B - f0) : a.K;
  const uint32_t tid = threadIdx.x, nt = blockDim.x;
  uint64_t* tbl = reinterpret_cast<uint64_t*>(lds + a.tbl);
  for (uint32_t e = tid; e < nf * a.nseg; e += nt) {
    const uint32_t k = a.fnseg.div(e);
    tbl[e] = seg_addr(a, e - k * a.nseg, f0 + k);
  }
  __syncthreads();
  const uint32_t n = nf * a.wpm;
  for (uint32_t t0 = tid; t0 < n; t0 += nt * UNR) {
    Word v[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const uint32_t t = t0 + u * nt;
      if (t < n) {
        const uint32_t k = a.fwpm.div(t), q = t - k * a.wpm;  // word q of matrix k
        const uint32_t s = a.fwps.div(q);
        v[u] = ld_nt<NT, Word>(a.data + tbl[k * a.nseg + s] + (uint64_t)(q - s * a.fwps.d) * w);
      }
    }
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const uint32_t t = t0 + u * nt;
      if (t < n) {
        const uint32_t k = a.fwpm.div(t), q = t - k * a.wpm;
        const uint32_t r = a.fwpr.div(q);
        lds_putw<Word>(lds + k * a.fpl + r * a.rs + (q - r * a.fwpr.d) * w, v[u]);
      }
    }
  }
  __syncthreads();
  for (uint32_t t = tid; t < n; t += nt) {
    const uint32_t k = a.fwpm.div(t), q = t - k * a.wpm;  // output word q of matrix k
    const uint32_t j = a.fwpo.div(q), x = q - j * a.fwpo.d;  // output row j (column j in LDS)
    const unsigned char* l = lds + k * a.fpl + j * a.E;
    Word val;
    if constexpr (V > 1) {
      using Elem = typename UIntOf<sizeof(Word) / V>::T;
      Packed<Elem, V> pk;
#pragma unroll
      for (int v = 0; v < V; ++v) pk.e[v] = *reinterpret_cast<const Elem*>(l + (x * V + v) * a.rs);
      val = pk.p;
    } else {
      const uint32_t i = a.fwpe.div(x);
      val = lds_getw<Word>(l + i * a.rs + (x - i * a.fwpe.d) * w);
    }
    const uint32_t s = a.fwps.div(q);
    st_nt<NT, Word>(a.data + tbl[k * a.nseg + s] + (uint64_t)(q - s * a.fwps.d) * w, val);
  }
}

template <typename Word, int V, int NT = 0>
__global__ __launch_bounds__(512) void k_segtile(SegArgs a, int remap) {
  extern __shared__ __align__(16) unsigned char lds[];
  seg_body<Word, V, NT>(a, xcd_logical(blockIdx.x, gridDim.x, remap) * a.K, lds);
}

// Workgroup size of a segmented tile pass: enough threads that each issues
// one full round of UNR loads for the workgroup's words (whole waves, at
// most 512), unless the segtile_threads knob fixes it.
int segtile_threads(const SegArgs& a, const LaunchCfg& cfg) {
  if (cfg.segtile_threads > 0) return cfg.segtile_threads;
  const uint64_t words = (uint64_t)a.K * a.wpm;
  const uint64_t t = (words + UNR - 1) / UNR;
  return (int)std::min<uint64_t>(512, std::max<uint64_t>(64, (t + 63) / 64 * 64));
}

// Word size of a segmented tile pass: the widest (at most 16 bytes) that
// divides the segments, both row lengths and the data pointer, and that
// divides the unit or is a multiple of it.
int64_t segtile_word_bytes(const Op& op, uintptr_t data) {
  int64_t w = 16;
  while (w > 1 && ((op.segb % w) || ((op.C * op.U) % w) || ((op.R * op.U) % w) || (data % w)))
    w /= 2;
  while (w > 1 && (op.U % w) && (w % op.U)) w /= 2;
  return w;
}

// Arguments of a segmented tile pass; *w: word bytes, *V: units per word,
// *lds: LDS bytes per workgroup, *groups: workgroups of K matrices.
hipError_t prepare_seg_args(const Op& op, char* data, const LaunchCfg& cfg, SegArgs& a,
                            int64_t* w_out, int64_t* V_out, int64_t* lds_out,
                            int64_t* groups_out) {
  const int64_t bytes = op.R * op.C * op.U;
  if (op.kind != OP_SEGTILE || op.segb <= 0 || op.B <= 0 || op.B >= (1ll << 32) ||
      bytes % op.segb || op.sin_n > 6 || op.sout_n > 6 || op.base != 0)
    return hipErrorInvalidValue;
  // the digits must number exactly nseg segments and B matrices, and the
  // farthest segment must end inside the B*R*C*U bytes of the tensor
  {
    unsigned __int128 pin = 1, pout = 1, last = (unsigned __int128)op.segb;
    for (int d = 0; d < op.sin_n; ++d) {
      if (op.sin_radix[d] < 1 || op.sin_stride[d] < 0) return hipErrorInvalidValue;
      pin *= (uint64_t)op.sin_radix[d];
      last += (unsigned __int128)(op.sin_radix[d] - 1) * (uint64_t)op.sin_stride[d];
    }
    for (int d = 0; d < op.sout_n; ++d) {
      if (op.sout_radix[d] < 1 || op.sout_stride[d] < 0) return hipErrorInvalidValue;
      pout *= (uint64_t)op.sout_radix[d];
      last += (unsigned __int128)(op.sout_radix[d] - 1) * (uint64_t)op.sout_stride[d];
    }
    if (pin != (unsigned __int128)(bytes / op.segb) || pout != (unsigned __int128)op.B ||
        last > (unsigned __int128)op.B * (uint64_t)bytes)
      return hipErrorInvalidValue;
  }
  const int64_t w = segtile_word_bytes(op, (uintptr_t)data);
  const int64_t V = w > op.U ? w / op.U : 1;
  int64_t uw = 1;  // largest power of two (at most 16) dividing the unit
  while (uw < 16 && op.U % (uw * 2) == 0) uw *= 2;
  const int64_t rs = tile_row_stride(op.C * op.U, uw);
  const int64_t fpl = (op.R * rs + 15) / 16 * 16;
  const int64_t nseg = bytes / op.segb;
  if (fpl + nseg * 8 > 160 * 1024) return hipErrorInvalidValue;
  int64_t K = pick_K(op.B, fpl + nseg * 8, cfg);
  while (K > 1 && K * (fpl + nseg * 8) > 160 * 1024) --K;
  std::memset(&a, 0, sizeof(a));
  a.data = data;
  a.B = (uint32_t)op.B;
  a.R = (uint32_t)op.R;
  a.C = (uint32_t)op.C;
  a.E = (uint32_t)op.U;
  a.nseg = (uint32_t)nseg;
  a.K = (uint32_t)K;
  a.rs = (uint32_t)rs;
  a.fpl = (uint32_t)fpl;
  a.tbl = (uint32_t)(K * fpl);
  a.wpm = (uint32_t)(bytes / w);
  a.nin = (uint32_t)op.sin_n;
  a.nout = (uint32_t)op.sout_n;
  for (int d = 0; d < 6; ++d) {
    a.fin[d].init(d < op.sin_n ? op.sin_radix[d] : 1);
    a.sin[d] = d < op.sin_n ? (uint64_t)op.sin_stride[d] : 0;
    a.fout[d].init(d < op.sout_n ? op.sout_radix[d] : 1);
    a.sout[d] = d < op.sout_n ? (uint64_t)op.sout_stride[d] : 0;
  }
  a.fnseg.init(nseg);
  a.fwpm.init(bytes / w);
  a.fwps.init(op.segb / w);
  a.fwpr.init(op.C * op.U / w);
  a.fwpo.init(op.R * op.U / w);
  a.fwpe.init(V > 1 ? 1 : op.U / w);
  const int64_t groups = (op.B + K - 1) / K;
  if (groups >= (1ll << 31)) return hipErrorInvalidValue;
  *w_out = w;
  *V_out = V;
  *lds_out = K * (fpl + nseg * 8);
  *groups_out = groups;
  return hipSuccess;
}

template <typename Word, int V>
hipError_t launch_seg_wv(const SegArgs& a, int64_t lds, int64_t groups, const LaunchCfg& cfg,
                         hipStream_t stream) {
  const int threads = segtile_threads(a, cfg);
  if constexpr (sizeof(Word) == 16) {
    if (cfg.segtile_nt == 3) {
      hipLaunchKernelGGL((k_segtile<Word, V, 3>), dim3((uint32_t)groups), dim3(threads),
                         (size_t)lds, stream, a, cfg.xcd_remap);
      return hipGetLastError();
    }
    if (cfg.segtile_nt == 2) {
      hipLaunchKernelGGL((k_segtile<Word, V, 2>), dim3((uint32_t)groups), dim3(threads),
                         (size_t)lds, stream, a, cfg.xcd_remap);
      return hipGetLastError();
    }
  }
  hipLaunchKernelGGL((k_segtile<Word, V>), dim3((uint32_t)groups), dim3(threads), (size_t)lds,
                     stream, a, cfg.xcd_remap);
  return hipGetLastError();
}

template <typename Word>
hipError_t launch_seg_w(const SegArgs& a, int64_t V, int64_t lds, int64_t groups,
                        const LaunchCfg& cfg, hipStream_t stream) {
  constexpr int64_t w = sizeof(Word);
  switch (V) {
    case 1: return launch_seg_wv<Word, 1>(a, lds, groups, cfg, stream);
    case 2: return launch_seg_wv<Word, (w >= 2 ? 2 : 1)>(a, lds, groups, cfg, stream);
    case 4: return launch_seg_wv<Word, (w >= 4 ? 4 : 1)>(a, lds, groups, cfg, stream);
    case 8: return launch_seg_wv<Word, (w >= 8 ? 8 : 1)>(a, lds, groups, cfg, stream);
    default: return launch_seg_wv<Word, (w >= 16 ? 16 : 1)>(a, lds, groups, cfg, stream);
  }
}

hipError_t launch_segtile(const Op& op, char* data, const LaunchCfg& cfg, hipStream_t stream) {
  SegArgs a;
  int64_t w = 1, V = 1, lds = 0, groups = 0;
  hipError_t e = prepare_seg_args(op, data, cfg, a, &w, &V, &lds, &groups);
  if (e != hipSuccess) return e;
  switch (w) {
    case 16: return launch_seg_w<W16>(a, V, lds, groups, cfg, stream);
    case 8: return launch_seg_w<uint64_t>(a, V, lds, groups, cfg, stream);
    case 4: return launch_seg_w<uint32_t>(a, V, lds, groups, cfg, stream);
    case 2: return launch_seg_w<uint16_t>(a, V, lds, groups, cfg, stream);
    default: return launch_seg_w<uint8_t>(a, V, lds, groups, cfg, stream);
  }
}

// ---------------------------------------------------------------------------
// persistent batched launch: items of many tensors' passes in one kernel
// ---------------------------------------------------------------------------
// An op is the arguments of one tensor's pass; its variant selects the body:
//   KM*10 + log2(word)*2 + vec  LDS tile/column/row passes (KM 0..6)
//   70 + log2(word)             cycle following (item: blockDim threads)
//   80 + log2(word)*2 + vec     fused block permutation (item: K blocks)
//   90 + log2(word)             boundary save of a cycle pass (item: blockDim words)
//   100 + log2(word)*5 + log2(V) segmented tile (item: K matrices)
//   120 + log2(word)            cycle following of cells that are not whole
//                               16-byte words, in shifted 16-byte chunks
//                               (item: blockDim threads; 32-bit offsets)
struct CycUOp {
  CycArgs a;
  CycUArgs u;
};
constexpr size_t cmax(size_t a, size_t b) { return a > b ? a : b; }
constexpr size_t kBatchArgBytes =
    cmax(cmax(cmax(sizeof(Args<uint32_t>), sizeof(CycArgs)), cmax(sizeof(BlkArgs), sizeof(SegArgs))),
         sizeof(CycUOp));
struct BatchOpDev {
  int variant;
  int pad[3];
  alignas(16) unsigned char raw[(kBatchArgBytes + 15) / 16 * 16];
};
struct BatchItemDev {
  uint32_t op;
  uint32_t f0;
};

template <int WL>
struct WordOf {
  using T = typename std::conditional<
      WL == 4, W16,
      typename std::conditional<
          WL == 3, uint64_t,
          typename std::conditional<WL == 2, uint32_t,
                                    typename std::conditional<WL == 1, uint16_t, uint8_t>::type>::
              type>::type>::type;
};

template <int KM, int WL, int VF>
__device__ __forceinline__ void batch_body(const BatchOpDev& op, uint32_t f0, unsigned char* lds) {
  using Word = typename WordOf<WL>::T;
  constexpr int V = VF ? (16 >> WL) : 1;
  const Args<uint32_t>& a = *reinterpret_cast<const Args<uint32_t>*>(op.raw);
  if (KM == 0) tile_body<Word, uint32_t, V>(a, f0, lds);
  else if (KM <= 4) col_body<Word, uint32_t, (KM <= 4 ? KM : 1), V>(a, f0, lds);
  else row_body<Word, uint32_t, (KM >= 5 ? KM : 5), V>(a, f0, lds);
}

#define EH_CASE(KM, WL, VF) \
  case KM * 10 + WL * 2 + VF: batch_body<KM, WL, VF>(op, item.f0, lds); break;
#define EH_CASES_W(KM, WL) EH_CASE(KM, WL, 0) EH_CASE(KM, WL, 1)
#define EH_CASES(KM) \
  EH_CASES_W(KM, 0) EH_CASES_W(KM, 1) EH_CASES_W(KM, 2) EH_CASES_W(KM, 3) EH_CASE(KM, 4, 0)
#define EH_CYC(WL)                                                                           \
  case 70 + WL:                                                                              \
    cycle_body<typename WordOf<WL>::T, 8, 0>(*reinterpret_cast<const CycArgs*>(op.raw),      \
                                             (uint64_t)item.f0 * blockDim.x + threadIdx.x); \
    break;                                                                                   \
  case 90 + WL:                                                                              \
    cycle_save_word<typename WordOf<WL>::T>(*reinterpret_cast<const CycArgs*>(op.raw),       \
                                            (uint64_t)item.f0 * blockDim.x + threadIdx.x);   \
    break;
#define EH_CYCU(WL)                                                                           \
  case 120 + WL: {                                                                            \
    const CycUOp& c = *reinterpret_cast<const CycUOp*>(op.raw);                               \
    cycle_u_body<typename WordOf<WL>::T, int32_t, 4>(c.a, c.u,                                \
                                                    (uint64_t)item.f0 * blockDim.x + threadIdx.x); \
    break;                                                                                    \
  }
#define EH_BLK(WL, VF)                                                                      \
  case 80 + WL * 2 + VF:                                                                    \
    block_body<typename WordOf<WL>::T, VF ? (16 >> WL) : 1, 8>(                                \
        *reinterpret_cast<const BlkArgs*>(op.raw), item.f0, lds);                           \
    break;

#define EH_SEG(WL, VL)                                                                      \
  case 100 + WL * 5 + VL:                                                                   \
    seg_body<typename WordOf<WL>::T, (1 << VL)>(*reinterpret_cast<const SegArgs*>(op.raw),  \
                                                item.f0, lds);                              \
    break;

__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void k_batch(const BatchOpDev* __restrict__ ops,
                                               const BatchItemDev* __restrict__ items,
                                               uint32_t nitems) {
  extern __shared__ __align__(16) unsigned char lds[];
  for (uint32_t it = blockIdx.x; it < nitems; it += gridDim.x) {
    const BatchItemDev item = items[it];
    const BatchOpDev& op = ops[item.op];
    switch (op.variant) {
      EH_CASES(0)
      EH_CASES(1)
      EH_CASES(2)
      EH_CASES(3)
      EH_CASES(4)
      EH_CASES(5)
      EH_CASES(6)
      EH_CYC(0)
      EH_CYC(1)
      EH_CYC(2)
      EH_CYC(3)
      EH_CYC(4)
      EH_BLK(0, 0)
      EH_BLK(0, 1)
      EH_BLK(1, 0)
      EH_BLK(1, 1)
      EH_BLK(2, 0)
      EH_BLK(2, 1)
      EH_BLK(3, 0)
      EH_BLK(3, 1)
      EH_BLK(4, 0)
      EH_SEG(0, 0)
      EH_SEG(1, 0)
      EH_SEG(1, 1)
      EH_SEG(2, 0)
      EH_SEG(2, 1)
      EH_SEG(2, 2)
      EH_SEG(3, 0)
      EH_SEG(3, 1)
      EH_SEG(3, 2)
      EH_SEG(3, 3)
      EH_SEG(4, 0)
      EH_SEG(4, 1)
      EH_SEG(4, 2)
      EH_SEG(4, 3)
      EH_SEG(4, 4)
      default: break;
    }
    __syncthreads();  // the next item reuses the LDS
  }
}

// the items without LDS (cycle following and boundary saves) in a kernel of
// their own: it needs a fraction of k_batch's registers, so more waves of
// these latency-bound items stay resident
__global__ __launch_bounds__(512) void k_batch_cycles(const BatchOpDev* __restrict__ ops,
                                                      const BatchItemDev* __restrict__ items,
                                                      uint32_t nitems) {
  for (uint32_t it = blockIdx.x; it < nitems; it += gridDim.x) {
    const BatchItemDev item = items[it];
    const BatchOpDev& op = ops[item.op];
    switch (op.variant) {
      EH_CYC(0)
      EH_CYC(1)
      EH_CYC(2)
      EH_CYC(3)
      EH_CYC(4)
      EH_CYCU(0)
      EH_CYCU(1)
      EH_CYCU(2)
      default: break;
    }
  }
}
#undef EH_CYCU
#undef EH_BLK
#undef EH_SEG
#undef EH_CYC
#undef EH_CASES
#undef EH_CASES_W
#undef EH_CASE

}  // namespace

hipError_t prepare_op(const Op& op) {
  hipError_t e = hipSuccess;
  if (op.kind == OP_CYCLE && op.cyc && !op.cyc->segs.empty()) cycle_table_device(*op.cyc, &e);
  if (op.kind == OP_BLOCK && op.blk && !op.blk->ndig) block_table_device(*op.blk, &e);
  return e;
}

hipError_t launch_op(const Op& op, char* data, char* scratch, int64_t scratch_bytes,
                     const LaunchCfg& cfg, hipStream_t stream, const void* table) {
  if (op.kind == OP_CYCLE)
    return launch_cycle(op, data, scratch, scratch_bytes, cfg.cycle_variant, cfg.cycle_u_depth, stream,
                        table);
  if (op.kind == OP_BLOCK) return launch_block(op, data, cfg, stream, table);
  if (op.kind == OP_SEGTILE) return launch_segtile(op, data, cfg, stream);
  if (op.kind == 0 || op.us <= 0 || op.U % op.us) return hipErrorInvalidValue;
  if (needs64(op)) return run_op<uint64_t>(op, data, scratch, scratch_bytes, cfg, stream);
  return run_op<uint32_t>(op, data, scratch, scratch_bytes, cfg, stream);
}


// ---------------------------------------------------------------------------
// persistent batched launches (host side)
// ---------------------------------------------------------------------------
struct BatchLaunch {
  // one k_batch launch per part, in order: boundary saves, cycle items (no
  // LDS), then LDS items in classes of footprint size so that a few large
  // footprints do not hold every workgroup of the stage to one per CU
  struct Part {
    uint32_t offset, count, grid;
    int64_t lds;
    bool cycles;  // only cycle and boundary-save items: k_batch_cycles
  };
  BatchOpDev* d_ops = nullptr;
  BatchItemDev* d_items = nullptr;
  std::vector<Part> parts;
  int threads = 256;
};

static int log2i(int64_t w) {
  int l = 0;
  while ((1ll << l) < w) ++l;
  return l;
}

int64_t batch_boundary_bytes(const Op& op) {
  if (op.kind != OP_CYCLE || !op.cyc) return 0;
  return ((int64_t)op.B * (int64_t)op.cyc->slots.size() * op.U + 255) / 256 * 256;
}

int batch_launch_build(const std::vector<std::pair<const Op*, char*>>& ops, const LaunchCfg& cfg_in,
                       int grid_per_cu, int cus, char* scratch, int64_t scratch_bytes,
                       BatchLaunch** out, std::vector<size_t>* rejected) {
  *out = nullptr;
  // the batch's passes share each launch: footprints pack to the LDS target
  // instead of spreading every pass over the whole GPU
  LaunchCfg cfg = cfg_in;
  cfg.min_groups = 1;
  // workgroup size: the larger one when any LDS footprint needs it
  bool big = false;
  for (const auto& pr : ops) {
    const Op& op = *pr.first;
    if (op.kind == OP_BLOCK) {
      BlkArgs a;
      int64_t w, lds, groups;
      int vec;
      if (prepare_block_args(op, pr.second, cfg, a, &w, &vec, &lds, &groups) == hipSuccess)
        big = big || lds > 96 * 1024;
    } else if (op.kind == OP_SEGTILE) {
      SegArgs a;
      int64_t w, V, lds, groups;
      if (prepare_seg_args(op, pr.second, cfg, a, &w, &V, &lds, &groups) == hipSuccess)
        big = big || lds > 96 * 1024;
    } else if (!op.staged && op.kind >= OP_TILE && op.kind <= OP_ROW && !needs64(op)) {
      Args<uint32_t> a;
      Prepared p;
      prepare_lds<uint32_t>(op, pr.second, cfg, a, p);
      big = big || p.threads > cfg.threads;
    }
  }
  const int threads = big ? cfg.threads_big : cfg.threads;
  std::vector<BatchOpDev> hops;
  std::vector<std::pair<int64_t, BatchItemDev>> items;
  std::vector<BatchItemDev> saves, cycles;  // cycle items run in a launch without LDS
  int64_t lds = 0, boff = 0;
  auto reject = [&](size_t i) {
    if (rejected) rejected->push_back(i);
  };
  for (size_t i = 0; i < ops.size(); ++i) {
    const Op& op = *ops[i].first;
    char* data = ops[i].second;
    BatchOpDev d{};
    if (op.kind == OP_CYCLE) {
      // each tensor's boundary cells get their own part of the scratch
      const int64_t need = batch_boundary_bytes(op);
      if (boff + need > scratch_bytes) {
        reject(i);
        continue;
      }
      CycArgs a;
      int64_t w = 1;
      int variant = 0;
      bool empty = false;
      if (prepare_cycle_args(op, data, need ? scratch + boff : nullptr, need, &variant, a, &w,
                             &empty) != hipSuccess) {
        reject(i);
        continue;
      }
      if (empty) continue;
      boff += need;
      // cells that are not whole 16-byte words move in shifted 16-byte
      // chunks (cells of 8-byte words as of 4-byte words, as in single
      // launches) when every offset fits 32 bits, else word by word
      CycUOp cu;
      cu.a = a;
      const int64_t wu = w == 8 ? 4 : w;
      const bool shifted = w < 16 && cfg.cycle_variant == 0 && cycle_u_geometry(cu.a, cu.u, (uint64_t)wu);
      uint64_t groups;
      if (shifted) {
        std::memcpy(d.raw, &cu, sizeof(cu));
        d.variant = 120 + log2i(wu);
        groups = (cycu_threads(cu.a, cu.u) + threads - 1) / threads;
      } else {
        std::memcpy(d.raw, &a, sizeof(a));
        d.variant = 70 + log2i(w);
        groups = ((a.items << a.lane_log2) + threads - 1) / threads;
      }
      const uint32_t idx = (uint32_t)hops.size();
      hops.push_back(d);
      for (uint64_t g = 0; g < groups; ++g) cycles.push_back(BatchItemDev{idx, (uint32_t)g});
      // the boundary save below reads the plain cycle arguments
      std::memcpy(d.raw, &a, sizeof(a));
      if (a.nslots) {
        d.variant = 90 + log2i(w);
        const uint32_t sidx = (uint32_t)hops.size();
        hops.push_back(d);
        const uint64_t words = a.B * a.nslots * (a.U / (uint64_t)w);
        for (uint64_t g = 0; g < (words + threads - 1) / threads; ++g)
          saves.push_back(BatchItemDev{sidx, (uint32_t)g});
      }
      continue;
    }
    if (op.kind == OP_BLOCK) {
      BlkArgs a;
      int64_t w = 1, blds = 0, groups = 0;
      int vec = 0;
      if (prepare_block_args(op, data, cfg, a, &w, &vec, &blds, &groups) != hipSuccess) {
        reject(i);
        continue;
      }
      std::memcpy(d.raw, &a, sizeof(a));
      d.variant = 80 + log2i(w) * 2 + (vec && w < 16 ? 1 : 0);
      const uint32_t idx = (uint32_t)hops.size();
      hops.push_back(d);
      lds = std::max(lds, blds);
      for (int64_t g = 0; g < groups; ++g)
        items.push_back({blds, BatchItemDev{idx, (uint32_t)(g * a.K)}});
      continue;
    }
    if (op.kind == OP_SEGTILE) {
      SegArgs a;
      int64_t w = 1, V = 1, slds = 0, groups = 0;
      if (prepare_seg_args(op, data, cfg, a, &w, &V, &slds, &groups) != hipSuccess) {
        reject(i);
        continue;
      }
      std::memcpy(d.raw, &a, sizeof(a));
      d.variant = 100 + log2i(w) * 5 + log2i(V);
      const uint32_t idx = (uint32_t)hops.size();
      hops.push_back(d);
      lds = std::max(lds, slds);
      for (int64_t g = 0; g < groups; ++g)
        items.push_back({slds, BatchItemDev{idx, (uint32_t)(g * a.K)}});
      continue;
    }
    if (op.staged || op.kind < OP_TILE || op.kind > OP_ROW || needs64(op)) {
      reject(i);
      continue;
    }
    Args<uint32_t> a;
    Prepared p;
    prepare_lds<uint32_t>(op, data, cfg, a, p);
    std::memcpy(d.raw, &a, sizeof(a));
    d.variant = p.variant;
    const uint32_t idx = (uint32_t)hops.size();
    hops.push_back(d);
    lds = std::max(lds, p.lds);
    for (int64_t g = 0; g < p.groups; ++g)
      items.push_back({p.lds, BatchItemDev{idx, (uint32_t)(g * p.K)}});
  }
  if (items.empty() && saves.empty() && cycles.empty()) return EITHOT_OK;
  // largest items first so that round-robin assignment balances the tail
  std::stable_sort(items.begin(), items.end(),
                   [](const std::pair<int64_t, BatchItemDev>& x,
                      const std::pair<int64_t, BatchItemDev>& y) { return x.first > y.first; });
  (void)lds;
  BatchLaunch* L = new BatchLaunch;
  L->threads = threads;
  std::vector<BatchItemDev> hitems;
  auto add_part = [&](const std::vector<BatchItemDev>& v, int64_t plds, int64_t per_cu,
                      bool cyc) {
    if (v.empty()) return;
    const uint32_t n = (uint32_t)v.size();
    const uint32_t grid =
        grid_per_cu > 0 ? (uint32_t)std::min<int64_t>(n, (int64_t)cus * per_cu) : n;
    L->parts.push_back({(uint32_t)hitems.size(), n, grid, plds, cyc});
    hitems.insert(hitems.end(), v.begin(), v.end());
  };
  // without LDS, up to 2048 threads per CU stay resident
  add_part(saves, 0, 2048 / threads, true);
  add_part(cycles, 0, 2048 / threads, true);
  // LDS items (sorted by footprint, largest first) in classes of at most
  // 40 KiB, 80 KiB and the rest
  for (int64_t cap : {(int64_t)1 << 40, (int64_t)80 * 1024, (int64_t)40 * 1024}) {
    std::vector<BatchItemDev> cls;
    int64_t clds = 0;
    for (const auto& it : items) {
      const int64_t lo = cap == (int64_t)1 << 40 ? 80 * 1024 : cap == 80 * 1024 ? 40 * 1024 : -1;
      if (it.first <= cap && it.first > lo) {
        cls.push_back(it.second);
        clds = std::max(clds, it.first);
      }
    }
    add_part(cls, clds, grid_per_cu, false);
  }
  if (hipMalloc((void**)&L->d_ops, hops.size() * sizeof(BatchOpDev)) != hipSuccess ||
      hipMalloc((void**)&L->d_items, hitems.size() * sizeof(BatchItemDev)) != hipSuccess) {
    batch_launch_free(L);
    return EITHOT_ERR_NOMEM;
  }
  if (hipMemcpy(L->d_ops, hops.data(), hops.size() * sizeof(BatchOpDev), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(L->d_items, hitems.data(), hitems.size() * sizeof(BatchItemDev),
                hipMemcpyHostToDevice) != hipSuccess) {
    batch_launch_free(L);
    return EITHOT_ERR_DEVICE;
  }
  *out = L;
  return EITHOT_OK;
}

hipError_t batch_launch_enqueue(const BatchLaunch* L, hipStream_t st) {
  if (!L) return hipSuccess;
  // the boundary-save part comes first: every cycle pass's cells are saved
  // before any of them writes
  for (const BatchLaunch::Part& p : L->parts) {
    if (p.cycles)
      hipLaunchKernelGGL(k_batch_cycles, dim3(p.grid ? p.grid : 1), dim3(L->threads), 0, st,
                         L->d_ops, L->d_items + p.offset, p.count);
    else
      hipLaunchKernelGGL(k_batch, dim3(p.grid ? p.grid : 1), dim3(L->threads), (size_t)p.lds, st,
                         L->d_ops, L->d_items + p.offset, p.count);
  }
  return hipGetLastError();
}

void batch_launch_free(BatchLaunch* L) {
  if (!L) return;
  if (L->d_ops) (void)hipFree(L->d_ops);
  if (L->d_items) (void)hipFree(L->d_items);
  delete L;
}

int64_t batch_launch_items(const BatchLaunch* L) {
  int64_t n = 0;
  if (L)
    for (const BatchLaunch::Part& p : L->parts) n += p.count;
  return n;
}

int batch_launch_count(const BatchLaunch* L) { return L ? (int)L->parts.size() : 0; }

hipError_t launch_fill(char* data, uint64_t n, int elem, uint64_t seed, int mode, uint64_t first,
                       hipStream_t stream) {
  const uint64_t mixed = sm64_host(seed);
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 256ull * 128);
  if (n == 0) return hipSuccess;
  switch (elem) {
    case 1: hipLaunchKernelGGL(k_fill<uint8_t>, dim3(blocks), dim3(256), 0, stream, (uint8_t*)data, n, mixed, mode, first); break;
    case 2: hipLaunchKernelGGL(k_fill<uint16_t>, dim3(blocks), dim3(256), 0, stream, (uint16_t*)data, n, mixed, mode, first); break;
    case 4: hipLaunchKernelGGL(k_fill<uint32_t>, dim3(blocks), dim3(256), 0, stream, (uint32_t*)data, n, mixed, mode, first); break;
    case 8: hipLaunchKernelGGL(k_fill<uint64_t>, dim3(blocks), dim3(256), 0, stream, (uint64_t*)data, n, mixed, mode, first); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_verify(const char* data, const int64_t* dims, const int* perm0, int rank,
                         int elem, uint64_t seed, int mode, unsigned long long* d_mismatch,
                         hipStream_t stream) {
  if (rank < 1 || rank > 16) return hipErrorInvalidValue;
  VerifyArgs v{};
  v.rank = rank;
  uint64_t os[16], s = 1, vol = 1;
  for (int k = 0; k < rank; ++k) {
    os[perm0[k]] = s;
    s *= (uint64_t)dims[perm0[k]];
  }
  for (int k = 0; k < rank; ++k) {
    v.fd[k].init((uint64_t)dims[k]);
    v.ostride[k] = os[k];
    vol *= (uint64_t)dims[k];
  }
  v.vol = vol;
  const uint64_t mixed = sm64_host(seed);
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((vol + 255) / 256, 256ull * 128);
  switch (elem) {
    case 1: hipLaunchKernelGGL(k_verify<uint8_t>, dim3(blocks), dim3(256), 0, stream, (const uint8_t*)data, v, mixed, mode, d_mismatch); break;
    case 2: hipLaunchKernelGGL(k_verify<uint16_t>, dim3(blocks), dim3(256), 0, stream, (const uint16_t*)data, v, mixed, mode, d_mismatch); break;
    case 4: hipLaunchKernelGGL(k_verify<uint32_t>, dim3(blocks), dim3(256), 0, stream, (const uint32_t*)data, v, mixed, mode, d_mismatch); break;
    case 8: hipLaunchKernelGGL(k_verify<uint64_t>, dim3(blocks), dim3(256), 0, stream, (const uint64_t*)data, v, mixed, mode, d_mismatch); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

namespace {
template <typename Word, typename Idx, int V>
void cfg_set(int64_t bytes, hipError_t& e) {
  auto set = [&](const void* fn) {
    hipError_t r = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (r != hipSuccess) e = r;
  };
  set((const void*)k_tile<Word, Idx, V>);
  set((const void*)k_col<Word, Idx, MAP_ROT, V>);
  set((const void*)k_col<Word, Idx, MAP_ROTINV, V>);
  set((const void*)k_col<Word, Idx, MAP_COLSH, V>);
  set((const void*)k_col<Word, Idx, MAP_COLSHINV, V>);
  set((const void*)k_row<Word, Idx, MAP_ROWG, V>);
  set((const void*)k_row<Word, Idx, MAP_ROWD, V>);
}
template <typename Word>
void cfg_word(int64_t bytes, hipError_t& e) {
  constexpr int V = 16 / sizeof(Word) > 1 ? 16 / sizeof(Word) : 1;
  cfg_set<Word, uint32_t, 1>(bytes, e);
  cfg_set<Word, uint64_t, 1>(bytes, e);
  if (V > 1) {
    cfg_set<Word, uint32_t, V>(bytes, e);
    cfg_set<Word, uint64_t, V>(bytes, e);
  }
  auto set = [&](const void* fn) {
    hipError_t r = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (r != hipSuccess) e = r;
  };
  set((const void*)k_tile_oop<Word, uint32_t, 1>);
  set((const void*)k_tile_oop<Word, uint64_t, 1>);
  set((const void*)k_block<Word, 1, 4>);
  set((const void*)k_block<Word, 1, 8>);
  if (V > 1) set((const void*)k_block<Word, V, 4>);
  if (V > 1) set((const void*)k_block<Word, V, 8>);
  constexpr int w = sizeof(Word);
  set((const void*)k_segtile<Word, 1>);
  if (w >= 2) set((const void*)k_segtile<Word, (w >= 2 ? 2 : 1)>);
  if (w >= 4) set((const void*)k_segtile<Word, (w >= 4 ? 4 : 1)>);
  if (w >= 8) set((const void*)k_segtile<Word, (w >= 8 ? 8 : 1)>);
  if (w >= 16) set((const void*)k_segtile<Word, (w >= 16 ? 16 : 1)>);
  if constexpr (w == 16) {
    set((const void*)k_segtile<Word, 1, 2>);
    set((const void*)k_segtile<Word, 2, 2>);
    set((const void*)k_segtile<Word, 4, 2>);
    set((const void*)k_segtile<Word, 8, 2>);
    set((const void*)k_segtile<Word, 16, 2>);
    set((const void*)k_segtile<Word, 1, 3>);
    set((const void*)k_segtile<Word, 2, 3>);
    set((const void*)k_segtile<Word, 4, 3>);
    set((const void*)k_segtile<Word, 8, 3>);
    set((const void*)k_segtile<Word, 16, 3>);
  }
}
}  // namespace

hipError_t configure_kernels(int64_t bytes) {
  hipError_t e = hipSuccess;
  {
    hipError_t r = hipFuncSetAttribute((const void*)k_batch,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (r != hipSuccess) e = r;
  }
  cfg_word<W16>(bytes, e);
  cfg_word<uint64_t>(bytes, e);
  cfg_word<uint32_t>(bytes, e);
  cfg_word<uint16_t>(bytes, e);
  cfg_word<uint8_t>(bytes, e);
  return e;
}

TableAccount::TableAccount(std::atomic<int64_t>* c) : prev(t_account) { t_account = c; }
TableAccount::~TableAccount() { t_account = prev; }

}  // namespace eithot
