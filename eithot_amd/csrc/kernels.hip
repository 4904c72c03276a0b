// gfx950 pass kernels for in-place transposition.
//
// Every pass permutes units inside disjoint "footprints":
//   OP_TILE  one B-matrix (R x C units), written back transposed (C x R)
//   OP_COL   one column panel: W adjacent columns x all R rows of the grid
//   OP_ROW   one row of C units
// optionally restricted to one `us`-byte slice of every unit (units larger
// than a footprint budget are cut into slices that are permuted alike).
//
// LDS kernel: a workgroup loads K footprints into LDS with coalesced reads,
// barriers, and writes them back in gather order (dest unit q <- LDS unit
// src(q)), so reads and writes are both full-line HBM transactions.
// Staged kernels: footprints larger than LDS are copied to a scratch buffer
// (stage-out) and gathered back (stage-in); two launches on one stream.
#include "kernels.hpp"

#include <algorithm>

#include "eithot_status.h"
#include "maps.hpp"

namespace eithot {
namespace {

struct alignas(16) W16 {
  uint32_t v[4];
};

template <typename Idx>
struct PassArgs {
  char* data;
  char* scratch;
  uint64_t base;  // byte offset of footprint origin (batch 0)
  uint64_t U, us, RCU, CU;
  Idx R, C, W, S, NU, wu, K, nfp, P, per_fp;
  uint64_t fp0, G;  // staged: first footprint and count of this group
  int contiguous;   // footprint group is one contiguous byte range (S == 1, not COL)
  typename FDsel<Idx>::type fS, fR, fW, fP, fwu, fper;
  GridMaps<Idx> g;
};

__device__ __forceinline__ uint32_t xcd_logical(uint32_t bid, uint32_t n, int remap) {
  if (!remap || n < 16) return bid;
  const uint32_t per = n / 8;
  if (bid >= per * 8) return bid;
  return (bid % 8) * per + bid / 8;
}

template <typename Idx, int KIND>
struct Foot {
  uint64_t addr;  // byte address (offset from data) of unit 0 of this footprint slice
  Idx row;        // OP_ROW: row within the grid; OP_COL: first column of the panel
  Idx weff;       // OP_COL: columns present in the panel
};

template <typename Idx, int KIND>
__device__ __forceinline__ Foot<Idx, KIND> decode(const PassArgs<Idx>& a, Idx f) {
  Foot<Idx, KIND> r;
  const Idx o = a.fS.div(f);
  const Idx s = f - o * a.S;
  if (KIND == OP_TILE) {
    r.addr = a.base + (uint64_t)o * a.RCU + (uint64_t)s * a.us;
    r.row = 0;
    r.weff = 0;
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
    r.weff = 0;
  }
  return r;
}

// byte offset of footprint unit q (LDS order == memory order of the footprint)
template <typename Idx, int KIND>
__device__ __forceinline__ uint64_t unit_off(const PassArgs<Idx>& a, Idx q, Idx* col) {
  if (KIND == OP_COL) {
    const Idx i = a.fW.div(q);
    const Idx jj = q - i * a.W;
    *col = jj;
    return (uint64_t)i * a.CU + (uint64_t)jj * a.U;
  }
  *col = 0;
  return (uint64_t)q * a.U;
}

// footprint unit that holds the data for destination unit q
template <typename Idx, int KIND, int MAP>
__device__ __forceinline__ Idx src_unit(const PassArgs<Idx>& a, const Foot<Idx, KIND>& f, Idx q) {
  if (KIND == OP_TILE) {  // q = j*R + i in the transposed C x R output
    const Idx j = a.fR.div(q);
    const Idx i = q - j * a.R;
    return i * a.C + j;
  } else if (KIND == OP_COL) {
    const Idx i = a.fW.div(q);
    const Idx jj = q - i * a.W;
    const Idx j = f.row + jj;
    Idx s;
    if (MAP == MAP_ROT) s = a.g.rot(i, j);
    else if (MAP == MAP_ROTINV) s = a.g.rotinv(i, j);
    else if (MAP == MAP_COLSH) s = a.g.colsh(i, j);
    else s = a.g.colshinv(i, j);
    return s * a.W + jj;
  } else {
    if (MAP == MAP_ROWG) return a.g.rowg(f.row, q);
    return a.g.rowd(f.row, q);
  }
}

template <typename Word, typename Idx, int KIND, int MAP>
__global__ __launch_bounds__(256) void k_lds(PassArgs<Idx> a, int remap) {
  extern __shared__ __align__(16) unsigned char lds[];
  Word* L = reinterpret_cast<Word*>(lds);
  const Idx wg = (Idx)xcd_logical(blockIdx.x, gridDim.x, remap);
  const Idx f0 = wg * a.K;
  const Idx nf = (a.nfp - f0) < a.K ? (a.nfp - f0) : a.K;
  const Idx total = nf * a.per_fp;
  const Idx wu = a.wu;
  const Idx tid = threadIdx.x, nt = blockDim.x;

  if (a.contiguous) {
    // the K footprints form one contiguous byte range: plain coalesced copy
    const Foot<Idx, KIND> ff = decode<Idx, KIND>(a, f0);
    const char* src = a.data + ff.addr;
    const uint64_t nbytes = (uint64_t)total * sizeof(Word);
    if ((((uintptr_t)src) & 15) == 0 && (nbytes & 15) == 0) {
      const W16* s16 = reinterpret_cast<const W16*>(src);
      W16* l16 = reinterpret_cast<W16*>(lds);
      const Idx n16 = (Idx)(nbytes / 16);
      for (Idx x = tid; x < n16; x += nt) l16[x] = s16[x];
    } else {
      const Word* sw = reinterpret_cast<const Word*>(src);
      for (Idx x = tid; x < total; x += nt) L[x] = sw[x];
    }
  } else {
    for (Idx idx = tid; idx < total; idx += nt) {
      const Idx fl = a.K == 1 ? 0 : a.fper.div(idx);
      const Idx rem = idx - fl * a.per_fp;
      const Idx q = wu == 1 ? rem : a.fwu.div(rem);
      const Idx x = rem - q * wu;
      const Foot<Idx, KIND> f = decode<Idx, KIND>(a, f0 + fl);
      Idx col;
      const uint64_t off = unit_off<Idx, KIND>(a, q, &col);
      if (KIND == OP_COL && col >= f.weff) continue;
      L[idx] = *reinterpret_cast<const Word*>(a.data + f.addr + off + (uint64_t)x * sizeof(Word));
    }
  }
  __syncthreads();
  for (Idx idx = tid; idx < total; idx += nt) {
    const Idx fl = a.K == 1 ? 0 : a.fper.div(idx);
    const Idx rem = idx - fl * a.per_fp;
    const Idx q = wu == 1 ? rem : a.fwu.div(rem);
    const Idx x = rem - q * wu;
    const Foot<Idx, KIND> f = decode<Idx, KIND>(a, f0 + fl);
    Idx col;
    const uint64_t off = unit_off<Idx, KIND>(a, q, &col);
    if (KIND == OP_COL && col >= f.weff) continue;
    const Idx s = src_unit<Idx, KIND, MAP>(a, f, q);
    *reinterpret_cast<Word*>(a.data + f.addr + off + (uint64_t)x * sizeof(Word)) =
        L[(uint64_t)fl * a.per_fp + (uint64_t)s * wu + x];
  }
}

// staged pass, part 1: footprints [fp0, fp0+G) -> scratch (footprint-major)
template <typename Word, typename Idx, int KIND>
__global__ __launch_bounds__(256) void k_stage_out(PassArgs<Idx> a) {
  const uint64_t total = a.G * (uint64_t)a.per_fp;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  Word* sc = reinterpret_cast<Word*>(a.scratch);
  for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += stride) {
    const uint64_t fl = idx / a.per_fp;
    const Idx rem = (Idx)(idx - fl * a.per_fp);
    const Idx q = a.wu == 1 ? rem : a.fwu.div(rem);
    const Idx x = rem - q * a.wu;
    const Foot<Idx, KIND> f = decode<Idx, KIND>(a, (Idx)(a.fp0 + fl));
    Idx col;
    const uint64_t off = unit_off<Idx, KIND>(a, q, &col);
    if (KIND == OP_COL && col >= f.weff) continue;
    sc[idx] = *reinterpret_cast<const Word*>(a.data + f.addr + off + (uint64_t)x * sizeof(Word));
  }
}

// staged pass, part 2: gather from scratch back into the footprints
template <typename Word, typename Idx, int KIND, int MAP>
__global__ __launch_bounds__(256) void k_stage_in(PassArgs<Idx> a) {
  const uint64_t total = a.G * (uint64_t)a.per_fp;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const Word* sc = reinterpret_cast<const Word*>(a.scratch);
  for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += stride) {
    const uint64_t fl = idx / a.per_fp;
    const Idx rem = (Idx)(idx - fl * a.per_fp);
    const Idx q = a.wu == 1 ? rem : a.fwu.div(rem);
    const Idx x = rem - q * a.wu;
    const Foot<Idx, KIND> f = decode<Idx, KIND>(a, (Idx)(a.fp0 + fl));
    Idx col;
    const uint64_t off = unit_off<Idx, KIND>(a, q, &col);
    if (KIND == OP_COL && col >= f.weff) continue;
    const Idx s = src_unit<Idx, KIND, MAP>(a, f, q);
    *reinterpret_cast<Word*>(a.data + f.addr + off + (uint64_t)x * sizeof(Word)) =
        sc[fl * a.per_fp + (uint64_t)s * a.wu + x];
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

template <typename Idx>
PassArgs<Idx> make_args(const Op& op, char* data, int64_t wbytes) {
  PassArgs<Idx> a{};
  a.data = data;
  a.scratch = nullptr;
  a.base = op.base;
  a.U = op.U;
  a.us = op.us;
  a.RCU = (uint64_t)op.R * op.C * op.U;
  a.CU = (uint64_t)op.C * op.U;
  a.R = (Idx)op.R;
  a.C = (Idx)op.C;
  a.W = (Idx)(op.kind == OP_COL ? op.W : 1);
  a.S = (Idx)(op.U / op.us);
  a.NU = (Idx)op.footprint_units();
  a.wu = (Idx)(op.us / wbytes);
  a.per_fp = a.NU * a.wu;
  a.nfp = (Idx)op.num_footprints();
  a.P = (Idx)((op.C + a.W - 1) / a.W);
  a.K = 1;
  a.fp0 = 0;
  a.G = 0;
  a.contiguous = 0;
  a.fS.init(a.S);
  a.fR.init(op.R);
  a.fW.init(a.W);
  a.fP.init(a.P);
  a.fwu.init(a.wu);
  a.fper.init(a.per_fp);
  if (op.kind != OP_TILE) a.g.init(op.m, op.n, op.c, op.a, op.b, op.k, op.binv);
  return a;
}

bool needs64(const Op& op) {
  const unsigned __int128 lim = (unsigned __int128)1 << 31;
  auto big = [&](unsigned __int128 v) { return v >= lim; };
  if (big((unsigned __int128)op.num_footprints())) return true;
  if (big((unsigned __int128)op.footprint_units() * (op.us))) return true;
  if (op.kind == OP_TILE) return big((unsigned __int128)op.R * op.C);
  const unsigned __int128 m = op.m, n = op.n, a = op.a, b = op.b, c = op.c;
  return big(m * m + n) || big(b * b) || big(a * a) || big(a * b) || big(n + 2 * m) ||
         big(c * b + m) || big(m * 3);
}

int64_t word_bytes(const Op& op, uintptr_t data) {
  int64_t w = 16;
  while (w > 1 && ((op.us % w) || (op.U % w) || ((data + op.base) % w))) w /= 2;
  return w;
}

template <typename Word, typename Idx, int KIND, int MAP>
hipError_t run_typed(const Op& op, char* data, char* scratch, int64_t scratch_bytes,
                     const LaunchCfg& cfg, hipStream_t st) {
  PassArgs<Idx> a = make_args<Idx>(op, data, sizeof(Word));
  const int64_t fpb = (int64_t)a.per_fp * (int64_t)sizeof(Word);
  if (!op.staged) {
    int64_t K = std::max<int64_t>(1, cfg.lds_target / std::max<int64_t>(fpb, 1));
    // keep at least ~4 workgroups per CU in flight when there is enough work
    const int64_t nfp = (int64_t)a.nfp;
    const int64_t want = 1024;
    if (nfp / K < want) K = std::max<int64_t>(1, nfp / want);
    a.K = (Idx)K;
    a.fper.init(a.per_fp);
    a.contiguous = (KIND != OP_COL && a.S == 1) ? 1 : 0;
    const int64_t nwg = (nfp + K - 1) / K;
    const size_t lds = (size_t)(K * fpb);
    hipLaunchKernelGGL((k_lds<Word, Idx, KIND, MAP>), dim3((uint32_t)nwg), dim3(cfg.threads), lds,
                       st, a, cfg.xcd_remap);
    return hipGetLastError();
  }
  if (!scratch || scratch_bytes < fpb) return hipErrorInvalidValue;
  a.scratch = scratch;
  const uint64_t nfp = (uint64_t)a.nfp;
  const uint64_t G = std::min<uint64_t>(nfp, (uint64_t)(scratch_bytes / fpb));
  for (uint64_t f0 = 0; f0 < nfp; f0 += G) {
    a.fp0 = f0;
    a.G = std::min<uint64_t>(G, nfp - f0);
    const uint64_t words = a.G * (uint64_t)a.per_fp;
    const uint64_t blocks = std::min<uint64_t>((words + 255) / 256, 256ull * 64);
    hipLaunchKernelGGL((k_stage_out<Word, Idx, KIND>), dim3((uint32_t)blocks), dim3(256), 0, st, a);
    hipLaunchKernelGGL((k_stage_in<Word, Idx, KIND, MAP>), dim3((uint32_t)blocks), dim3(256), 0, st,
                       a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

template <typename Word, typename Idx>
hipError_t run_map(const Op& op, char* d, char* s, int64_t sb, const LaunchCfg& c, hipStream_t st) {
  if (op.kind == OP_TILE) return run_typed<Word, Idx, OP_TILE, MAP_NONE>(op, d, s, sb, c, st);
  if (op.kind == OP_COL) {
    switch (op.map) {
      case MAP_ROT: return run_typed<Word, Idx, OP_COL, MAP_ROT>(op, d, s, sb, c, st);
      case MAP_ROTINV: return run_typed<Word, Idx, OP_COL, MAP_ROTINV>(op, d, s, sb, c, st);
      case MAP_COLSH: return run_typed<Word, Idx, OP_COL, MAP_COLSH>(op, d, s, sb, c, st);
      case MAP_COLSHINV: return run_typed<Word, Idx, OP_COL, MAP_COLSHINV>(op, d, s, sb, c, st);
    }
  }
  if (op.kind == OP_ROW) {
    if (op.map == MAP_ROWG) return run_typed<Word, Idx, OP_ROW, MAP_ROWG>(op, d, s, sb, c, st);
    if (op.map == MAP_ROWD) return run_typed<Word, Idx, OP_ROW, MAP_ROWD>(op, d, s, sb, c, st);
  }
  return hipErrorInvalidValue;
}

template <typename Word>
hipError_t run_idx(const Op& op, char* d, char* s, int64_t sb, const LaunchCfg& c, hipStream_t st) {
  if (needs64(op)) return run_map<Word, uint64_t>(op, d, s, sb, c, st);
  return run_map<Word, uint32_t>(op, d, s, sb, c, st);
}

// ---- pattern fill / verify ----------------------------------------------------

__device__ __forceinline__ uint64_t sm64(uint64_t x) {
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
  if (op.staged || op.kind == 0) return 0;
  const int64_t fpb = op.footprint_units() * op.us;
  int64_t K = std::max<int64_t>(1, cfg.lds_target / std::max<int64_t>(fpb, 1));
  const int64_t nfp = op.num_footprints();
  if (nfp / K < 1024) K = std::max<int64_t>(1, nfp / 1024);
  return K * fpb;
}

hipError_t launch_op(const Op& op, char* data, char* scratch, int64_t scratch_bytes,
                     const LaunchCfg& cfg, hipStream_t stream) {
  if (op.kind == 0 || op.us <= 0 || op.U % op.us) return hipErrorInvalidValue;
  switch (word_bytes(op, (uintptr_t)data)) {
    case 16: return run_idx<W16>(op, data, scratch, scratch_bytes, cfg, stream);
    case 8: return run_idx<uint64_t>(op, data, scratch, scratch_bytes, cfg, stream);
    case 4: return run_idx<uint32_t>(op, data, scratch, scratch_bytes, cfg, stream);
    case 2: return run_idx<uint16_t>(op, data, scratch, scratch_bytes, cfg, stream);
    default: return run_idx<uint8_t>(op, data, scratch, scratch_bytes, cfg, stream);
  }
}

hipError_t launch_fill(char* data, uint64_t n, int elem, uint64_t seed, int mode, uint64_t first,
                       hipStream_t stream) {
  const uint64_t mixed = [&] {
    uint64_t x = seed + 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
  }();
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
  const uint64_t mixed = [&] {
    uint64_t x = seed + 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
  }();
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
template <typename Word, typename Idx>
hipError_t cfg_words(int64_t bytes) {
  hipError_t e = hipSuccess;
  auto set = [&](const void* fn) {
    hipError_t r = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (r != hipSuccess) e = r;
  };
  set((const void*)k_lds<Word, Idx, OP_TILE, MAP_NONE>);
  set((const void*)k_lds<Word, Idx, OP_COL, MAP_ROT>);
  set((const void*)k_lds<Word, Idx, OP_COL, MAP_ROTINV>);
  set((const void*)k_lds<Word, Idx, OP_COL, MAP_COLSH>);
  set((const void*)k_lds<Word, Idx, OP_COL, MAP_COLSHINV>);
  set((const void*)k_lds<Word, Idx, OP_ROW, MAP_ROWG>);
  set((const void*)k_lds<Word, Idx, OP_ROW, MAP_ROWD>);
  return e;
}
template <typename Word>
hipError_t cfg_idx(int64_t bytes) {
  hipError_t e = cfg_words<Word, uint32_t>(bytes);
  hipError_t f = cfg_words<Word, uint64_t>(bytes);
  return e != hipSuccess ? e : f;
}
}  // namespace

hipError_t configure_kernels(int64_t bytes) {
  hipError_t e[5] = {cfg_idx<W16>(bytes), cfg_idx<uint64_t>(bytes), cfg_idx<uint32_t>(bytes),
                     cfg_idx<uint16_t>(bytes), cfg_idx<uint8_t>(bytes)};
  for (hipError_t x : e)
    if (x != hipSuccess) return x;
  return hipSuccess;
}

}  // namespace eithot
