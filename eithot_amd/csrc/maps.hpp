// Index maps of the device passes and division by run-time invariant
// integers.  Host and device share this header: the planner evaluates the
// same maps on the host for its checks, the kernels on gfx950.
//
// Catanzaro decomposition of the transpose of a row-major m x n grid
// (Catanzaro, Keller, Garland, "A Decomposition for In-place Matrix
// Transposition", PPoPP 2014), with c = gcd(m, n), a = m/c, b = n/c,
// k = a^{-1} mod b, binv = b^{-1} mod a.  The element at (i, j) must reach
// linear position j*m + i.  Every map below is in gather form: the pass
// writes dest[i][j] = old[src(i, j)].
//
//   C2R (m x n row-major -> n x m):
//     ROT    column j:  src row = (i + floor(j/b)) mod m         (if c > 1)
//     ROWG   row i:     src col = q*b + ((rho*k) mod b), with
//                       q = (j - i) mod c, t = (i + q) mod m,
//                       rho = ((j - t) mod n) / c
//     COLSH  column j:  src row = (i*n + j - floor(i/a)) mod m
//   R2C is the inverse sequence on the same grid:
//     COLSHINV column j: src row = a*t + ((v * binv) mod a), with
//                       t = (j - i) mod c, v = ((i - j + t) mod m) / c
//     ROWD   row i:     src col = (c*((r*a) mod b) + (i + q) mod m) mod n,
//                       j = q*b + r
//     ROTINV column j:  src row = (i - floor(j/b)) mod m         (if c > 1)
// ROWG is the inverse of ROWD and COLSHINV of COLSH (derivation in the
// design notes); tests/ check each map against the reference's functors.
#pragma once

#include <cstdint>

#include <hip/hip_runtime.h>
#define EH_HD __host__ __device__ __forceinline__

namespace eithot {

EH_HD uint64_t mulhi64(uint64_t x, uint64_t y) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul64hi(x, y);
#else
  return (uint64_t)(((unsigned __int128)x * y) >> 64);
#endif
}

EH_HD uint32_t mulhi32(uint32_t x, uint32_t y) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umulhi(x, y);
#else
  return (uint32_t)(((uint64_t)x * y) >> 32);
#endif
}

// Division by an invariant d >= 1 (Granlund & Montgomery, "Division by
// invariant integers using multiplication", PLDI 1994, round-up method):
// l = ceil(log2 d), mul = floor(2^N (2^l - d) / d) + 1,
// q = (t + ((x - t) >> min(l,1))) >> max(l-1,0) with t = mulhi(x, mul).
// Exact for every x < 2^N.
struct FD64 {
  uint64_t d = 1, mul = 1;
  uint32_t s1 = 0, s2 = 0;
  FD64() = default;
  explicit FD64(uint64_t dv) { init(dv); }
  void init(uint64_t dv) {
    d = dv ? dv : 1;
    uint32_t l = 0;
    while (l < 64 && ((uint64_t)1 << l) < d) ++l;
    unsigned __int128 num = ((unsigned __int128)1 << 64) * ((((unsigned __int128)1) << l) - d);
    mul = (uint64_t)(num / d) + 1;
    s1 = l < 1 ? l : 1;
    s2 = l > 0 ? l - 1 : 0;
  }
  EH_HD uint64_t div(uint64_t x) const {
    uint64_t t = mulhi64(x, mul);
    return (t + ((x - t) >> s1)) >> s2;
  }
  EH_HD uint64_t mod(uint64_t x) const { return x - div(x) * d; }
};

struct FD32 {
  uint32_t d = 1, mul = 1;
  uint32_t s1 = 0, s2 = 0;
  FD32() = default;
  explicit FD32(uint64_t dv) { init(dv); }
  void init(uint64_t dv) {
    d = dv ? (uint32_t)dv : 1;
    uint32_t l = 0;
    while (l < 32 && ((uint64_t)1 << l) < d) ++l;
    uint64_t num = ((uint64_t)1 << 32) * ((((uint64_t)1) << l) - d);
    mul = (uint32_t)(num / d + 1);
    s1 = l < 1 ? l : 1;
    s2 = l > 0 ? l - 1 : 0;
  }
  EH_HD uint32_t div(uint32_t x) const {
    uint32_t t = mulhi32(x, mul);
    return (t + ((x - t) >> s1)) >> s2;
  }
  EH_HD uint32_t mod(uint32_t x) const { return x - div(x) * d; }
};

// Source cell of a digit permutation, stepped cell by cell.  A cell index
// o has mixed-radix digits d_k (radix frad[k].d, fastest first) and its
// source is sum d_k * stride[k]; step() moves to o + 1 with one add per
// digit that changes (an odometer), instead of ndig divisions.  ND bounds
// ndig at compile time: kernels instantiated for at most 4 digits keep the
// radices of digits 4..7 out of their scalar registers.
template <int ND>
struct DigitOdometerN {
  uint32_t dg[ND];
  uint32_t src;
  EH_HD void init(const FD32* frad, const uint32_t* stride, uint32_t ndig, uint32_t o) {
    src = 0;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      dg[d] = 0;
      if (d >= (int)ndig) continue;
      const uint32_t nq = frad[d].div(o);
      dg[d] = o - nq * frad[d].d;
      src += dg[d] * stride[d];
      o = nq;
    }
  }
  EH_HD void step(const FD32* frad, const uint32_t* stride, uint32_t ndig) {
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      if (d >= (int)ndig) break;
      src += stride[d];
      if (++dg[d] < frad[d].d) break;
      dg[d] = 0;
      src -= frad[d].d * stride[d];
    }
  }
};
using DigitOdometer = DigitOdometerN<8>;

template <typename Idx>
struct FDsel;
template <>
struct FDsel<uint32_t> {
  using type = FD32;
};
template <>
struct FDsel<uint64_t> {
  using type = FD64;
};

// Constants of the m x n grid of one Catanzaro pass.
template <typename Idx>
struct GridMaps {
  using FD = typename FDsel<Idx>::type;
  Idx m = 1, n = 1, c = 1, a = 1, b = 1, k = 0, binv = 0, nm = 0;  // nm = n mod m
  FD fm, fc, fa, fb;
  // b as a 64-bit divisor: the row maps' products rho * k < b * b and
  // r * a < a * b are reduced mod b in 64 bits even when Idx is 32-bit
  FD64 fb64;
  // x * y mod b for x, y < 2^32 (exact in 64 bits)
  EH_HD Idx mulmod_b(Idx x, Idx y) const {
    if constexpr (sizeof(Idx) == 4) return (Idx)fb64.mod((uint64_t)x * (uint64_t)y);
    else return fb.mod(x * y);
  }

  void init(uint64_t m_, uint64_t n_, uint64_t c_, uint64_t a_, uint64_t b_, uint64_t k_,
            uint64_t binv_) {
    m = (Idx)m_;
    n = (Idx)n_;
    c = (Idx)c_;
    a = (Idx)a_;
    b = (Idx)b_;
    k = (Idx)k_;
    binv = (Idx)binv_;
    nm = (Idx)(n_ % m_);
    fm.init(m_);
    fc.init(c_);
    fa.init(a_);
    fb.init(b_);
    fb64.init(b_);
  }

  EH_HD Idx rot(Idx i, Idx j) const {
    Idx r = i + fb.div(j);
    return r >= m ? r - m : r;
  }
  EH_HD Idx rotinv(Idx i, Idx j) const {
    Idx q = fb.div(j);
    return i >= q ? i - q : i + m - q;
  }
  EH_HD Idx colsh(Idx i, Idx j) const {
    Idx x = fm.mod(i * nm + j);
    Idx t = fa.div(i);
    return x >= t ? x - t : x + m - t;
  }
  EH_HD Idx colshinv(Idx s, Idx j) const {
    Idx jc = fc.mod(j), sc = fc.mod(s);
    Idx t = jc >= sc ? jc - sc : jc + c - sc;
    Idx jm = fm.mod(j);
    Idx y = s + t + m - jm;  // < 3m
    if (y >= m) y -= m;
    if (y >= m) y -= m;
    Idx v = fc.div(y);
    Idx u = fa.mod(v * binv);
    return a * t + u;
  }
  EH_HD Idx rowg(Idx i, Idx j) const {
    Idx ic = fc.mod(i), jc = fc.mod(j);
    Idx q = jc >= ic ? jc - ic : jc + c - ic;
    Idx t = i + q;
    if (t >= m) t -= m;
    Idx y = j >= t ? j - t : j + n - t;
    Idx rho = fc.div(y);
    Idx r = mulmod_b(rho, k);
    return q * b + r;
  }
  EH_HD Idx rowd(Idx i, Idx j) const {
    Idx q = fb.div(j);
    Idx r = j - q * b;
    Idx rho = mulmod_b(r, a);
    Idx t = i + q;
    if (t >= m) t -= m;
    Idx y = c * rho + t;
    return y >= n ? y - n : y;
  }
};

}  // namespace eithot
