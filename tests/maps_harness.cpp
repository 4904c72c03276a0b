// Host-side evaluation of the product's index maps and magic division
// (eithot_amd/csrc/maps.hpp) for the CPU tests; compiled by the test with
// hipcc, never part of the product.
#include <cstdint>

#include "../eithot_amd/csrc/maps.hpp"

using namespace eithot;

extern "C" int64_t mh_eval(int which, int wide, int64_t m, int64_t n, int64_t c, int64_t a,
                           int64_t b, int64_t k, int64_t binv, int64_t i, int64_t j) {
  if (wide) {
    GridMaps<uint64_t> g;
    g.init(m, n, c, a, b, k, binv);
    switch (which) {
      case 1: return g.rot(i, j);
      case 2: return g.rotinv(i, j);
      case 3: return g.colsh(i, j);
      case 4: return g.colshinv(i, j);
      case 5: return g.rowg(i, j);
      case 6: return g.rowd(i, j);
    }
  } else {
    GridMaps<uint32_t> g;
    g.init(m, n, c, a, b, k, binv);
    switch (which) {
      case 1: return g.rot((uint32_t)i, (uint32_t)j);
      case 2: return g.rotinv((uint32_t)i, (uint32_t)j);
      case 3: return g.colsh((uint32_t)i, (uint32_t)j);
      case 4: return g.colshinv((uint32_t)i, (uint32_t)j);
      case 5: return g.rowg((uint32_t)i, (uint32_t)j);
      case 6: return g.rowd((uint32_t)i, (uint32_t)j);
    }
  }
  return -1;
}

extern "C" uint64_t mh_div64(uint64_t d, uint64_t x) { return FD64(d).div(x); }
extern "C" uint32_t mh_div32(uint32_t d, uint32_t x) { return FD32(d).div(x); }

// The fused block pass's digit odometer: for every cell o of a block with
// the given radices and strides, init(o) and the steps from init(0) must
// both give sum digit_k(o) * stride_k; returns the first o where they do
// not, or -1.
extern "C" int64_t mh_odometer(const uint32_t* radix, const uint32_t* stride, int ndig) {
  FD32 frad[8];
  uint32_t st[8] = {0};
  uint64_t n = 1;
  for (int d = 0; d < ndig; ++d) {
    frad[d].init(radix[d]);
    st[d] = stride[d];
    n *= radix[d];
  }
  DigitOdometer walk;
  walk.init(frad, st, (uint32_t)ndig, 0);
  for (uint64_t o = 0; o < n; ++o) {
    uint64_t q = o, want = 0;
    for (int d = 0; d < ndig; ++d) {
      want += (q % radix[d]) * stride[d];
      q /= radix[d];
    }
    DigitOdometer direct;
    direct.init(frad, st, (uint32_t)ndig, (uint32_t)o);
    if (direct.src != want || walk.src != want) return (int64_t)o;
    walk.step(frad, st, (uint32_t)ndig);
  }
  // a full turn comes back to cell 0
  return walk.src == 0 ? -1 : (int64_t)n;
}
