// C entry points over the reference's own integer-math sources
// (ITNT/src/gcd.cpp, ITNT/src/reduced_math.cpp), which this recipe compiles
// unmodified from /root/reference.  The declarations repeat the reference's
// signatures (ITNT/src/gcd.h:7, ITNT/src/reduced_math.h:15-35); no reference
// header is replaced.  Test infrastructure only.
#include <stdexcept>

namespace inplace {
void extended_gcd(int a, int b, int& gcd, int& mmi);
namespace detail {
void find_divisor(unsigned int denom, unsigned int& mul_coeff, unsigned int& shift_coeff);
void find_divisor(unsigned long long denom, unsigned long long& mul_coeff,
                  unsigned int& shift_coeff);
unsigned long long host_umulhi(unsigned long long x, unsigned long long y);
}  // namespace detail
}  // namespace inplace

extern "C" void ref_extended_gcd(int a, int b, int* gcd, int* mmi) {
  int g = 0, m = 0;
  inplace::extended_gcd(a, b, g, m);
  *gcd = g;
  *mmi = m;
}

// reduced_divisor_impl<unsigned>::div (reduced_math.h:56-58), host path
extern "C" int ref_div32(unsigned d, unsigned x, unsigned* q) {
  unsigned mul = 0, shift = 0;
  try {
    inplace::detail::find_divisor(d, mul, shift);
  } catch (const std::invalid_argument&) {
    return -1;
  }
  unsigned long long z = (unsigned long long)x * (unsigned long long)mul;
  *q = mul ? (unsigned)(z >> 32) >> shift : x;
  return 0;
}

extern "C" int ref_div64(unsigned long long d, unsigned long long x, unsigned long long* q) {
  unsigned long long mul = 0;
  unsigned shift = 0;
  try {
    inplace::detail::find_divisor(d, mul, shift);
  } catch (const std::invalid_argument&) {
    return -1;
  }
  *q = mul ? inplace::detail::host_umulhi(x, mul) >> shift : x;
  return 0;
}
