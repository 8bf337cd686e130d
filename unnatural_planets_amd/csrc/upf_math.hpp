// upf_math.hpp -- pow and atan for the library's FP32 paths, computed in double
// arithmetic from IEEE basic operations only (+, -, *, /, comparisons, bit moves), so
// that every compiler and target that honours IEEE double rounding gives the same bits.
//
// The elevation modes (sources/elevationModes.cpp: pow), the material stack
// (sources/coloringModes.cpp: pow, atan) and the further shapes use these in place of the
// C library's powf / atanf, whose last bit differs between implementations.  The double
// result is rounded once to float, so the float result is the correctly rounded value
// except within ~1e-15 relative of a rounding boundary.
//
// Methods (textbook):
//   ln:   x = m 2^k with m in [sqrt(1/2), sqrt(2)), s = (m - 1) / (m + 1), |s| <= 0.1716,
//         ln m = 2 atanh s = 2 (s + s^3/3 + ... + s^21/21)  (next term < 2^-60 relative)
//   exp:  y = n ln2 + r with n = round(y / ln2), |r| <= 0.347, e^r by its Taylor series
//         to r^16/16!  (next term < 2^-56), times 2^n assembled from bits
//   atan: |x| > 1 through pi/2 - atan(1/|x|); then table-driven reduction (Muller,
//         "Elementary Functions", ch. 6): c = round(8a)/8, atan a = atan c + atan t with
//         t = (a - c) / (1 + a c), |t| <= 1/16, atan t by its series to t^15/15; atan(i/8)
//         are double constants.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define UPF_MATH_FN __host__ __device__ inline
#else
#define UPF_MATH_FN inline
#endif

namespace upf_math {

constexpr double kLn2 = 0.6931471805599453;
constexpr double kInvLn2 = 1.4426950408889634;
constexpr double kHalfPi = 1.5707963267948966;

// ln x for finite x > 0 (a normal double: every float > 0 is one)
UPF_MATH_FN double ln_pos(double x) {
  const uint64_t b = __builtin_bit_cast(uint64_t, x);
  int k = (int)((b >> 52) & 0x7ff) - 1023;
  double m = __builtin_bit_cast(double, (b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
  if (m > 1.4142135623730951) {
    m = m * 0.5;
    k = k + 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double s2 = s * s;
  double p = 1.0 / 21.0;
  p = p * s2 + 1.0 / 19.0;
  p = p * s2 + 1.0 / 17.0;
  p = p * s2 + 1.0 / 15.0;
  p = p * s2 + 1.0 / 13.0;
  p = p * s2 + 1.0 / 11.0;
  p = p * s2 + 1.0 / 9.0;
  p = p * s2 + 1.0 / 7.0;
  p = p * s2 + 1.0 / 5.0;
  p = p * s2 + 1.0 / 3.0;
  p = p * s2 + 1.0;
  return (double)k * kLn2 + 2.0 * s * p;
}

// e^y for |y| <= 700
UPF_MATH_FN double exp_small(double y) {
  const double nf = y * kInvLn2 + 0.5;
  int n = (int)nf;
  if ((double)n > nf) n = n - 1;  // floor
  const double r = y - (double)n * kLn2;
  double p = 1.0 / 20922789888000.0;  // 1/16!
  p = p * r + 1.0 / 1307674368000.0;
  p = p * r + 1.0 / 87178291200.0;
  p = p * r + 1.0 / 6227020800.0;
  p = p * r + 1.0 / 479001600.0;
  p = p * r + 1.0 / 39916800.0;
  p = p * r + 1.0 / 3628800.0;
  p = p * r + 1.0 / 362880.0;
  p = p * r + 1.0 / 40320.0;
  p = p * r + 1.0 / 5040.0;
  p = p * r + 1.0 / 720.0;
  p = p * r + 1.0 / 120.0;
  p = p * r + 1.0 / 24.0;
  p = p * r + 1.0 / 6.0;
  p = p * r + 0.5;
  p = p * r + 1.0;
  p = p * r + 1.0;
  const double two_n = __builtin_bit_cast(double, (uint64_t)(n + 1023) << 52);
  return p * two_n;
}

// x^e in FP32 (the C library's powf for the cases the library meets): x > 0 through
// e^(e ln x); x = 0 gives 0 (e > 0), 1 (e = 0) or inf (e < 0); x < 0 with an integer e
// gives the signed power, with any other e NaN; NaN stays NaN.
UPF_MATH_FN float powf_(float x, float e) {
  if (e == 0.0f) return 1.0f;
  if (x != x || e != e) return x + e;
  if (x == 0.0f) return e > 0.0f ? 0.0f : __builtin_inff();
  double a = (double)x;
  bool neg = false;
  if (x < 0.0f) {
    const double ed = (double)e;
    const double half = ed * 0.5;
    const int64_t ei = (int64_t)ed;
    if ((double)ei != ed) return __builtin_nanf("");
    neg = (double)(int64_t)half != half;  // odd exponent
    a = -a;
  }
  if (a == (double)__builtin_inff()) return neg ? -__builtin_inff() : (e > 0.0f ? __builtin_inff() : 0.0f);
  const double y = (double)e * ln_pos(a);
  float r;
  if (y > 89.0) r = __builtin_inff();        // above FLT_MAX
  else if (y < -104.0) r = 0.0f;             // below the smallest float denormal
  else r = (float)exp_small(y);
  return neg ? -r : r;
}

// atan x in FP32
UPF_MATH_FN float atanf_(float x) {
  if (x != x) return x;
  double a = (double)x;
  const bool neg = a < 0.0;
  if (neg) a = -a;
  const bool inv = a > 1.0;
  if (inv) a = 1.0 / a;  // a = inf gives 0: atan = pi/2
  const int i = (int)(a * 8.0 + 0.5);
  const double c = (double)i * 0.125;
  const double t = (a - c) / (1.0 + a * c);
  const double t2 = t * t;
  double p = -1.0 / 15.0;
  p = p * t2 + 1.0 / 13.0;
  p = p * t2 - 1.0 / 11.0;
  p = p * t2 + 1.0 / 9.0;
  p = p * t2 - 1.0 / 7.0;
  p = p * t2 + 1.0 / 5.0;
  p = p * t2 - 1.0 / 3.0;
  p = p * t2 + 1.0;
  // atan(i / 8), i = 0 .. 8, as double constants (a select chain: no array in memory)
  const double ac = i == 0 ? 0.0
                  : i == 1 ? 0.12435499454676144
                  : i == 2 ? 0.24497866312686414
                  : i == 3 ? 0.35877067027057225
                  : i == 4 ? 0.4636476090008061
                  : i == 5 ? 0.5585993153435624
                  : i == 6 ? 0.6435011087932844
                  : i == 7 ? 0.7188299996216245
                           : 0.7853981633974483;
  double r = ac + t * p;
  if (inv) r = kHalfPi - r;
  return (float)(neg ? -r : r);
}

}  // namespace upf_math
