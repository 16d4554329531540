// exp2_f64.hpp -- 2^x in FP64 for the soft-min and weighted-sum passes.
//
// Range reduction x = k + r with k = rint(x) and r in [-1/2, 1/2], 2^r by a
// polynomial of degree 11, 2^k applied exactly with ldexp. The coefficients
// are the Chebyshev fit of 2^r on [-1/2, 1/2] (mpmath.chebyfit, 50 digits),
// rounded to double: 2.0e-17 relative error before the rounding of the
// Horner evaluation, which adds about one ulp. The constant term is exactly
// 1, so integer arguments give exact powers of two.
//
// Arguments below -1100 are clamped there and give 0 (so does -inf, where r
// would be inf - inf); NaN gives NaN (the comparison is false for it).
// Callers pass arguments <= 0 or modestly above it (a fixed shift a little
// off a row's maximum), never above 1100.
//
// The passes evaluate one exponential per (entry, slice), and the library
// exp2 spends about 26 FP64 instructions on it, with its overflow and
// underflow cases. This form takes 18, all straight-line.
//
// Plain C++ (std::fma, std::rint, std::ldexp), so that a host build computes
// the same values bit for bit as the device build.
#pragma once

#include <cmath>

#ifdef __HIPCC__
#define RWIG_HD __host__ __device__
#else
#define RWIG_HD
#endif

namespace rwig {

RWIG_HD inline double exp2_f64(double x) {
  x = x < -1100.0 ? -1100.0 : x;
  const double k = std::rint(x);
  const double r = x - k;
  double p = 4.4558179083360645e-10;
  p = std::fma(p, r, 7.074194297288521e-09);
  p = std::fma(p, r, 1.0178057087733941e-07);
  p = std::fma(p, r, 1.3215432535912375e-06);
  p = std::fma(p, r, 1.5252733841556773e-05);
  p = std::fma(p, r, 1.5403530463724353e-04);
  p = std::fma(p, r, 1.333355814640647e-03);
  p = std::fma(p, r, 9.618129107587256e-03);
  p = std::fma(p, r, 5.5504108664821625e-02);
  p = std::fma(p, r, 2.4022650695910158e-01);
  p = std::fma(p, r, 6.931471805599453e-01);
  p = std::fma(p, r, 1.0);
  return std::ldexp(p, (int)k);
}

}  // namespace rwig
