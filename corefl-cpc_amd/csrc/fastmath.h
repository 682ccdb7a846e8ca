// Natural logarithm and exponential for the transport fits, written from the
// textbook argument reductions (J.-M. Muller, Elementary Functions, ch. 11):
//
//   log x = e ln 2 + 2 atanh(s),  x = m 2^e,  m in [sqrt(1/2), sqrt(2)),
//           s = (m - 1) / (m + 1),  |s| <= 0.1716,  atanh by its odd series
//           to s^21 (the next term is below 1e-19 relative);
//   exp x = 2^n e^r,  n = rint(x / ln 2),  |r| <= ln(2) / 2,  e^r by its
//           Taylor series to r^13 (the next term is below 5e-18).
//
// ln 2 is split into a 33-bit head and a tail so that n ln 2 is exact in
// the reduction.  Both stay within a few ulp of the correctly rounded result
// for normal positive finite arguments (log) and |x| < 700 (exp), which is
// all the transport fits pass them; they cost about a third (log) and half
// (exp) of the general-purpose routines.  Host and device share the code, so
// the accuracy is checked on the CPU.
#pragma once

#include <cmath>

#ifdef __HIPCC__
#define COREFL_FM inline __host__ __device__
#else
#define COREFL_FM inline
#endif

namespace corefl {

constexpr double kFmLn2Hi = 6.93147180369123816490e-01;  // ln 2 rounded to 33 bits
constexpr double kFmLn2Lo = 1.90821492927058770002e-10;  // ln 2 - kFmLn2Hi
constexpr double kFmLog2e = 1.44269504088896338700e+00;

COREFL_FM double fm_log(double x) {
  int e;
  double m = std::frexp(x, &e);  // x = m 2^e, m in [1/2, 1)
  if (m < 0.70710678118654752440) {
    m *= 2.0;
    e -= 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double s2 = s * s;
  double p = 1.0 / 21;
  p = std::fma(p, s2, 1.0 / 19);
  p = std::fma(p, s2, 1.0 / 17);
  p = std::fma(p, s2, 1.0 / 15);
  p = std::fma(p, s2, 1.0 / 13);
  p = std::fma(p, s2, 1.0 / 11);
  p = std::fma(p, s2, 1.0 / 9);
  p = std::fma(p, s2, 1.0 / 7);
  p = std::fma(p, s2, 1.0 / 5);
  p = std::fma(p, s2, 1.0 / 3);
  const double lm = 2.0 * std::fma(s * s2, p, s);
  const double de = static_cast<double>(e);
  return std::fma(de, kFmLn2Hi, std::fma(de, kFmLn2Lo, lm));
}

COREFL_FM double fm_exp(double x) {
  const double n = std::rint(x * kFmLog2e);
  double r = std::fma(-n, kFmLn2Hi, x);
  r = std::fma(-n, kFmLn2Lo, r);
  double p = 1.0 / 6227020800.0;          // 1/13!
  p = std::fma(p, r, 1.0 / 479001600.0);  // 1/12!
  p = std::fma(p, r, 1.0 / 39916800.0);
  p = std::fma(p, r, 1.0 / 3628800.0);
  p = std::fma(p, r, 1.0 / 362880.0);
  p = std::fma(p, r, 1.0 / 40320.0);
  p = std::fma(p, r, 1.0 / 5040.0);
  p = std::fma(p, r, 1.0 / 720.0);
  p = std::fma(p, r, 1.0 / 120.0);
  p = std::fma(p, r, 1.0 / 24.0);
  p = std::fma(p, r, 1.0 / 6.0);
  p = std::fma(p, r, 0.5);
  p = std::fma(p, r, 1.0);
  p = std::fma(p, r, 1.0);
  return std::ldexp(p, static_cast<int>(n));
}

}  // namespace corefl
