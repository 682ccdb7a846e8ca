// Accuracy of corefl::fm_log / fm_exp against the C library on the host:
// prints the largest error in units of the last place over a sweep of
// arguments (log: 1e-300 .. 1e300 and a dense sweep of [0.5, 2]; exp:
// -700 .. 700).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "fastmath.h"

static double ulp_err(double got, double ref) {
  if (got == ref) return 0.0;
  const double u = std::nextafter(std::fabs(ref), INFINITY) - std::fabs(ref);
  return std::fabs(got - ref) / u;
}

int main() {
  double worst_log = 0, worst_exp = 0, at_log = 0, at_exp = 0;
  uint64_t st = 0x9E3779B97F4A7C15ull;
  auto rnd = [&]() {
    st ^= st << 13;
    st ^= st >> 7;
    st ^= st << 17;
    return (st >> 11) * (1.0 / 9007199254740992.0);
  };
  for (int i = 0; i < 2000000; ++i) {
    const double x = std::pow(10.0, -300.0 + 600.0 * rnd());
    const double e = ulp_err(corefl::fm_log(x), std::log(x));
    if (e > worst_log) worst_log = e, at_log = x;
    const double y = 0.5 + 1.5 * rnd();
    const double e2 = ulp_err(corefl::fm_log(y), std::log(y));
    if (e2 > worst_log) worst_log = e2, at_log = y;
    const double z = -700.0 + 1400.0 * rnd();
    const double e3 = ulp_err(corefl::fm_exp(z), std::exp(z));
    if (e3 > worst_exp) worst_exp = e3, at_exp = z;
  }
  std::printf("%.3f %.17g %.3f %.17g\n", worst_log, at_log, worst_exp, at_exp);
  return 0;
}
