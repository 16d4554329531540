// Host check of rwig::exp2_f64 against long-double exp2l: prints the largest
// error in units of the last place over a sweep of arguments, and the values
// at the special arguments, one "name value" line each.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <initializer_list>

#include "../rwig_amd/csrc/exp2_f64.hpp"

static double ulp_err(double got, long double want) {
  const double w = (double)want;
  if (w == 0.0) return got == 0.0 ? 0.0 : 1e300;
  const double ulp = std::nextafter(std::fabs(w), INFINITY) - std::fabs(w);
  return (double)(std::fabs((long double)got - want) / ulp);
}

int main() {
  double worst = 0.0, worst_x = 0.0;
  uint64_t s = 88172645463325252ull;
  auto next = [&] {  // xorshift64
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return s;
  };
  // arguments down to -1022 (normal results) and up to 60, uniformly, plus
  // the neighbourhood of each integer
  for (int n = 0; n < 4000000; ++n) {
    const double u = (next() >> 11) * (1.0 / 9007199254740992.0);
    const double x = -1022.0 + u * 1082.0;
    const double e = ulp_err(rwig::exp2_f64(x), exp2l((long double)x));
    if (e > worst) worst = e, worst_x = x;
  }
  for (int k = -1022; k <= 60; ++k)
    for (double d : {-0.5, -0.25, -1e-9, 0.0, 1e-9, 0.25, 0.4999999}) {
      const double x = k + d;
      const double e = ulp_err(rwig::exp2_f64(x), exp2l((long double)x));
      if (e > worst) worst = e, worst_x = x;
    }
  printf("max_ulp %.6g\n", worst);
  printf("max_ulp_at %.17g\n", worst_x);
  printf("at_0 %.17g\n", rwig::exp2_f64(0.0));
  printf("at_-1 %.17g\n", rwig::exp2_f64(-1.0));
  printf("at_-1074 %.17g\n", rwig::exp2_f64(-1074.0));
  printf("at_-1075.5 %.17g\n", rwig::exp2_f64(-1075.5));
  printf("at_-2000 %.17g\n", rwig::exp2_f64(-2000.0));
  printf("at_-inf %.17g\n", rwig::exp2_f64(-INFINITY));
  printf("at_nan_is_nan %d\n", std::isnan(rwig::exp2_f64(NAN)) ? 1 : 0);
  printf("at_1023_5 %.17g\n", rwig::exp2_f64(1023.5));
  // subnormal results: relative error grows as the precision shrinks, so
  // the absolute error is measured in units of the smallest subnormal
  double worst_sub = 0.0;
  for (int n = 0; n < 200000; ++n) {
    const double x = -1074.0 + (next() >> 11) * (1.0 / 9007199254740992.0) * 52.0;
    const long double w = exp2l((long double)x);
    const double e = (double)(std::fabs((long double)rwig::exp2_f64(x) - w) / 4.9406564584124654e-324L);
    if (e > worst_sub) worst_sub = e;
  }
  printf("max_subnormal_err_in_min_subnormals %.6g\n", worst_sub);
  return 0;
}
