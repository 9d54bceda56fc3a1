// Host check of log_bounded (gray_math.h) against the x87 extended-precision
// logarithm: log-uniform samples over the normal range, samples near 1 (where
// log x is small and f = x - 1 carries it), and every power of two.  Prints
// the largest error in units in the last place of the double result, the
// fraction correctly rounded, and exits non-zero if any error reaches one ulp.
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <random>

#include "../../bns_nurates_amd/csrc/gray_math.h"

namespace {

double ulp_error(double x, double& worst_x, double& worst, long& rounded) {
  const long double ref = logl(static_cast<long double>(x));
  const double got = bns::log_bounded(x);
  const double nearest = static_cast<double>(ref);
  if (got == nearest) ++rounded;
  const double r = static_cast<double>(ref);
  const double u = r == 0. ? DBL_TRUE_MIN : std::nextafter(std::fabs(r), INFINITY) - std::fabs(r);
  const double e = static_cast<double>(std::fabs(static_cast<long double>(got) - ref) / u);
  if (e > worst) {
    worst = e;
    worst_x = x;
  }
  return e;
}

}  // namespace

int main() {
  std::mt19937_64 g(20261204);
  std::uniform_real_distribution<double> ue(-1021.9 * 0.6931471805599453, 1023.9 * 0.6931471805599453);
  std::uniform_real_distribution<double> un(-0.3, 0.42);
  double worst = 0., worst_x = 0.;
  long rounded = 0, n = 0;
  for (long i = 0; i < 20000000; ++i, ++n) ulp_error(std::exp(ue(g)), worst_x, worst, rounded);
  for (long i = 0; i < 5000000; ++i, ++n) ulp_error(1. + un(g), worst_x, worst, rounded);
  for (long i = 0; i < 2000000; ++i, ++n) {
    // within a few thousand ulps of 1, where log x ~ x - 1
    const double x = 1. + (static_cast<double>(i) - 1e6) * 0x1p-52;
    ulp_error(x, worst_x, worst, rounded);
  }
  for (int e = -1022; e <= 1023; ++e, ++n) ulp_error(std::ldexp(1., e), worst_x, worst, rounded);
  std::printf("log_bounded: %ld samples, max error %.3f ulp at x = %a, correctly rounded %.4f\n",
              n, worst, worst_x, double(rounded) / double(n));
  return worst < 1. ? 0 : 1;
}
