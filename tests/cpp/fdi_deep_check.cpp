// Host check of the deep-tail form of the Fermi integrals (gray_math.h):
// below kFdiDeep the exponential form F = e^x (a0 - e^x p/q) must round to
// e^x a0 for every order, on both sides of the reflection.  Samples the
// boundary region densely and the rest of the tail at random; prints the
// number of mismatching bit patterns and exits non-zero if there are any.
// The margin of the bound (it holds from x = -36.7 on) is checked too: at
// x = -36 the identity must fail, or the test would not see a wrong bound.
#include <cstdio>
#include <cstring>
#include <random>

#include "../../bns_nurates_amd/csrc/gray_math.h"

using namespace bns;

template <int K>
double deep_form(double y) {
  const double x = -fabs(y);
  const double fd = exp_bounded(fmax(x, -746.)) * Fdi<K>::a0;
  return y > 0. ? fdi_reflect<K>(fd, y) : fd;
}

template <int K>
long mismatches(double lo, double hi, long n, std::mt19937_64& g) {
  std::uniform_real_distribution<double> u(lo, hi);
  long bad = 0;
  for (long i = 0; i < n; ++i) {
    const double x = (i & 1) ? u(g) : hi - double(i) * 1e-7;
    for (double y : {x, -x}) {
      const double a = fdi<K>(y), b = deep_form<K>(y);
      if (std::memcmp(&a, &b, sizeof a) != 0) ++bad;
    }
  }
  return bad;
}

long all_orders(double lo, double hi, long n, std::mt19937_64& g) {
  return mismatches<0>(lo, hi, n, g) + mismatches<1>(lo, hi, n, g) + mismatches<2>(lo, hi, n, g) +
         mismatches<3>(lo, hi, n, g) + mismatches<4>(lo, hi, n, g) + mismatches<5>(lo, hi, n, g);
}

int main() {
  std::mt19937_64 g(12345);
  const long tail = all_orders(-746.5, kFdiDeep, 400000, g);
  const long outside = all_orders(-36.2, -36.0, 1000, g);
  std::printf("deep tail mismatches %ld; at x = -36: %ld\n", tail, outside);
  return (tail == 0 && outside > 0) ? 0 : 1;
}
