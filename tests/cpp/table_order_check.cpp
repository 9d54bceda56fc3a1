// Host check of the NEPS half-cell visit orders of the resident tables
// (bns_nurates_amd/csrc/gray_tables.{h,cpp}).  For quadrature orders
// 4..32 it builds the table block and checks that
//   - the NEPS records are stored in ascending w (= min(w, w'));
//   - every degeneracy bin's order is a permutation of the records;
//   - the last bin (|eta| >= kNpBinHi) visits them in stored order;
//   - the other bins visit them grouped by the form classes of their six
//     Fermi-integral arguments at the bin's central |eta| (the class key is
//     non-decreasing along the order, ties in ascending w);
//   - the bin index of np_bin_of covers [kNpBinLo, kNpBinHi) with the
//     ranges the host builder uses;
//   - the order block starts 16-byte aligned (an even double offset).
// Prints "table order failures 0" and exits 0 when all hold.
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../bns_nurates_amd/csrc/gray_tables.h"

using namespace bns;

namespace {

// Gauss-Legendre nodes and weights on [0, 1] (Newton on P_n)
void gauss_legendre01(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; ++k) {
        const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
      const double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    x[i] = 0.5 * (1.0 - z);
    w[i] = 1.0 / ((1.0 - z * z) * dp * dp);
  }
}

int fdi_class_of(double el, double v) {
  const double a = el - v;
  return (std::fabs(a) > 2.0 ? 0 : 2) + (a > 0.0 ? 1 : 0);
}

}  // namespace

int main() {
  long failures = 0;
  auto fail = [&](const char* what, int n, int bin, long k) {
    if (failures < 20) std::printf("FAIL n=%d bin=%d k=%ld: %s\n", n, bin, k, what);
    ++failures;
  };
  for (int n : {4, 5, 8, 10, 12, 16, 20, 32}) {
    std::vector<double> x, w;
    gauss_legendre01(n, x, w);
    const TableLayout L = gray_table_layout(n, n);
    std::vector<double> block(L.total);
    build_gray_tables(n, x.data(), w.data(), n, x.data(), w.data(), block.data());
    if (L.npp % 2 != 0) fail("order block offset not even", n, -1, L.npp);
    const double* rec = block.data() + L.np;
    for (int k = 1; k < L.Q; ++k)
      if (!(rec[NP_REC * (k - 1) + NP_W] <= rec[NP_REC * k + NP_W]))
        fail("records not in ascending w", n, -1, k);
    std::vector<uint16_t> ord(static_cast<size_t>(kNpBins) * L.Q);
    std::memcpy(ord.data(), block.data() + L.npp, ord.size() * sizeof(uint16_t));
    for (int bin = 0; bin < kNpBins; ++bin) {
      const uint16_t* o = ord.data() + static_cast<size_t>(bin) * L.Q;
      std::vector<int> seen(L.Q, 0);
      for (int k = 0; k < L.Q; ++k) {
        if (o[k] >= L.Q) {
          fail("index out of range", n, bin, k);
          continue;
        }
        ++seen[o[k]];
      }
      for (int k = 0; k < L.Q; ++k)
        if (seen[k] != 1) fail("not a permutation", n, bin, k);
      if (bin == kNpBins - 1) {
        for (int k = 0; k < L.Q; ++k)
          if (o[k] != k) fail("last bin not in stored order", n, bin, k);
        continue;
      }
      // the bin's central |eta| (bin 0: 10^-0.2 decade below kNpBinLo)
      const double lo = bin == 0 ? std::log10(kNpBinLo) - 1.0
                                 : std::log10(kNpBinLo) + (bin - 1) / kNpBinsPerDecade;
      const double eb = std::pow(10.0, lo + 0.5 / kNpBinsPerDecade);
      auto key = [&](int r) {
        const double wv = rec[NP_REC * r + NP_W], wpv = rec[NP_REC * r + NP_WP];
        const double ad = std::fabs(wv - wpv);
        return std::array<int, 7>{fdi_class_of(eb, wv), fdi_class_of(eb, wpv),
                                  fdi_class_of(eb, ad), fdi_class_of(-eb, wv),
                                  fdi_class_of(-eb, wpv), fdi_class_of(-eb, ad), r};
      };
      for (int k = 1; k < L.Q; ++k)
        if (!(key(o[k - 1]) < key(o[k]))) fail("not grouped by form class", n, bin, k);
    }
  }
  // bin index: |eta| in [kNpBinLo, kNpBinHi) -> 1 .. kNpBins - 2, monotone
  int prev = 1;
  for (int i = 0; i <= 6000; ++i) {
    const double a = kNpBinLo * std::pow(kNpBinHi / kNpBinLo, i / 6000.0);
    if (!(a < kNpBinHi)) break;
    const int b = np_bin_of(std::log10(a));
    if (b < 1 || b > kNpBins - 2 || b < prev) fail("bin index", -1, b, i);
    prev = b;
  }
  if (prev != kNpBins - 2) fail("top range does not reach the last graded bin", -1, prev, 0);
  std::printf("table order failures %ld\n", failures);
  return failures == 0 ? 0 : 1;
}
