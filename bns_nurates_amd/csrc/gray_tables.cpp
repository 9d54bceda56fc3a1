// gray_tables.cpp -- host construction of the resident quadrature tables.
#include "gray_tables.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

#include "gray_math.h"

namespace bns {

namespace {

using NpRec = std::array<double, NP_REC>;

// Form class of one Fermi-integral argument el - v of the NEPS bundles:
// exponential (|el - v| > 2) or rational form, and reflected (el - v > 0).
int fdi_class(double el, double v) {
  const double a = el - v;
  return (std::fabs(a) > 2.0 ? 0 : 2) + (a > 0.0 ? 1 : 0);
}

// The order of bin `bin` (see kNpBins): below |eta| = 10^-0.2 the
// half-cells grouped by the form classes of their six arguments at the
// bin's central |eta| (ties in ascending w); above, ascending w.  Measured
// on points of the 128^3 workload selected by |eta| (n = 16, 2^19 points a
// range), the grouped orders cut the launch time by 6.0% below 1e-3, 5.2%
// from 1e-3 to 0.01, 4.3% from 0.01 to 0.05 and 1.5% from 0.1 to 0.63;
// grouped orders above 0.63 made it slower (+0.6% from 1 to 10, +1.1% from
// 10 to 100), where ascending w keeps the rounds whose positron arguments
// all lie in the deep tail together.
std::vector<int> neps_order_for_bin(const std::vector<NpRec>& rec, const std::vector<int>& by_w,
                                    int bin) {
  if (bin == kNpBins - 1) return by_w;
  const double lo = bin == 0 ? std::log10(kNpBinLo) - 1.0
                             : std::log10(kNpBinLo) + (bin - 1) / kNpBinsPerDecade;
  const double eb = std::pow(10.0, lo + 0.5 / kNpBinsPerDecade);
  std::vector<std::array<int, 6>> key(rec.size());
  for (size_t k = 0; k < rec.size(); ++k) {
    const double w = rec[k][NP_W], wp = rec[k][NP_WP], ad = std::fabs(w - wp);
    key[k] = {fdi_class(eb, w), fdi_class(eb, wp), fdi_class(eb, ad),
              fdi_class(-eb, w), fdi_class(-eb, wp), fdi_class(-eb, ad)};
  }
  std::vector<int> ord = by_w;  // ties keep ascending w
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return key[a] < key[b]; });
  return ord;
}

}  // namespace

TableLayout build_gray_tables(int n1, const double* x1, const double* w1, int n2,
                              const double* x2, const double* w2, double* block) {
  TableLayout L = gray_table_layout(n1, n2);
  std::fill(block, block + L.total, 0.0);
  for (int i = 0; i < n1; ++i) {
    block[L.q1 + Q1_REC * i + Q1_X] = x1[i];
    block[L.q1 + Q1_REC * i + Q1_W] = w1[i];
    block[L.q1 + Q1_REC * i + Q1_RX] = 1.0 / x1[i];
  }
  const int n = n2, N = L.N;
  // 2D nodes a < n: x_a, a >= n: 1/x_{a-n} (m1_opacities.hpp:783-787); the
  // quadrature weight of the second half carries the Jacobian 1/x^2
  // (integration.hpp:651-673).  The weight stored includes y^2 so that the
  // per-cell factor nu^2 nubar^2 / T^4 is the product of two node weights.
  std::vector<double> y(N), W(N);
  for (int i = 0; i < n; ++i) {
    y[i] = c::PairScale * x2[i];
    y[n + i] = c::PairScale / x2[i];
    W[i] = w2[i];
    W[n + i] = w2[i] / (x2[i] * x2[i]);
  }
  for (int a = 0; a < N; ++a) {
    block[L.nd + ND_REC * a + ND_RX] = 1.0 / x2[a < n ? a : a - n];
    block[L.nd + ND_REC * a + ND_WQ] = W[a] * y[a] * y[a];
    block[L.nd + ND_REC * a + ND_X] = x2[a < n ? a : a - n];
    block[L.nd + ND_REC * a + ND_INV] = a < n ? 0.0 : 1.0;
  }
  // pair cells a <= b, sorted by S = y_a + y_b so that the 64 lanes of a
  // round evaluate Fermi integrals at nearby arguments (same branch)
  std::vector<std::pair<int, int>> cells;
  for (int a = 0; a < N; ++a)
    for (int b = a; b < N; ++b) cells.emplace_back(a, b);
  std::stable_sort(cells.begin(), cells.end(), [&](const auto& u, const auto& v) {
    return y[u.first] + y[u.second] < y[v.first] + y[v.second];
  });
  for (int c = 0; c < L.P; ++c) {
    int a = cells[c].first, b = cells[c].second;
    double S = y[a] + y[b];
    double* r = block + L.cl + CL_REC * c;
    r[CL_AB] = double(a * 256 + b);
    r[CL_S] = S;
    // pair: detailed balance e^{(w+w')/T} and 1/(1 - e^{(w+w')/T})
    // (kernel_pair.hpp:145, 193-194)
    r[CL_DB] = safe_exp(S);
    r[CL_PAUX] = 1.0 / (1.0 - safe_exp(S));
    // BRT06: x = (w+w')/(2T), K1(x)/x e^{-+x} (kernel_brem_BRT06.hpp:52-59)
    double xb = 0.5 * S;
    double k1x = bessk1_fit(xb) / xb;
    r[CL_K1EM] = k1x * safe_exp(-xb);
    r[CL_K1AB] = k1x * safe_exp(xb);
    // HR98: x = (w+w')/T, clamped as BremKernelS does (kernel_brem_HR98.hpp:77)
    double xh = S > c::BremMin ? S : c::BremMin;
    r[CL_X11] = std::pow(xh, 1.1);
    r[CL_XM4] = std::pow(xh, -4.0);
    r[CL_X093] = std::pow(xh, 0.93);
    r[CL_X12] = std::pow(xh, 1.2);
    r[CL_X28] = std::pow(xh, 2.8);
    r[CL_LSD] = std::log((xh * xh + c::FourPiSquared) * xh /
                         (c::FourPiSquared * (1.0 - safe_exp(-xh))));
    r[CL_EMX] = safe_exp(-S);
  }
  // NEPS half-cells (i, j, half): w = nu/T, w' = nubar/T with
  // nu = t x_i (1 -+ x_j)/2 (half 0) or t (1 -+ x_j)/(2 x_i) (half 1),
  // t = 4 * 2.182 T (m1_opacities.hpp:910-912, 982-984, 1167).  The number
  // weight folds w_i w_j, the Jacobian x_i or 1/x_i^3
  // (integration.hpp:730-756), nu^2 nubar^2 / T^4, and the factor
  // (1/2) t_neps^2 / t_pair^2 = 8 that puts both 2D sums on one scale.
  // The half-cells are stored once per degeneracy bin, each copy in the
  // order neps_order_for_bin picks (ascending w = min(w, w'), or grouped by
  // the Fermi-integral forms the bin's arguments take), so that the 64
  // lanes of a round take the same branches.
  const double tn = 4.0 * c::PairScale;
  std::vector<NpRec> rec(size_t(2) * n * n);
  int h = 0;
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) {
      for (int half = 0; half < 2; ++half, ++h) {
        double xi = x2[i], xj = x2[j];
        double w, wp, jac;
        if (half == 0) {
          w = 0.5 * tn * xi * (1.0 - xj);
          wp = 0.5 * tn * xi * (1.0 + xj);
          jac = xi;
        } else {
          w = 0.5 * tn * (1.0 - xj) / xi;
          wp = 0.5 * tn * (1.0 + xj) / xi;
          jac = 1.0 / (xi * xi * xi);
        }
        double* r = rec[h].data();
        r[NP_W] = w;
        r[NP_WP] = wp;
        r[NP_RXI] = 1.0 / xi;
        r[NP_EF] = neps_exp_func(wp - w);
        r[NP_EFX] = neps_exp_func(w - wp);
        r[NP_INVDEN] = 1.0 / (w * w * wp * wp);
        r[NP_WQ] = 8.0 * w2[i] * w2[j] * jac * (w * w) * (wp * wp);
        r[NP_SIGN] = std::signbit(wp - w) ? 1.0 : -1.0;
        r[NP_XI] = xi;
        r[NP_OMX] = 1.0 - xj;
        r[NP_OPX] = 1.0 + xj;
        r[NP_HALF] = half;
        r[NP_LW] = std::log(w);
        r[NP_LWP] = std::log(wp);
        r[NP_EW] = std::exp(-w);
        r[NP_EWP] = std::exp(-wp);
        r[NP_EAD] = std::exp(-std::fabs(w - wp));
      }
    }
  }
  // the base order: ascending w
  std::vector<int> by_w(rec.size());
  std::iota(by_w.begin(), by_w.end(), 0);
  std::stable_sort(by_w.begin(), by_w.end(),
                   [&](int a, int b) { return rec[a][NP_W] < rec[b][NP_W]; });
  // the records in ascending w, and per degeneracy bin (see kNpBins) the
  // order of visit as indices into them
  std::vector<int> pos(rec.size());
  for (size_t k = 0; k < by_w.size(); ++k) {
    std::copy(rec[by_w[k]].begin(), rec[by_w[k]].end(), block + L.np + NP_REC * k);
    pos[by_w[k]] = static_cast<int>(k);
  }
  std::vector<uint16_t> perm(size_t(kNpBins) * L.Q);
  for (int bin = 0; bin < kNpBins; ++bin) {
    const std::vector<int> ord = neps_order_for_bin(rec, by_w, bin);
    for (size_t k = 0; k < ord.size(); ++k)
      perm[size_t(bin) * L.Q + k] = static_cast<uint16_t>(pos[ord[k]]);
  }
  std::memcpy(block + L.npp, perm.data(), perm.size() * sizeof(uint16_t));
  return L;
}

double gamma_stirling_36() {
  const double x = 0.6 + 3.0;
  return std::sqrt(6.283185307179586 / x) * std::pow(x / 2.718281828459045235360287471352, x);
}

}  // namespace bns
