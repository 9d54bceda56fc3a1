// gray_tables.h -- point-independent geometry of the gray M1 quadratures,
// built once per quadrature order on the host and kept resident in HBM.
//
// The energies of the 2D pair/bremsstrahlung grid are y_a = t x~_a / T with
// t = 2.182 T (m1_opacities.hpp:774-787, 1087-1088), x~ = x_i or 1/x_i; the
// NEPS grid uses t = 4 * 2.182 T and the change of variables of
// m1_opacities.hpp:900-1054.  In units of T both grids are independent of the
// thermodynamic point, so every quantity that depends on a cell alone (sums of
// energies, detailed-balance exponentials, the BRT06 K1 factor, the HR98
// powers of x, NEPSExpFunc, quadrature weights) is tabulated here once, as
// array-of-records so that one lane reads a cell with wide loads.
#ifndef BNS_NURATES_AMD_GRAY_TABLES_H_
#define BNS_NURATES_AMD_GRAY_TABLES_H_

#include <cstdint>

namespace bns {

constexpr int kMaxQuad = 32;  // quadrature order n <= 32 (2n = 64 nodes)

// 1D quadrature record: node, weight, 1 / node correctly rounded
enum { Q1_X = 0, Q1_W, Q1_RX, Q1_REC = 3 };
// 2D node record: 1 / x correctly rounded, number weight times y^2 (y =
// nu / T), the quadrature node x and whether the energy is t x (0) or t / x (1)
enum { ND_RX = 0, ND_WQ, ND_X, ND_INV, ND_REC = 4 };
// pair/brem cell record (a <= b), cells sorted by S = y_a + y_b
enum {
  CL_AB = 0,  // a * 256 + b
  CL_S,       // y_a + y_b
  CL_DB,      // SafeExp(S)                 pair detailed balance
  CL_PAUX,    // 1 / (1 - SafeExp(S))        pair prefactor
  CL_K1EM,    // K1(S/2)/(S/2) e^{-S/2}      BRT06 emission
  CL_K1AB,    // K1(S/2)/(S/2) e^{+S/2}      BRT06 absorption
  CL_X11,     // S^1.1                       HR98 powers of x = S
  CL_XM4,     // S^-4
  CL_X093,    // S^0.93
  CL_X12,     // S^1.2
  CL_X28,     // S^2.8
  CL_LSD,     // log((S^2 + 4 pi^2) S / (4 pi^2 (1 - e^-S)))
  CL_EMX,     // SafeExp(-S)                 HR98 detailed balance
  CL_REC = 16
};
// NEPS half-cell record (i, j, half)
enum {
  NP_W = 0,   // nu / T
  NP_WP,      // nubar / T
  NP_RXI,     // 1 / x_i correctly rounded
  NP_EF,      // NEPSExpFunc(w' - w)
  NP_EFX,     // NEPSExpFunc(w - w')
  NP_INVDEN,  // 1 / (w^2 w'^2)
  NP_WQ,      // 8 w_i w_j (x_i or 1/x_i^3) w^2 w'^2
  NP_SIGN,    // 2 signbit(w' - w) - 1
  NP_XI,      // x_i
  NP_OMX,     // 1 - x_j
  NP_OPX,     // 1 + x_j
  NP_HALF,    // 0: nu = t x_i (1 -+ x_j)/2, 1: nu = t (1 -+ x_j)/(2 x_i)
  NP_LW,      // log(w): log(nu) = log(T) + log(w)
  NP_LWP,     // log(w')
  NP_EW,      // e^-w, e^-w' and e^-|w - w'|: with e^-|eta| of the point the
  NP_EWP,     // exponentials of the Fermi integrals at -|eta| - w, -|eta| - w'
  NP_EAD,     // and -|eta| - |w - w'| (see the NEPS loop)
  NP_REC = 17
};

struct TableLayout {
  int n1, n2;   // 1D and 2D quadrature orders
  int N, P, Q;  // 2D nodes (2 n2), pair cells (N(N+1)/2), NEPS half-cells (2 n2^2)
  int q1, nd, cl, np;  // offsets (in doubles) of the four record arrays
  int npp;             // offset (in doubles) of the half-cell orders: uint16 [kNpBins][Q]
  int total;           // doubles in the block
};

// The NEPS half-cells are visited in one order per range of the electron
// degeneracy |eta| ("bin"; uint16 record indices at npp): a round of 64 lanes evaluates
// a Fermi-integral bundle in its exponential form, its rational form or
// both, as its lanes' arguments eta - v, -eta - v fall, and an order that
// groups the half-cells by the forms they take at that |eta| has fewer
// rounds that pay for both.  Bin 0: |eta| < 1e-3 (and NaN); bins 1..7:
// 2.5 per decade up to 10^-0.2; bin 8: 10^-0.2 and above.
constexpr int kNpBins = 9;
constexpr double kNpBinLo = 1e-3, kNpBinHi = 0.63095734448019325, kNpBinsPerDecade = 2.5;

#if defined(__HIPCC__)
__host__ __device__
#endif
inline int np_bin_of(double abs_eta_log10) {
  // abs_eta_log10 = log10(|eta|) for kNpBinLo <= |eta| < kNpBinHi
  int b = 1 + static_cast<int>(kNpBinsPerDecade * (abs_eta_log10 + 3.0));
  return b < 1 ? 1 : (b > kNpBins - 2 ? kNpBins - 2 : b);
}

// record counts and offsets for quadrature orders n1 (1D) and n2 (2D); a
// constant expression, so that a kernel instance for fixed orders gets its
// offsets and loop bounds at compile time
#if defined(__HIPCC__)
__host__ __device__
#endif
constexpr TableLayout gray_table_layout(int n1, int n2) {
  TableLayout L{};
  L.n1 = n1;
  L.n2 = n2;
  L.N = 2 * n2;
  L.P = L.N * (L.N + 1) / 2;
  L.Q = 2 * n2 * n2;
  L.q1 = 0;
  L.nd = L.q1 + Q1_REC * n1 + (n1 & 1) * Q1_REC;
  L.cl = L.nd + ND_REC * L.N;
  L.np = L.cl + CL_REC * L.P;
  L.npp = (L.np + NP_REC * L.Q + 1) / 2 * 2;  // even: 16-byte aligned orders
  L.total = L.npp + (kNpBins * L.Q + 3) / 4;
  return L;
}

// Fills `block` (layout.total doubles) for the quadratures (x1, w1: n1 nodes)
// and (x2, w2: n2 nodes) on [0, 1].
TableLayout build_gray_tables(int n1, const double* x1, const double* w1, int n2,
                              const double* x2, const double* w2, double* block);

// GammaStirling (functions.hpp:4460-4470) at c_f + 3 = 3.6, the Gamma
// factor of the optically thin distribution normalisation
double gamma_stirling_36();

}  // namespace bns

#endif  // BNS_NURATES_AMD_GRAY_TABLES_H_
