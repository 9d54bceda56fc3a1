// oracle/flop_count.cpp -- the CPU restatement instantiated over a counting
// real type, to measure the algorithmic work of the gray M1 path per point.
//
// TEST INFRASTRUCTURE ONLY (used by tests/ and bench.py to state FLOP/point).
//
// Convention: +, -, *, / and sqrt count one FLOP each (the restatement is
// written without fused multiply-adds); exp, log, pow, atan and cbrt are
// tallied separately and not folded into the FLOP figure.  Comparisons,
// sign flips and fabs are not counted.
#include <cmath>
#include <cstdint>

namespace orc {
struct Cnt;
}

namespace {
struct Tally {
  int64_t add = 0, mul = 0, div = 0, sqrt = 0, exp = 0, log = 0, pow = 0, atan = 0,
          cbrt = 0;
};
thread_local Tally g_t;
}  // namespace

namespace orc {
struct Cnt {
  double v;
  Cnt() : v(0.) {}
  Cnt(double x) : v(x) {}  // NOLINT: implicit, mirrors double
  Cnt& operator+=(Cnt o) { ++g_t.add; v += o.v; return *this; }
  Cnt& operator-=(Cnt o) { ++g_t.add; v -= o.v; return *this; }
  Cnt& operator*=(Cnt o) { ++g_t.mul; v *= o.v; return *this; }
  Cnt& operator/=(Cnt o) { ++g_t.div; v /= o.v; return *this; }
};
inline Cnt operator+(Cnt a, Cnt b) { ++g_t.add; return Cnt(a.v + b.v); }
inline Cnt operator-(Cnt a, Cnt b) { ++g_t.add; return Cnt(a.v - b.v); }
inline Cnt operator*(Cnt a, Cnt b) { ++g_t.mul; return Cnt(a.v * b.v); }
inline Cnt operator/(Cnt a, Cnt b) { ++g_t.div; return Cnt(a.v / b.v); }
inline Cnt operator-(Cnt a) { return Cnt(-a.v); }
inline bool operator<(Cnt a, Cnt b) { return a.v < b.v; }
inline bool operator>(Cnt a, Cnt b) { return a.v > b.v; }
inline bool operator<=(Cnt a, Cnt b) { return a.v <= b.v; }
inline bool operator>=(Cnt a, Cnt b) { return a.v >= b.v; }
inline Cnt rexp(Cnt x) { ++g_t.exp; return Cnt(std::exp(x.v)); }
inline Cnt rlog(Cnt x) { ++g_t.log; return Cnt(std::log(x.v)); }
inline Cnt rpow(Cnt x, Cnt y) { ++g_t.pow; return Cnt(std::pow(x.v, y.v)); }
inline Cnt rsqrt(Cnt x) { ++g_t.sqrt; return Cnt(std::sqrt(x.v)); }
inline Cnt ratan(Cnt x) { ++g_t.atan; return Cnt(std::atan(x.v)); }
inline Cnt rcbrt(Cnt x) { ++g_t.cbrt; return Cnt(std::cbrt(x.v)); }
inline Cnt rfabs(Cnt x) { return Cnt(std::fabs(x.v)); }
inline bool rsignbit(Cnt x) { return std::signbit(x.v); }
inline double rval(Cnt x) { return x.v; }
}  // namespace orc

#include "m1_oracle.hpp"

using namespace orc;

extern "C" {

// Runs the gray path over np points (same arguments as orc_m1_opacities) with
// the counting type and writes the totals:
// counts[0..8] = add, mul, div, sqrt, exp, log, pow, atan, cbrt.
int orc_count_m1_opacities(int64_t np, int n1, const double* x1, const double* w1, int n2,
                           const double* x2, const double* w2, const double* pts,
                           const int* flags, const int* params, int stim, double* out,
                           int64_t* counts) {
  g_t = Tally{};
  Quad q1{n1, x1, w1}, q2{n2, x2, w2};
  Flags fl{flags[0], flags[1], flags[2], flags[3], flags[4]};
  Params pr{params[0], params[1], params[2], params[3], params[4], params[5], params[6], params[7]};
  Scratch<Cnt> M;
  for (int64_t i = 0; i < np; ++i) {
    Eos<Cnt> e{pts[P_NB * np + i], pts[P_T * np + i], pts[P_YP * np + i],
               pts[P_YN * np + i], pts[P_MUP * np + i], pts[P_MUN * np + i],
               pts[P_MUE * np + i], pts[P_DU * np + i], pts[P_DMEFF * np + i]};
    M1<Cnt> m;
    for (int s = 0; s < 4; ++s) {
      m.n[s] = pts[(P_N0 + s) * np + i];
      m.J[s] = pts[(P_J0 + s) * np + i];
      m.chi[s] = pts[(P_CHI0 + s) * np + i];
    }
    Distr<Cnt> d;
    distr_from_m1(m, d);
    Cnt o[20];
    m1_opacities_point(q1, q2, e, m, d, fl, pr, stim, M, o);
    for (int f = 0; f < 20; ++f) out[f * np + i] = o[f].v;
  }
  counts[0] = g_t.add;
  counts[1] = g_t.mul;
  counts[2] = g_t.div;
  counts[3] = g_t.sqrt;
  counts[4] = g_t.exp;
  counts[5] = g_t.log;
  counts[6] = g_t.pow;
  counts[7] = g_t.atan;
  counts[8] = g_t.cbrt;
  return 0;
}

// The spectral path over np points x nnu energies (arguments of
// orc_spectral_opacities, distributions reconstructed from M1); counts as
// for orc_count_m1_opacities.
int orc_count_spectral_opacities(int64_t np, int n, const double* x, const double* w,
                                 const double* pts, int64_t nnu, const double* nu,
                                 const int* flags, const int* params, int stim, double* out,
                                 int64_t* counts) {
  if (n < 1 || n > kSpecMaxNodes / 2) return 1;
  g_t = Tally{};
  Quad q{n, x, w};
  Flags fl{flags[0], flags[1], flags[2], flags[3], flags[4]};
  Params pr{params[0], params[1], params[2], params[3], params[4], params[5], params[6], params[7]};
  for (int64_t i = 0; i < np; ++i) {
    Eos<Cnt> e{pts[P_NB * np + i], pts[P_T * np + i], pts[P_YP * np + i],
               pts[P_YN * np + i], pts[P_MUP * np + i], pts[P_MUN * np + i],
               pts[P_MUE * np + i], pts[P_DU * np + i], pts[P_DMEFF * np + i]};
    M1<Cnt> m;
    for (int s = 0; s < 4; ++s) {
      m.n[s] = pts[(P_N0 + s) * np + i];
      m.J[s] = pts[(P_J0 + s) * np + i];
      m.chi[s] = pts[(P_CHI0 + s) * np + i];
    }
    Distr<Cnt> d;
    distr_from_m1(m, d);
    spectral_point(
        q, e, d, fl, pr, nnu, [&](int64_t k) { return Cnt(nu[k]); }, stim,
        [&](int64_t k, const Cnt* o) {
          for (int f = 0; f < 16; ++f) out[(f * nnu + k) * np + i] = o[f].v;
        });
  }
  counts[0] = g_t.add;
  counts[1] = g_t.mul;
  counts[2] = g_t.div;
  counts[3] = g_t.sqrt;
  counts[4] = g_t.exp;
  counts[5] = g_t.log;
  counts[6] = g_t.pow;
  counts[7] = g_t.atan;
  counts[8] = g_t.cbrt;
  return 0;
}

}  // extern "C"
