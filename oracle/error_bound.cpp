// oracle/error_bound.cpp -- the CPU restatement instantiated over a
// magnitude-propagating real type: a first-order running error bound
// (Wilkinson's running error analysis) for each output of the gray M1 path.
//
// TEST INFRASTRUCTURE ONLY: the parity tests use it to tell rounding noise
// from a real difference at outputs formed by cancellation.
//
// Every value carries m >= |v|, a bound such that any IEEE double evaluation
// of the same formulas (any operation order, with or without fused
// multiply-adds, any faithfully rounded libm) differs from the exact result
// by at most about k u m, u = 2^-53, k a small multiple of the longest
// dependency chain.  Propagation (first order):
//   a +- b : m = m_a + m_b + |a +- b|
//   a * b  : m = |a| m_b + |b| m_a + |a b|
//   a / b  : m = m_a / |b| + |a| m_b / b^2 + |a / b|
//   f(x)   : m = |f'(x)| m_x + |f(x)|            (exp, log, sqrt, atan, cbrt)
//   x^y    : m = |y x^(y-1)| m_x + |x^y log x| m_y + |x^y|
// Every operation that can underflow (*, /, the functions) also adds
// DBL_MIN = 2^-1022: its absolute error below the normal range is at most
// 2^-1075 = u DBL_MIN (IEEE gradual underflow), which in the units of u m is
// DBL_MIN.  Sums of doubles are exact in the subnormal range.
// Constants and inputs start with m = |v|.  Outputs whose m is many orders of
// magnitude above |v| are cancellation-dominated: their digits beyond
// u m / |v| are not determined by the formulas in double precision.
#include <cfloat>
#include <cmath>
#include <cstdint>

#include <omp.h>

namespace orc {
// products of magnitudes where a zero meets an infinity contribute nothing
// (an exact zero carries no error), never NaN
inline double mm(double a, double b) { return (a == 0. || b == 0.) ? 0. : a * b; }
inline double md(double a, double b) { return (a == 0. || std::isinf(b)) ? 0. : a / b; }
constexpr double kUnder = DBL_MIN;  // underflow term, see the header
struct Mag {
  double v, m;
  Mag() : v(0.), m(0.) {}
  Mag(double x) : v(x), m(std::fabs(x)) {}  // NOLINT: implicit, mirrors double
  Mag(double x, double mm) : v(x), m(mm) {}
  Mag& operator+=(Mag o) { *this = Mag(v + o.v, m + o.m + std::fabs(v + o.v)); return *this; }
  Mag& operator-=(Mag o) { *this = Mag(v - o.v, m + o.m + std::fabs(v - o.v)); return *this; }
  Mag& operator*=(Mag o) {
    double r = v * o.v;
    *this = Mag(r, mm(std::fabs(v), o.m) + mm(std::fabs(o.v), m) + std::fabs(r) + kUnder);
    return *this;
  }
  Mag& operator/=(Mag o) {
    double r = v / o.v, ab = std::fabs(o.v);
    *this = Mag(r, md(m, ab) + md(mm(std::fabs(v), o.m), mm(ab, ab)) + std::fabs(r) + kUnder);
    return *this;
  }
};
inline Mag operator+(Mag a, Mag b) { return a += b; }
inline Mag operator-(Mag a, Mag b) { return a -= b; }
inline Mag operator*(Mag a, Mag b) { return a *= b; }
inline Mag operator/(Mag a, Mag b) { return a /= b; }
inline Mag operator-(Mag a) { return Mag(-a.v, a.m); }
inline bool operator<(Mag a, Mag b) { return a.v < b.v; }
inline bool operator>(Mag a, Mag b) { return a.v > b.v; }
inline bool operator<=(Mag a, Mag b) { return a.v <= b.v; }
inline bool operator>=(Mag a, Mag b) { return a.v >= b.v; }
inline Mag rexp(Mag x) {
  double e = std::exp(x.v);
  return Mag(e, mm(e, x.m) + e + kUnder);
}
inline Mag rlog(Mag x) {
  double l = std::log(x.v);
  return Mag(l, md(x.m, std::fabs(x.v)) + std::fabs(l));
}
inline Mag rsqrt(Mag x) {
  double s = std::sqrt(x.v);
  return Mag(s, (s > 0. ? 0.5 * x.m / s : 0.) + s);
}
inline Mag ratan(Mag x) {
  double a = std::atan(x.v);
  return Mag(a, x.m / (1. + x.v * x.v) + std::fabs(a));
}
inline Mag rcbrt(Mag x) {
  double c = std::cbrt(x.v);
  return Mag(c, (c != 0. ? x.m / (3. * c * c) : 0.) + std::fabs(c));
}
inline Mag rpow(Mag x, Mag y) {
  double p = std::pow(x.v, y.v);
  double dx = (x.v != 0.) ? std::fabs(y.v * p / x.v) : 0.;
  double dy = (x.v > 0.) ? std::fabs(p * std::log(x.v)) : 0.;
  return Mag(p, mm(dx, x.m) + mm(dy, y.m) + std::fabs(p) + kUnder);
}
inline Mag rfabs(Mag x) { return Mag(std::fabs(x.v), x.m); }
inline bool rsignbit(Mag x) { return std::signbit(x.v); }
inline double rval(Mag x) { return x.v; }
}  // namespace orc

#include "m1_oracle.hpp"

using namespace orc;

extern "C" {

// the GP19 table of this library (orc_set_gp19_table of the double build)
void orc_bound_set_gp19_table(int nnb, const double* nb, const double* lognb, int nye,
                              const double* ye, int nT, const double* T, int nw, const double* w,
                              const double* logw, const double* data) {
  orc::gp19_assign(nnb, nb, lognb, nye, ye, nT, T, nw, w, logw, data);
}

// Same arguments as orc_m1_opacities (distr NULL: reconstruct from M1);
// writes the outputs to out and their magnitudes m to mag ([20][np] each).
int orc_bound_m1_opacities(int64_t np, int n1, const double* x1, const double* w1, int n2,
                           const double* x2, const double* w2, const double* pts,
                           const double* distr, const int* flags, const int* params, int stim,
                           double* out, double* mag, int nthreads) {
  Quad q1{n1, x1, w1}, q2{n2, x2, w2};
  Flags fl{flags[0], flags[1], flags[2], flags[3], flags[4]};
  Params pr{params[0], params[1], params[2], params[3], params[4], params[5], params[6], params[7]};
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
  {
    Scratch<Mag> M;
#pragma omp for schedule(dynamic, 8)
    for (int64_t i = 0; i < np; ++i) {
      Eos<Mag> e{pts[P_NB * np + i], pts[P_T * np + i], pts[P_YP * np + i],
                 pts[P_YN * np + i], pts[P_MUP * np + i], pts[P_MUN * np + i],
                 pts[P_MUE * np + i], pts[P_DU * np + i], pts[P_DMEFF * np + i]};
      M1<Mag> m;
      for (int s = 0; s < 4; ++s) {
        m.n[s] = pts[(P_N0 + s) * np + i];
        m.J[s] = pts[(P_J0 + s) * np + i];
        m.chi[s] = pts[(P_CHI0 + s) * np + i];
      }
      Distr<Mag> d;
      if (distr) {
        for (int s = 0; s < 4; ++s) {
          d.w_t[s] = distr[(D_WT + s) * np + i];
          d.temp_t[s] = distr[(D_TT + s) * np + i];
          d.eta_t[s] = distr[(D_ETAT + s) * np + i];
          d.w_f[s] = distr[(D_WF + s) * np + i];
          d.temp_f[s] = distr[(D_TF + s) * np + i];
          d.c_f[s] = distr[(D_CF + s) * np + i];
          d.beta_f[s] = distr[(D_BF + s) * np + i];
        }
      } else {
        distr_from_m1(m, d);
      }
      Mag o[20];
      m1_opacities_point(q1, q2, e, m, d, fl, pr, stim, M, o);
      for (int f = 0; f < 20; ++f) {
        out[f * np + i] = o[f].v;
        mag[f * np + i] = o[f].m;
      }
    }
  }
  return 0;
}

// Same arguments as orc_spectral_opacities; writes the outputs to out and
// their magnitudes m to mag ([16][nnu][np] each).
int orc_bound_spectral_opacities(int64_t np, int n, const double* x, const double* w,
                                 const double* pts, const double* distr, int64_t nnu,
                                 const double* nu, int64_t nu_ld, const int* flags,
                                 const int* params, int stim, double* out, double* mag,
                                 int nthreads) {
  Quad q{n, x, w};
  Flags fl{flags[0], flags[1], flags[2], flags[3], flags[4]};
  Params pr{params[0], params[1], params[2], params[3], params[4], params[5], params[6], params[7]};
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  if (n > kSpecMaxNodes / 2) return 1;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
  for (int64_t i = 0; i < np; ++i) {
    Eos<Mag> eos{pts[P_NB * np + i], pts[P_T * np + i], pts[P_YP * np + i],
                 pts[P_YN * np + i], pts[P_MUP * np + i], pts[P_MUN * np + i],
                 pts[P_MUE * np + i], pts[P_DU * np + i], pts[P_DMEFF * np + i]};
    Distr<Mag> d;
    if (distr) {
      for (int s = 0; s < 4; ++s) {
        d.w_t[s] = distr[(D_WT + s) * np + i];
        d.temp_t[s] = distr[(D_TT + s) * np + i];
        d.eta_t[s] = distr[(D_ETAT + s) * np + i];
        d.w_f[s] = distr[(D_WF + s) * np + i];
        d.temp_f[s] = distr[(D_TF + s) * np + i];
        d.c_f[s] = distr[(D_CF + s) * np + i];
        d.beta_f[s] = distr[(D_BF + s) * np + i];
      }
    } else {
      M1<Mag> m;
      for (int s = 0; s < 4; ++s) {
        m.n[s] = pts[(P_N0 + s) * np + i];
        m.J[s] = pts[(P_J0 + s) * np + i];
        m.chi[s] = pts[(P_CHI0 + s) * np + i];
      }
      distr_from_m1(m, d);
    }
    spectral_point(
        q, eos, d, fl, pr, nnu, [&](int64_t e) { return Mag(nu[nu_ld ? e * nu_ld + i : e]); },
        stim, [&](int64_t e, const Mag* o) {
          for (int f = 0; f < 16; ++f) {
            out[(f * nnu + e) * np + i] = o[f].v;
            mag[(f * nnu + e) * np + i] = o[f].m;
          }
        });
  }
  return 0;
}

// CalculateDistrParamsFromM1 over the point block: the 28 distribution
// parameters and their magnitudes ([28][np] each, layout of orc_distr_from_m1).
void orc_bound_distr_from_m1(int64_t np, const double* pts, double* distr, double* mag) {
  for (int64_t i = 0; i < np; ++i) {
    M1<Mag> m;
    for (int s = 0; s < 4; ++s) {
      m.n[s] = pts[(P_N0 + s) * np + i];
      m.J[s] = pts[(P_J0 + s) * np + i];
      m.chi[s] = pts[(P_CHI0 + s) * np + i];
    }
    Distr<Mag> d;
    distr_from_m1(m, d);
    const Mag* f[7] = {d.w_t, d.temp_t, d.eta_t, d.w_f, d.temp_f, d.c_f, d.beta_f};
    const int off[7] = {D_WT, D_TT, D_ETAT, D_WF, D_TF, D_CF, D_BF};
    for (int k = 0; k < 7; ++k)
      for (int s = 0; s < 4; ++s) {
        distr[(off[k] + s) * np + i] = f[k][s].v;
        mag[(off[k] + s) * np + i] = f[k][s].m;
      }
  }
}

}  // extern "C"
