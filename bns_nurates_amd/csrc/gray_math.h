// gray_math.h -- constants and special functions of the gray M1 opacity path
// for CDNA4 (gfx950) device code and the host-side table builders.
//
// The physical constants are evaluated from the fundamental ones with the
// same expressions as constants.hpp:31-244 so the derived values agree with
// the reference to the last bit.  The Fermi-Dirac coefficient tables are the
// published piecewise minimax rationals of Fukushima (2015) as tabulated in
// functions.hpp; they are data that must be identical for parity.
#ifndef BNS_NURATES_AMD_GRAY_MATH_H_
#define BNS_NURATES_AMD_GRAY_MATH_H_

#include <hip/hip_runtime.h>

#include <cmath>

#define BNS_HD __host__ __device__ __forceinline__

namespace bns {

// Wave-uniform per-point values kept in LDS and read where they are used:
// volatile keeps the compiler from hoisting the reads out of the cell loops
// (held in registers through them they spill), and the LDS address space
// keeps them LDS reads.  Through a generic pointer a volatile access is left
// in the generic address space: a flat load that waits for its completion
// before the next instruction.
template <class T>
using lds_cv = const volatile __attribute__((address_space(3))) T;
template <class T>
__device__ __forceinline__ lds_cv<T>* lds_volatile(T* p) {
  return (lds_cv<T>*)p;
}

// x as a value the compiler cannot see through: work formed from it inside a
// rarely taken branch stays in that branch instead of being hoisted to where
// every lane pays for it.
__device__ __forceinline__ double opaque(double x) {
  __asm__ volatile("" : "+v"(x));
  return x;
}

// ---------------------------------------------------------------------------
// Constants (constants.hpp)
// ---------------------------------------------------------------------------
namespace c {
constexpr double MeV = 1.6021766341182763e+8;
constexpr double Me = 0.51099895069;
constexpr double Mn = 939.56542194;
constexpr double Mp = 938.27208943;
constexpr double Q = 1.29333251;
constexpr double Mb = 1.6737747126723517e-24;
constexpr double Pi = 3.14159265358979323846264338328;
constexpr double H = 4.135667696e-21;
constexpr double Hbar = 0.5 * H / Pi;
constexpr double Clight = 2.99792458e+17;
constexpr double Gf0 = 1.1663787e-11;
constexpr double Gf = Gf0 * ((Hbar * Clight) * (Hbar * Clight) * (Hbar * Clight));
constexpr double SaturationN = 0.15e+18;
constexpr double Ga = 1.23, Gv = 1., Gs = 0., SinThW2 = 0.2325;
constexpr double Hnv = -0.5, Hna = -0.5 * Ga, Hpv = 0.5 - 2. * SinThW2, Hpa = 0.5 * Ga;
constexpr double BetaConst = ((Clight * Clight * Clight) * ((Gf0 * Hbar) * (Gf0 * Hbar))) *
                             ((Gv * Gv) + 3. * (Ga * Ga)) / Pi;
constexpr double IsoEF = 4.785390000313652 * (Hbar * Hbar) / Mb * MeV;
constexpr double IsoConst =
    2. * Pi * (Gf * Gf) / Hbar / ((H * Clight) * (H * Clight) * (H * Clight));
constexpr double IsoC0p = IsoConst * ((Hpv * Hpv) + 3. * (Hpa * Hpa));
constexpr double IsoC1p = IsoConst * ((Hpv * Hpv) - (Hpa * Hpa));
constexpr double IsoC0n = IsoConst * ((Hnv * Hnv) + 3. * (Hna * Hna));
constexpr double IsoC1n = IsoConst * ((Hnv * Hnv) - (Hna * Hna));
constexpr double BremMin = 1.0e-10;
constexpr double BremAux1 = 0.25 * (H * H) * MeV / (Pi * Pi);
constexpr double BremCa = -0.63;
constexpr double BremConst = (BremCa * BremCa) * (Gf * Gf) / Hbar;
constexpr double BremC4BRT06 = 0.1409912109375;
constexpr double PairPhiC = 1.55e-12 / Pi;
constexpr double PairA1E = 1. + 2. * SinThW2, PairA2E = 2. * SinThW2;
constexpr double PairA1X = -1. + 2. * SinThW2, PairA2X = 2. * SinThW2;
constexpr double NEPSConst =
    2. * (Gf0 * Gf0) * Clight * ((Hbar * Clight) * (Hbar * Clight)) / (3. * Pi);
constexpr double NEPSBPlus = (2. * SinThW2 + 1.) * (2. * SinThW2 + 1.);
constexpr double NEPSBMinus = (2. * SinThW2 - 1.) * (2. * SinThW2 - 1.);
constexpr double NEPSBZero = 4. * (SinThW2 * SinThW2);
constexpr double ExpLim = 700.;
constexpr double HC6FourPiSq = ((H * Clight) * (H * Clight) * (H * Clight) * (H * Clight) *
                                (H * Clight) * (H * Clight)) /
                               (16. * (Pi * Pi));
constexpr double WMEScale = MeV / (Mb * (Clight * Clight));
constexpr double FourPiHc3 = (4. * Pi) / ((H * Clight) * (H * Clight) * (H * Clight));
constexpr double FourPiHc3Sqr = FourPiHc3 * FourPiHc3;
constexpr double MnGrams = Mn * MeV / (Clight * Clight);
constexpr double MpGrams = Mp * MeV / (Clight * Clight);
constexpr double MAvgGrams = MeV / (Clight * Clight) * 938.918532994116257276238085018813;
constexpr double SqrtPi = 1.77245385090551602729816748334106;
constexpr double PiSquared = 9.86960440108935861883449099987572;
constexpr double FourPiSquared = 39.4784176043574344753379639995029;
constexpr double Pi2OneEighth = 1.15383506784998943054096521314981;
constexpr double PiHalfToFiveHalves = 3.09242868139914350627854469835251;
constexpr double ThresholdN = 1e-21, ThresholdJ = 1e-25;
constexpr double E2 = 7.38905609893065023;  // e^2, functions.hpp:23
// energy scale of the 2D pair/brem grid and of the NEPS grid, in units of T
// (m1_opacities.hpp:1087-1092, 1167): t = 2.182 T and 4 * 2.182 T
constexpr double PairScale = 0.5 * 4.364;
}  // namespace c

// ---------------------------------------------------------------------------
// Elementary helpers
// ---------------------------------------------------------------------------
// Products that must be rounded on their own (never fused into a following
// add), where the reference's two roundings decide a threshold or an exact
// zero: the energies of the quadrature grids and 3 chi - 1.
BNS_HD double mul_rn(double a, double b) {
#pragma clang fp contract(off)
  return a * b;
}
// e^x for x in [-746, 710] (callers bound x; e^x is 0 below -745.14 and
// inf above 709.79), on the device without the range checks of the library
// exponential: Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, with ln2
// split into 42 + 53 bits so that k ln2_hi is exact; e^r by a degree-11
// polynomial, a Chebyshev fit of e^r on [-ln2/2, ln2/2] whose relative error
// with the coefficients rounded to double is 1.7e-17; 2^k applied by ldexp.
// Within one ulp of e^x (correctly rounded for 93% of 6e7 samples).
BNS_HD double exp_bounded(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double k = rint(x * 0x1.71547652b82fep+0);
  double r = fma(-k, 0x1.62e42fefa3800p-1, x);
  r = fma(-k, 0x1.ef35793c76730p-45, r);
  double p = 0x1.af631d0059becp-26;
  p = fma(r, p, 0x1.28b4057f44145p-22);
  p = fma(r, p, 0x1.71ddf5749d126p-19);
  p = fma(r, p, 0x1.a01991ac8730ap-16);
  p = fma(r, p, 0x1.a01a01b14378fp-13);
  p = fma(r, p, 0x1.6c16c187fbe02p-10);
  p = fma(r, p, 0x1.111111110f225p-7);
  p = fma(r, p, 0x1.555555554f0cfp-5);
  p = fma(r, p, 0x1.555555555555ap-3);
  p = fma(r, p, 0x1.0000000000011p-1);
  p = fma(r, p, 1.0);
  p = fma(r, p, 1.0);
  return ldexp(p, static_cast<int>(k));
#else
  return std::exp(x);
#endif
}

// SafeExp (functions.hpp:66-75): NaN maps to the lower clamp
BNS_HD double safe_exp(double x) {
  double lo = (x > -c::ExpLim) ? x : -c::ExpLim;
  return exp_bounded((lo < c::ExpLim) ? lo : c::ExpLim);
}

// Division p/q for normal-range operands whose quotient is normal (no
// scaling for overflow or denormals): the hardware reciprocal (better than
// 2^-18), one Newton step, and one residual correction of the quotient,
// which rounds as p / q does for all but near-tie quotients.  6 VALU
// operations against 10 for the IEEE sequence.
BNS_HD double div_normal(double p, double q) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(q);
  double e = fma(-q, r, 1.0);
  r = fma(r, e, r);
  double y = p * r;
  double res = fma(-q, y, p);
  return fma(r, res, y);
#else
  return p / q;
#endif
}
// a / b given rb = 1 / b correctly rounded: q = a rb is within an ulp of
// a / b and the residual step gives the correctly rounded quotient
// (Markstein's theorem).  Used where the reference's rounding decides a
// threshold, with rb formed once per point or stored in a table.
BNS_HD double div_by_recip(double a, double b, double rb) {
  double q = a * rb;
  double r = fma(-q, b, a);
  return fma(r, rb, q);
}

// log(x) for positive, finite, normal x (callers guarantee it; the HR98
// bremsstrahlung logarithms and neutrino energies), without the library's
// special-case handling: x = 2^k m with m in [sqrt(1/2), sqrt(2)), f = m - 1
// (exact), s = f / (2 + f), log m = 2 atanh(s) = f - (f^2/2 - s (f^2/2 + R))
// with R = 2 sum_{j=1..10} s^{2j} / (2j + 1) (the series' first omitted term
// is below 1e-18 relative at |s| <= 0.1716), and k ln 2 with ln 2 split so
// that k ln2_hi is exact.  The same operations on the host and the device
// (contraction off), so that the host check covers the device code: within
// one ulp of log(x) (tests/cpp/log_bounded_check.cpp).
BNS_HD double log_bounded(double x) {
#pragma clang fp contract(off)
  int k;
  double m = frexp(x, &k);  // [0.5, 1)
  if (m < 0x1.6a09e667f3bcdp-1) {
    m += m;
    --k;
  }
  const double f = m - 1.;
  const double s = div_normal(f, 2. + f);
  const double z = s * s;
  double r = 0x1.8618618618618p-4;  // 2/21
  r = fma(z, r, 0x1.af286bca1af28p-4);  // 2/19
  r = fma(z, r, 0x1.e1e1e1e1e1e1ep-4);  // 2/17
  r = fma(z, r, 0x1.1111111111111p-3);  // 2/15
  r = fma(z, r, 0x1.3b13b13b13b14p-3);  // 2/13
  r = fma(z, r, 0x1.745d1745d1746p-3);  // 2/11
  r = fma(z, r, 0x1.c71c71c71c71cp-3);  // 2/9
  r = fma(z, r, 0x1.2492492492492p-2);  // 2/7
  r = fma(z, r, 0x1.999999999999ap-2);  // 2/5
  r = fma(z, r, 0x1.5555555555555p-1);  // 2/3
  r = r * z;
  const double hfsq = 0.5 * (f * f);
  const double dk = static_cast<double>(k);
  const double ln2_hi = 0x1.62e42ff000000p-1, ln2_lo = -0x1.718432a1b0e26p-35;
  const double corr = fma(s, hfsq + r, dk * ln2_lo);
  return fma(dk, ln2_hi, f - (hfsq - corr));
}

// log(x) for any x: the bounded form where x is positive, finite and normal,
// the library logarithm otherwise (zero, subnormals, inf and NaN keep its
// special values; a wavefront none of whose lanes needs it skips it)
BNS_HD double log_pos(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return (x >= 0x1p-1022 && x <= 0x1.fffffffffffffp+1023) ? log_bounded(x) : log(x);
#else
  return std::log(x);
#endif
}

// e^x for any x: the bounded form for |x| < 746 (where it also returns the
// library's underflow to 0 and overflow to inf), the library exponential for
// larger, infinite or NaN arguments
BNS_HD double exp_any(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return fabs(x) < 746. ? exp_bounded(x) : exp(x);
#else
  return std::exp(x);
#endif
}

// FermiDistr (functions.hpp:4376-4395) given the already-formed argument
// (e - mu) / T.  Written without a branch: 1/(e^a + 1) = e^-a/(e^-a + 1);
// the denominator lies in [1, 2] and the numerator in [e^-700, 1].
BNS_HD double fermi_of_arg(double arg) {
  double t = safe_exp(-fabs(arg));
  double num = (arg > 0.) ? t : 1.;
  return div_normal(num, t + 1.);
}

// NEPSExpFunc (functions.hpp:4401-4412)
BNS_HD double neps_exp_func(double x) {
  double ex = safe_exp(-fabs(x));
  bool neg = std::signbit(x);
  return ((neg ? 1. : 0.) - (neg ? 0. : 1.) * ex) / (1. - ex + (x == 0. ? 1. : 0.));
}

// NEPSExpFunc(x) and NEPSExpFunc(x2) for x2 = -x formed as a separate
// difference (w - w' and w' - w): one exponential, the same bits as two calls
BNS_HD void neps_exp_func_pair(double x, double x2, double& f, double& f2) {
  double ex = safe_exp(-fabs(x));
  double den = 1. - ex + (x == 0. ? 1. : 0.);
  f = (std::signbit(x) ? 1. : -ex) / den;
  f2 = (std::signbit(x2) ? 1. : -ex) / den;
}

// bessi1 / bessk1: the Abramowitz & Stegun polynomial fits of
// functions.hpp:99-180 (used by BRT06 bremsstrahlung)
BNS_HD double bessi1_fit(double x) {
  double ax = fabs(x), ans, y;
  if (ax < 3.75) {
    y = x / 3.75;
    y = y * y;
    ans = ax * (0.5 + y * (0.87890594 + y * (0.51498869 + y * (0.15084934 +
                 y * (0.2658733e-1 + y * (0.301532e-2 + y * 0.32411e-3))))));
  } else {
    y = 3.75 / ax;
    ans = 0.2282967e-1 + y * (-0.2895312e-1 + y * (0.1787654e-1 - y * 0.420059e-2));
    ans = 0.39894228 + y * (-0.3988024e-1 + y * (-0.362018e-2 + y * (0.163801e-2 +
          y * (-0.1031555e-1 + y * ans))));
    ans *= (exp(ax) / sqrt(ax));
  }
  return x < 0. ? -ans : ans;
}
BNS_HD double bessk1_fit(double x) {
  if (x <= 2.) {
    double y = x * x / 4.;
    return (log(x / 2.) * bessi1_fit(x)) +
           (1.0 / x) * (1.0 + y * (0.15443144 + y * (-0.67278579 + y * (-0.18156897 +
            y * (-0.1919402e-1 + y * (-0.110404e-2 + y * (-0.4686e-4)))))));
  }
  double y = 2. / x;
  return (exp(-x) / sqrt(x)) *
         (1.25331414 + y * (0.23498619 + y * (-0.3655620e-1 + y * (0.1504268e-1 +
          y * (-0.780353e-2 + y * (0.325614e-2 + y * (-0.68245e-3)))))));
}

// ---------------------------------------------------------------------------
// Complete Fermi-Dirac integrals F_k, k = 0..5 (functions.hpp:1551-2837).
// Coefficients in Horner order (highest power first), signs folded in:
//   x = -|y| < -2:  t = e^x e^2,  F = e^x (a0 - e^x Pa(t) / Qa(t)),
//                   Qa monic;
//   -2 <= x <= 0:   s = -x/2, t = 1 - s,  F = Pb(t) / Qb(s),
//                   Qb leading coefficient `qb_lead` (+-1);
//   y > 0:          exact reflection polynomial.
// ---------------------------------------------------------------------------
template <int K>
struct Fdi;
template <>
struct Fdi<0> {
  static constexpr int na = 5, nqa = 4, nb = 9, nqb = 7;
  static constexpr double a0 = 1., qb_lead = 1.;
  static constexpr double pa[5] = {0.00200096064827815813, 6.9167792879948140,
                                   357.623326425354522, 5222.0667923565138,
                                   22696.2126132366633};
  static constexpr double qa[4] = {71.072178562726798, 1611.36476693109675,
                                   14539.5980679273792, 45392.4252264733267};
  static constexpr double pb[9] = {0.000915454570009894267, -0.0734798777056723512,
                                   -0.450072959113928254, -0.441998676614933572,
                                   3.78181326142271599, 10.5181677709577503,
                                   0.377783268730614356, 23.7193942338278703,
                                   159.601717762460980};
  static constexpr double qb[8] = {8.1762315442738975, 35.456591489081173, 96.77898293084927,
                                   206.21640678892182, 310.2713981221035, 315.2592651624449,
                                   284.26032127745967, 0.};
};
template <>
struct Fdi<1> {
  static constexpr int na = 5, nqa = 4, nb = 9, nqb = 8;
  static constexpr double a0 = 1., qb_lead = -1.;
  static constexpr double pa[5] = {0.00397641173774375092, 5.9897442965804548,
                                   322.901386168881348, 4915.92700908746777,
                                   22189.1070807945062};
  static constexpr double qa[4] = {88.376214553692756, 2389.06277237306633,
                                   25002.3197546553836, 88756.428323178025};
  static constexpr double pb[9] = {-0.0792966701498222697, -0.403173132925886253,
                                   0.199701180197912643, 7.09659390228556164,
                                   20.1369115558099802, 17.9918985363509694,
                                   56.6537141912783024, 251.392824471576922,
                                   145.488167182330098};
  static constexpr double qb[8] = {-7.1995819520154718, -25.678454878692950,
                                   -53.72117554363975, -61.57766112137513, 27.2746245830016,
                                   252.1367051536344, 374.1806357435014, 606.0757707716040};
};
template <>
struct Fdi<2> {
  static constexpr int na = 4, nqa = 3, nb = 9, nqb = 8;
  static constexpr double a0 = 2., qb_lead = -1.;
  static constexpr double pa[5] = {0.0161890243763741414, 8.5861610217850095,
                                   273.085756700981399, 1914.06748184935743, 0.};
  static constexpr double qa[4] = {72.929152915475392, 1399.35442210906621,
                                   7656.2699273974454, 0.};
  static constexpr double pb[9] = {-0.326188299771397236, -2.14077421411719935,
                                   -2.68393549333878715, 24.0376482128898634,
                                   112.951038040682055, 172.881855215582924,
                                   222.606134197895041, 1299.85460914884154,
                                   2711.49678259128843};
  static constexpr double qb[8] = {-3.9142702096919080, 8.395768655115050, 122.54595216479181,
                                   531.62378035996132, 1428.0589853413436, 2541.7823512372631,
                                   3038.7689794575778, 2517.1726659917047};
};
template <>
struct Fdi<3> {
  static constexpr int na = 4, nqa = 3, nb = 8, nqb = 7;
  static constexpr double a0 = 6., qb_lead = 1.;
  static constexpr double pa[5] = {0.0410982603688952131, 19.0856927562699544,
                                   664.28706260743472, 5121.6401850302408, 0.};
  static constexpr double qa[4] = {92.376788603062645, 2136.54222460571183,
                                   13657.7071600806539, 0.};
  static constexpr double pb[9] = {1.59407954898394322, 19.9559488532742796,
                                   113.373708671587772, 354.630774329461644,
                                   653.359212389160499, 1260.13125873282465,
                                   4323.07526636309661, 7881.24597452900838, 0.};
  static constexpr double qb[8] = {16.468882811659000, 112.60906419590854, 459.86413596901097,
                                   1259.0724833462608, 2393.9995533270879, 2972.7443644211129,
                                   2570.7250703533430, 0.};
};
template <>
struct Fdi<4> {
  static constexpr int na = 4, nqa = 3, nb = 8, nqb = 7;
  static constexpr double a0 = 24., qb_lead = 1.;
  static constexpr double pa[5] = {0.121876197098273914, 54.659507299984584,
                                   2120.56302902849207, 18247.2542465629199, 0.};
  static constexpr double qa[4] = {117.071576489684022, 3261.01909656925521,
                                   24329.6723287505678, 0.};
  static constexpr double pb[9] = {6.67374311450268357, 77.4108960876508190,
                                   418.662172475927160, 1268.40017257978070,
                                   2393.94262931609398, 5431.08357152245897,
                                   19031.4783357798975, 33093.9102025608137, 0.};
  static constexpr double qb[8] = {15.216887271751039, 105.86020755838874, 448.99020896813485,
                                   1278.0109577275445, 2500.5977847175497, 3236.2237702166948,
                                   2645.4885670047153, 0.};
};
template <>
struct Fdi<5> {
  static constexpr int na = 4, nqa = 3, nb = 8, nqb = 7;
  static constexpr double a0 = 120., qb_lead = 1.;
  static constexpr double pa[5] = {0.413800735538960261, 190.753830813273698,
                                   8368.4990332049831, 81190.938912603315, 0.};
  static constexpr double qa[4] = {148.484432990224213, 4977.68099709243407,
                                   43301.8340867217726, 0.};
  static constexpr double pb[9] = {23.6526046298891619, 272.164427980501432,
                                   1461.45267097859337, 4445.76333033698006,
                                   9274.54751848368696, 26923.5693307648389,
                                   96307.2005742063042, 159651.547840031111, 0.};
  static constexpr double qb[8] = {13.491911254479298, 91.84661231161838, 397.7596246691212,
                                   1176.7202478443275, 2403.0532924519756, 3244.5527999477567,
                                   2522.7839609396783, 0.};
};

// Numerator and denominator of the two rational forms of F_k on x = -|y|:
// x < -2: F = e^x (a0 - e^x p(t) / q(t)), t = e^{x+2};
// -2 <= x <= 0: F = p(t) / q(s), s = -x/2, t = 1 - s.
template <int K>
BNS_HD void fdi_exp_pq(double t, double& p, double& q) {
  using C = Fdi<K>;
  p = C::pa[0];
#pragma unroll
  for (int i = 1; i < C::na; ++i) p = fma(t, p, C::pa[i]);
  q = t + C::qa[0];
#pragma unroll
  for (int i = 1; i < C::nqa; ++i) q = fma(t, q, C::qa[i]);
}
template <int K>
BNS_HD void fdi_rat_pq(double s, double t, double& p, double& q) {
  using C = Fdi<K>;
  p = C::pb[0];
#pragma unroll
  for (int i = 1; i < C::nb; ++i) p = fma(t, p, C::pb[i]);
  q = C::qb_lead;
#pragma unroll
  for (int i = 0; i < C::nqb; ++i) q = fma(s, q, C::qb[i]);
}
// exact polynomial added on reflection to y > 0 (functions.hpp:1606-2834)
template <int K>
BNS_HD double fdi_reflect(double fd, double y) {
  if constexpr (K == 0) {
    return fd + y;
  } else if constexpr (K == 1) {
    return -fd + fma(0.5 * y, y, 1.64493406684822644);
  } else if constexpr (K == 2) {
    return fma(y, fma(0.333333333333333333 * y, y, 3.28986813369645287), fd);
  } else if constexpr (K == 3) {
    double y2 = y * y;
    return fma(y2, fma(y2, 0.25, 4.93480220054467931), 11.3643939539669510) - fd;
  } else if constexpr (K == 4) {
    double y2 = y * y;
    return fma(y, fma(y2, fma(y2, 0.2, 6.57973626739290575), 45.4575758158678040), fd);
  } else {
    double y2 = y * y;
    return fma(y2, fma(y2, fma(y2, 0.166666666666666667, 8.22467033424113218),
                       113.643939539669510),
               236.532261911384425) -
           fd;
  }
}

// One order.  Only the polynomials are evaluated under the branch; the
// division and the combination F = A p/q + a0 e^x are shared (exp form: p
// carries e^x and A = -e^x; rational form: A = 1, e^x = 0; NaN: A = e^x = 0
// with finite p and q), so that a wavefront whose lanes take both forms pays
// for one division per order.
template <int K>
struct FdiUnroll {
  template <int KHI>
  BNS_HD static void eval(bool ebr, double ex, double A, double t, double s, double* F) {
    double p, q;
    if (ebr) {
      fdi_exp_pq<K>(t, p, q);
      p *= ex;
    } else {
      fdi_rat_pq<K>(s, t, p, q);
    }
    double fd = fma(A, div_normal(p, q), ex * Fdi<K>::a0);
#if defined(__HIP_DEVICE_COMPILE__)
    // keeps this order's division here: sunk to the uses, the divisions of
    // all orders would keep every (p, q) pair live at once
    asm volatile("" : "+v"(fd));
#endif
    F[K] = fd;
    if constexpr (K < KHI) FdiUnroll<K + 1>::template eval<KHI>(ebr, ex, A, t, s, F);
  }
};

// Deep tail of the exponential form: for x < kFdiDeep the correction
// e^{2x} p/q is below 2^-54 e^x a0 for every order (|p/q| <= 1.9 as t -> 0,
// a0 >= 1, so the bound holds from x = -36.7 on), and the fused
// F = fma(-e^x, e^x p/q, e^x a0) rounds to e^x a0: the same bits without the
// rational.
constexpr double kFdiDeep = -40.;
template <int K>
struct FdiDeepUnroll {
  template <int KHI>
  BNS_HD static void eval(double ex, double* F) {
    F[K] = ex * Fdi<K>::a0;
    if constexpr (K < KHI) FdiDeepUnroll<K + 1>::template eval<KHI>(ex, F);
  }
};

// The reflection to y > 0 of every order of a bundle, under one branch: the
// lanes that do not reflect are masked off rather than given a select per
// order, and a wavefront none of whose lanes reflects skips the polynomials
// (for eta >= 0 every positron argument of the NES kernels is <= 0, for
// eta < 0 every electron one).
template <int K>
struct FdiReflect {
  template <int KHI>
  BNS_HD static void apply(double y, double* F) {
    F[K] = fdi_reflect<K>(F[K], y);
    if constexpr (K < KHI) FdiReflect<K + 1>::template apply<KHI>(y, F);
  }
};

// F_k(y) for k = KLO..KHI at one argument, sharing the branch decision and
// the exponential (F[k] receives F_k; F must hold KHI + 1 entries).
// Inlined: with MachineLICM disabled at build time the coefficient
// materializations stay next to their use inside the cell loops.  (A
// batched reciprocal of the five denominators saves divisions but keeps
// more values live; at three waves per SIMD the extra spills cost more.)
// With DEEP, a wavefront whose active lanes all lie in the deep tail skips
// the rationals and the divisions (a wave-uniform branch, same results);
// callers enable it where whole wavefronts commonly fall in the tail, since
// the test itself costs a few instructions per bundle.
template <int KLO, int KHI = 5, bool DEEP = false>
BNS_HD void fdi_bundle(double y, double* F, bool try_deep = true) {
  double x = -fabs(y);
#if defined(__HIP_DEVICE_COMPILE__)
  if (DEEP && try_deep && __all(x < kFdiDeep)) {
    FdiDeepUnroll<KLO>::template eval<KHI>(exp_bounded(fmax(x, -746.)), F);
    if (y > 0.) FdiReflect<KLO>::template apply<KHI>(y, F);
    return;
  }
#endif
  bool ebr = x < -2.;
  bool rbr = !ebr && (x <= 0.);  // NaN arguments give 0, as the reference
  double ex = 0., A = rbr ? 1. : 0., t, s;
  if (ebr) {
    ex = exp_bounded(fmax(x, -746.));  // x < -2 here, never NaN
    A = -ex;
    t = ex * c::E2;
    s = 0.;
  } else {
    s = rbr ? -0.5 * x : 0.;
    t = 1. - s;
  }
  FdiUnroll<KLO>::template eval<KHI>(ebr, ex, A, t, s, F);
  if (y > 0.) FdiReflect<KLO>::template apply<KHI>(y, F);
}

// The bundle at an argument y <= 0 whose exponential e^y the caller supplies
// (formed otherwise than by exp_bounded, within a few ulps of it): the same
// forms, branch decisions and operations as fdi_bundle after its
// exponential; no reflection (y <= 0).
template <int KLO, int KHI = 5, bool DEEP = false>
BNS_HD void fdi_bundle_ex(double y, double ex_in, double* F) {
  const double x = y;
#if defined(__HIP_DEVICE_COMPILE__)
  if (DEEP && __all(x < kFdiDeep)) {
    FdiDeepUnroll<KLO>::template eval<KHI>(ex_in, F);
    return;
  }
#endif
  bool ebr = x < -2.;
  bool rbr = !ebr && (x <= 0.);  // NaN arguments give 0, as the reference
  double ex = 0., A = rbr ? 1. : 0., t, s;
  if (ebr) {
    ex = ex_in;
    A = -ex;
    t = ex * c::E2;
    s = 0.;
  } else {
    s = rbr ? -0.5 * x : 0.;
    t = 1. - s;
  }
  FdiUnroll<KLO>::template eval<KHI>(ebr, ex, A, t, s, F);
}

template <int K>
BNS_HD double fdi(double y) {
  double F[6];
  fdi_bundle<K, K>(y, F);
  return F[K];
}

}  // namespace bns

#endif  // BNS_NURATES_AMD_GRAY_MATH_H_
