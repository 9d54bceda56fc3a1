// point_physics.h -- per-point device helpers shared by the gray and the
// spectral opacity kernels: wavefront utilities, the distribution function
// (TotalNuF and its reconstruction from M1 moments), weak-magnetism and
// nucleon-degeneracy factors of the beta and isoenergetic processes, and the
// per-point constants of the HR98 bremsstrahlung fit.
#ifndef BNS_NURATES_AMD_POINT_PHYSICS_H_
#define BNS_NURATES_AMD_POINT_PHYSICS_H_

#include <hip/hip_runtime.h>

#include <cstdint>

#include "gray_math.h"

namespace bns {

namespace {

constexpr int kWave = 64;
// slots of 1 / T and log(T) in a kernel's 64-double per-point LDS block
// (the inputs occupy slots 0..48)
constexpr int kInvT = 63, kLogT = 62;

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Give XCD x the contiguous range of points
// [x q + min(x, r), ...) of length q + (x < r), q = G / 8, r = G % 8, so that
// the 8 points sharing a 64-byte line of a field row are read and written
// through one L2.  A bijection of [0, G) for any grid size G.
constexpr int kXcds = 8;
__device__ __forceinline__ int64_t xcd_point(int64_t b, int64_t G) {
  int64_t x = b % kXcds, slot = b / kXcds, q = G / kXcds, r = G % kXcds;
  return x * q + (x < r ? x : r) + slot;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Distribution parameters of one species, in the form the kernel evaluates
// TotalNuF (distribution.hpp:298-315):
//   f(w) = wt / (e^{(w - mut)/Tt} + 1) + wf bf w^cf e^{-w/Tf}
struct Spec {
  double wt, mut, invTt, wf, bf, cf, invTf;
  double lbf;  // log(bf), the mixed-precision form folds it into the exponent
};

__device__ __forceinline__ double nuf(double w, double lw, const Spec& d) {
  double arg = (w - d.mut) * d.invTt;
  // the exponent is below 0.6 log(w) + log(bf) (c_f = 0.6 from M1; a caller's
  // c_f beyond 709 / log(w) gives inf, as exp does); below -746 e^x is 0
  double thin = d.bf * exp_bounded(fmax(fma(d.cf, lw, -w * d.invTf), -746.));
  return fma(d.wt, fermi_of_arg(arg), d.wf * thin);
}

// A distribution value f and 1 - f.
struct GOmg {
  double g, omg;
};

// e^x with FP32 accuracy over the whole FP64 range: x log2(e) = k + r with
// k integral and |r| <= 1/2, 2^r by the hardware FP32 exponential, 2^k
// applied exactly in FP64.  NaN maps to the lower clamp (e^x = 0), as
// SafeExp does.
__device__ __forceinline__ double exp_mixed(double x) {
  x = fmin(fmax(x, -1100.), 709.);
  const double u = x * 1.4426950408889634;
  const double k = rint(u);
  const float e2 = __builtin_amdgcn_exp2f(static_cast<float>(u - k));
  return ldexp(static_cast<double>(e2), static_cast<int>(k));
}

// The mixed-precision TotalNuF: the exponent arguments are formed in FP64
// (they can be hundreds, and w - mut cancels near the Fermi surface), the
// exponentials, reciprocal and logarithm carry FP32 accuracy (relative error
// ~1e-7 in f, no cancellation downstream) over the FP64 range, and 1 - f is
// formed from 1 - FD = FD(-arg) without the cancellation of 1 - f at
// degenerate energies.
__device__ __forceinline__ GOmg nuf_mixed(double w, float lw, const Spec& d) {
  const double arg = (w - d.mut) * d.invTt;
  const double t = exp_mixed(fmax(-fabs(arg), -c::ExpLim));  // SafeExp's clamp
  const double r = static_cast<double>(__builtin_amdgcn_rcpf(static_cast<float>(t + 1.)));
  const double fd = (arg > 0. ? t : 1.) * r;  // 1 / (e^arg + 1)
  const double fc = (arg > 0. ? 1. : t) * r;  // 1 - fd
  const double thin = exp_mixed(fma(d.cf, static_cast<double>(lw), d.lbf - w * d.invTf));
  GOmg o;
  o.g = fma(d.wt, fd, d.wf * thin);
  o.omg = fma(d.wt, fc, 1. - d.wt) - d.wf * thin;
  return o;
}

// TotalNuF at w (and 1 - f) in the precision of the kernel variant
template <bool MIX>
__device__ __forceinline__ GOmg dist(double w, const Spec& d) {
  GOmg o;
  if constexpr (MIX) {
    o = nuf_mixed(w, logf(static_cast<float>(w)), d);
  } else {
    o.g = nuf(w, log_pos(w), d);
    o.omg = 1. - o.g;
  }
  return o;
}

__device__ __forceinline__ Spec pick(const Spec* D, int s) {
  Spec r = D[0];
  if (s == 1) r = D[1];
  if (s == 2) r = D[2];
  if (s == 3) r = D[3];
  return r;
}

// NucFrmFac (weak_magnetism.hpp:45-104)
__device__ __forceinline__ void form_factors(double E, int reac, double& cv, double& ca,
                                             double& F2) {
  const double lamp = 1.793, lamn = -1.913;
  // every denominator below is >= 1
  double e = E * c::WMEScale;
  double tau = div_normal(0.5 * (e * e), 1. + e);
  double eta = div_normal(1., fma(5.6, tau, 1.));
  double g = fma(4.97, tau, 1.);
  double G = div_normal(1., g * g);
  double itau = div_normal(1., 1. + tau);
  double Fp1 = fma(tau, 1. + lamp, 1.) * G * itau;
  double Fp2 = lamp * G * itau;
  double Fn1 = tau * lamn * (1. - eta) * G * itau;
  double Fn2 = lamn * fma(tau, eta, 1.) * G * itau;
  double c3 = fma(3.53, tau, 1.);
  double ic3 = div_normal(1., c3 * c3);
  const double hw = 0.5 - 2. * c::SinThW2;
  if (reac == 1) {
    cv = hw * Fp1 - 0.5 * Fn1;
    ca = 0.5 * (c::Ga - c::Gs) * ic3;
    F2 = hw * Fp2 - 0.5 * Fn2;
  } else if (reac == 2) {
    cv = hw * Fn1 - 0.5 * Fp1;
    ca = -0.5 * (c::Ga + c::Gs) * ic3;
    F2 = hw * Fn2 - 0.5 * Fp2;
  } else {
    cv = Fp1 - Fn1;
    ca = c::Ga * ic3;
    F2 = Fp2 - Fn2;
  }
}

// WMScatt R0 - R1/3 combination of IsoScattNucleon (formulas_beta_iso.hpp:
// 411-451, weak_magnetism.hpp:165-210) for one nucleon.
__device__ __forceinline__ double iso_leg(double omega, int reac, bool wm) {
  double c0 = reac == 1 ? c::IsoC0p : c::IsoC0n;
  double c1 = reac == 1 ? c::IsoC1p : c::IsoC1n;
  if (!wm) return c0 - c1 / 3.;
  double cv, ca, F2;
  form_factors(omega, reac, cv, ca, F2);
  double hv = reac == 1 ? c::Hpv : c::Hnv, ha = reac == 1 ? c::Hpa : c::Hna;
  double h0 = hv * hv + 3. * (ha * ha), h1 = hv * hv - ha * ha;
  double e = omega * c::WMEScale;
  double eF = e * F2;
  // h0, h1 and 3 are constants: their reciprocals fold at compile time
  double R0 = (cv * cv + 3. * (ca * ca) + 1.5 * (eF * eF) + 2. * (e * e) * cv * F2) * (1. / h0);
  double R1 = (cv * cv - ca * ca - 2. * (eF * eF) - 4. * (e * e) * cv * F2) * (1. / h1);
  return c0 * R0 - c1 * R1 * (1. / 3.);
}

// R0 of WMScatt (weak_magnetism.hpp:165-210): the correction of the zeroth
// Legendre coefficient of isoenergetic scattering on one nucleon
__device__ __forceinline__ double wm_scatt_r0(double omega, int reac) {
  double cv, ca, F2;
  form_factors(omega, reac, cv, ca, F2);
  double hv = reac == 1 ? c::Hpv : c::Hnv, ha = reac == 1 ? c::Hpa : c::Hna;
  double h0 = hv * hv + 3. * (ha * ha);
  double e = omega * c::WMEScale;
  double eF = e * F2;
  return (cv * cv + 3. * (ca * ca) + 1.5 * (eF * eF) + 2. * (e * e) * cv * F2) * (1. / h0);
}

// WMAbsEm (weak_magnetism.hpp:123-160): R (neutrino) or Rbar (antineutrino)
__device__ __forceinline__ double wm_abs(double omega, int s) {
  double cv, ca, F2;
  form_factors(omega, 3, cv, ca, F2);
  double e = omega * c::WMEScale;
  double a = fma(4. / 3., e, 1.);
  double tmp1 = (cv * cv) * (1. + 4. * e + (16. / 3.) * (e * e)) + 3. * ca * ca * (a * a) +
                (8. / 3.) * cv * F2 * (e * e) + (5. / 3.) * (e * e) * fma(2. / 5., e, 1.) * (F2 * F2);
  double tmp2 = 4. * (cv + F2) * ca * e * a;
  double b = fma(2., e, 1.);
  double tmp3 = (c::Gv * c::Gv + 3. * (c::Ga * c::Ga)) * (b * b * b);
  return div_normal(s == 0 ? tmp1 + tmp2 : tmp1 - tmp2, tmp3);
}

// EtaNNAbs (formulas_beta_iso.hpp:64-97)
__device__ __forceinline__ double eta_nn_abs(double n_in, double n_out, double mu_hat,
                                             double T) {
  if (fabs(mu_hat) < 1.0e-02) return n_in;
  double e = (n_out - n_in) / (safe_exp(-mu_hat / T) - 1.);
  return e < 0. ? n_in : e;
}

// Per-point scalars of the beta processes (AbsOpacity,
// formulas_beta_iso.hpp:291-319)
struct BetaPoint {
  double etanp, etapn;  // EtaNNAbs for n -> p and p -> n
  double Qp;            // dQ + dU
  double munu;          // mu_p + mu_e - mu_n
  double mu_e, T, rT;   // rT = 1 / T
  bool wm, decay;       // use_WM_ab, use_decay
};

__device__ __forceinline__ BetaPoint beta_point(double nb, double T, double yp, double yn,
                                                double mu_p, double mu_n, double mu_e,
                                                double dU_in, double dm_eff, bool use_dU,
                                                bool use_dm_eff, bool wm, bool decay) {
  BetaPoint bp;
  double nn_ = nb * yn, np_ = nb * yp;
  double dU = use_dU ? dU_in : 0.;
  double dQ = use_dm_eff ? dm_eff : c::Q;
  double mu_hat = mu_n - mu_p - dQ;
  bp.Qp = dQ + dU;
  double mu_np = mu_hat - dU;
  bp.etanp = eta_nn_abs(nn_, np_, mu_np, T);
  bp.etapn = eta_nn_abs(np_, nn_, -mu_np, T);
  bp.munu = mu_p + mu_e - mu_n;
  bp.mu_e = mu_e;
  bp.T = T;
  bp.rT = 1. / T;
  bp.wm = wm;
  bp.decay = decay;
  return bp;
}

// AbsOpacitySingleLep (formulas_beta_iso.hpp:138-279) for species s (0 nue,
// 1 anue) at energy nu: emissivity em and (not stimulated) absorptivity ab.
// The captured lepton has energy Ec (e- for nue, e+ for anue), the decay
// lepton Ed.
__device__ __forceinline__ void beta_em_ab(double nu, int s, const BetaPoint& bp, double& em,
                                           double& ab) {
  const double m = c::Me, m2 = m * m;
  const double T = bp.T;
  double R = bp.wm ? wm_abs(nu, s) : 1.;
  double sg = s == 0 ? 1. : -1.;
  double Ec = nu + sg * bp.Qp, Ed = -Ec;
  double muc = sg * bp.mu_e;
  double fc = fermi_of_arg(div_by_recip(Ec - muc, T, bp.rT));
  double fdd = fermi_of_arg(div_by_recip(Ed + muc, T, bp.rT));
  double cap = 0., dec = 0.;
  if (Ec - m >= 0.) cap = Ec * Ec * sqrt(1. - div_normal(m2, Ec * Ec)) * R;
  if (bp.decay && (Ed - m >= 0.)) dec = Ed * Ed * sqrt(1. - div_normal(m2, Ed * Ed));
  em = c::BetaConst * (s == 0 ? bp.etapn : bp.etanp) * (cap * fc + dec * (1. - fdd));
  ab = em * safe_exp(div_by_recip(nu - sg * bp.munu, T, bp.rT));
}

// EtaNNSc (formulas_beta_iso.hpp:371-395)
__device__ __forceinline__ double eta_nn_sc(double nb, double T, double yN) {
  double nN = yN * nb;
  if (nN <= 0.) return 0.;
  double eF = c::IsoEF * exp_any((2. / 3.) * log_pos(nN));
  double aux = 1.5 * T / eF;
  return nN * aux / sqrt(1. + aux * aux);
}

// Two species with bitwise identical distribution parameters have identical
// TotalNuF values at every energy (callers commonly give anux the moments of
// nux, as the reference's drivers do, tests/tests.hpp:551-553).  The
// comparison is made once per point; the result is wave-uniform.
__device__ __forceinline__ bool spec_same(const Spec& a, const Spec& b) {
  auto eq = [](double x, double y) { return __double_as_longlong(x) == __double_as_longlong(y); };
  return eq(a.wt, b.wt) && eq(a.mut, b.mut) && eq(a.invTt, b.invTt) && eq(a.wf, b.wf) &&
         eq(a.bf, b.bf) && eq(a.cf, b.cf) && eq(a.invTf, b.invTf) && eq(a.lbf, b.lbf);
}

// NuDistributionParams fields of one species (bns_nurates.hpp:387-401) as
// the caller gives them
__device__ __forceinline__ Spec spec_from_fields(double w_t, double temp_t, double eta_t,
                                                 double w_f, double temp_f, double c_f,
                                                 double beta_f) {
  Spec d;
  d.wt = w_t;
  d.mut = temp_t * eta_t;
  d.invTt = 1. / temp_t;
  d.wf = w_f;
  d.invTf = 1. / temp_f;
  d.cf = c_f;
  d.bf = beta_f;
  d.lbf = 0.;
  return d;
}

// CalculateThickParamsFromM1 + CalculateThinParamsFromM1
// (distribution.hpp:51-243) for one species from its moments n, J, chi;
// gamma36 = GammaStirling(3.6)
__device__ __forceinline__ Spec spec_from_m1(double n0, double J0, double chi, double gamma36) {
  double nn_ = n0 > 1e-100 ? n0 : 1e-100;
  double JJ = J0 > 1e-100 ? J0 : 1e-100;
  double r = nn_ / JJ;
  double y = nn_ * (r * r * r) / c::FourPiHc3;
  double et;
  if (y < 0.005) {
    double num = y * (y * (y * (y * (y * (y + 0.19926987701997) + 38865.0149220478) +
                 14364.6331099737) + 5750.1878570758) + 1120.71610972194) + 1.60356108438235e-8;
    double den = y * (y * (y * (y * (y * (y + 1.0) + 38840.0743174942) - 99.9009680656931) -
                 171.874844843596) + 75.7101579899442) + 83.0160130941424;
    et = log_pos(num / den);
  } else if (y <= 0.7) {
    double num = y * (y * (y * (y * (y * (41.3836568203438 * y + 32.5515666786612) -
                 157.774993512235) + 66.5726772426253) + 14.4883415579211) +
                 0.315360380575709) + 0.000660414331285249;
    double den = y * (y * (y * (y * (y * (y + 1.8888797407042) - 5.35488690539183) +
                 1.94673781342617) + 0.483128792035557) + 0.0113386564109086) +
                 2.64160073447322e-5;
    et = num / den - 30.;
  } else if (y > 0.7 && y < 0.7901234567745267) {
    double num = y * (y * (y * (y * (y * (3852.81416018959 * y - 5316.18895799799) +
                 1102.91561586553) + 1.54082262710661e-6) + 1732.89925128741) -
                 1769.59868329086) + 586.406885304906;
    double den = y * (y * (y * (y * (y * (y + 255.936658313629) + 9.42360945627147e-5) -
                 81.2467063138386) - 180.100197053091) - 89.0343496217014) +
                 143.849128123195;
    et = exp_any(num / den);
  } else {
    et = 20.;
  }
  et = et < 20. ? et : 20.;
  double F[6];
  fdi_bundle<2, 3>(et, F);
  double Tt = F[2] * JJ / (F[3] * nn_);
  Spec d;
  d.wt = 1.5 * (1. - chi);
  d.mut = Tt * et;
  d.invTt = 1. / Tt;
  d.wf = 0.5 * (mul_rn(3., chi) - 1.);
  double Tnu = JJ / (nn_ * 3.6);
  d.invTf = 1. / Tnu;
  d.cf = 0.6;
  d.bf = nn_ / (c::FourPiHc3 * gamma36 * exp_any(3.6 * log_pos(Tnu)));
  d.lbf = 0.;
  return d;
}

// Per-point, per-channel constants of the HR98 fit that do not depend on
// the cell (kernel_brem_HR98.hpp:239-338)
struct HrChan {
  double weight;  // n_N * s-weights of the channel sum (incl. 28/3 for np)
  double gamma_half_gb;
  double gamma_over_T;
  double lsd_pt;  // log of s_d / (cell factor): log(3 (pi/2)^2.5 es^-2.5 f_u)
  double f_pt;    // (1 + es^2)(1 + y^4)
  double g0;      // c1 y / (0.8 + 0.06 y^1.05) sqrt(es) / (es + 0.2)
  double h;       // h_brem
  double sd_pt;   // e^{-p lsd_pt}: s_d^-p = sd_pt e^{-p log(cell factor)}
  double rf_pt;   // 1 / f_pt
};

// The per-point scalars of the HR98 fit shared by the three channels
struct HrPoint {
  double y;       // 1.94 / (T / 10)
  double emy;     // SafeExp(-y / 12)
  double p;       // p_brem = 0.67 + 0.18 y^0.4
  double pinv;    // -1 / p_brem
  double medium;  // NN medium correction factor, or 1
};

__device__ __forceinline__ HrPoint hr98_point(double T, double nb, bool medium_corr) {
  HrPoint hp;
  double T10 = T * 0.1;
  hp.y = 1.94 / T10;
  hp.emy = safe_exp(-hp.y / 12.);
  hp.p = 0.67 + 0.18 * exp(0.4 * log(hp.y));
  hp.pinv = -1. / hp.p;
  hp.medium = 1.;
  if (medium_corr) {
    double cc = 1. + cbrt(nb / c::SaturationN) / 3.;
    double c2 = cc * cc;
    hp.medium = 1. / (c2 * c2 * c2);
  }
  return hp;
}

// Channel ch (0 nn, 1 pp, 2 np) of BremAllChannelsAbsKernel: the
// x-independent parts of BremSingleChannelAbsKernel, BremKernelG and
// BremKernelS (kernel_brem_HR98.hpp:33-338); y = HrPoint::y
__device__ __forceinline__ HrChan hr98_channel(int ch, double T, double nb, double yp,
                                               double yn, double y, double p) {
  const double T10 = T * 0.1;
  double xmean = sqrt(yn * yp);
  double nNc = ch == 0 ? nb * yn : ch == 1 ? nb * yp : nb * xmean;
  double mNc = ch == 0 ? c::MnGrams : ch == 1 ? c::MpGrams : c::MAvgGrams;
  double cwc = ch == 2 ? 28. / 3. : 1.;
  double ye = y > c::BremMin ? y : c::BremMin;
  double y4 = ye * ye * ye * ye;
  double y105 = exp(1.05 * log(ye));
  double es = exp((2. / 3.) * log(3. * c::PiSquared * nNc)) * c::BremAux1 / (2. * mNc * T);
  double gamma = 1.63 * exp(1.5 * log(es)) * T10;
  if (!(nNc > 0.)) {  // pow(0, 2/3) = 0 in the reference
    es = 0.;
    gamma = 0.;
  }
  // BremKernelG (kernel_brem_HR98.hpp:165-236), clamped arguments
  double yg = ye;
  double eg = es > c::BremMin ? es : c::BremMin;
  double y2 = yg * yg, ei = 1. / eg, a1d = 25. * y2 + 1.;
  double alpha1 = (0.5 + ei) / (1. + ei) * (1. / a1d) + (0.5 + eg / 15.6) * 25. * y2 / a1d;
  double alpha2 = (0.63 + 0.04 * exp(1.45 * log(eg))) / (1. + 0.02 * exp(2.5 * log(eg)));
  double eg15 = exp(1.5 * log(eg));
  double alpha3 = 1.2 * safe_exp(0.6 * eg - 0.4 * eg15);
  double p1 = (1.8 + 0.45 * eg) / (1. + 0.15 * eg15);
  double p2 = 2.3 - 0.05 * eg / (1. + 0.025 * eg);
  double ly = log(yg);
  double gb = (alpha1 + alpha2 * exp(p1 * ly)) /
              (1. + alpha3 * exp(p2 * ly) + alpha2 * exp((p1 + 2.) * ly) / 13.75);
  // BremKernelS, x-independent factors
  double u = sqrt(ye / (2. * eg)) + 1.e-10;
  double u2 = u * u;
  double uarg = u2 / (2. * sqrt(2. * u2 + 4.));
  double fu = (1. - 5. * u * atan(2. / u) / 6. + u2 / (3. * (u2 + 4.)) +
               atan(1. / uarg) * uarg / 3.);
  fu = fabs(fu) < 1e-14 ? 1e-14 : fu;
  HrChan h;
  h.lsd_pt = log(3. * c::PiHalfToFiveHalves * exp(-2.5 * log(eg)) * fu);
  h.f_pt = (1. + eg * eg) * (1. + y4);
  h.sd_pt = exp_any(-p * h.lsd_pt);
  h.rf_pt = 1. / h.f_pt;
  h.g0 = 0.0044 * ye / (0.8 + 0.06 * y105) * sqrt(eg) / (eg + 0.2);
  h.h = 0.1 * eg / (2.39 + 0.1 * exp(1.1 * log(eg)));
  h.gamma_half_gb = 0.5 * gamma * gb;
  h.gamma_over_T = gamma / T;
  h.weight = cwc * nNc;
  return h;
}


}  // namespace

}  // namespace bns

#endif  // BNS_NURATES_AMD_POINT_PHYSICS_H_
