// oracle/ref_driver.cpp -- C entry points over the genuine reference
// (RelNucAs/bns_nurates, header-only C++), compiled from its own headers
// where they lie under /root/reference/include (recipe: oracle/Makefile,
// output: oracle/_ref/libbnsref.so, git-ignored).
//
// TEST INFRASTRUCTURE ONLY: used by tests/ and tests/golden/make_golden.py to
// pin the CPU restatement (oracle/m1_oracle.hpp) against the reference.
//
// The reference #includes a GP19 table blob that is absent from the tree
// (kernel_brem_GP19.hpp:18-20).  The reference ships its own configuration
// without that table -- the device/SYCL branch of the same guards, in which
// BremKernelAbsGP19 returns zero (kernel_brem_GP19.hpp:483-493) -- and the
// recipe selects it with -D__SYCL_DEVICE_ONLY__.  GP19 is not on the path
// under test; nothing else in the reference reads that macro.
//
// BS_N_MAX is 20 in the reference (bns_nurates.hpp:68), so the gray entry
// point accepts quadrature orders n <= 10 and the spectral one n <= 20; both
// return 2 otherwise.
#include <cstdint>
#include <cstring>

#include "m1_opacities.hpp"

namespace {
MyEOSParams eos_at(const double* pts, int64_t np, int64_t i) {
  MyEOSParams e;
  std::memset(&e, 0, sizeof(e));
  e.nb = pts[0 * np + i];
  e.temp = pts[1 * np + i];
  e.yp = pts[2 * np + i];
  e.yn = pts[3 * np + i];
  e.mu_p = pts[4 * np + i];
  e.mu_n = pts[5 * np + i];
  e.mu_e = pts[6 * np + i];
  e.dU = pts[7 * np + i];
  e.dm_eff = pts[8 * np + i];
  return e;
}
OpacityParams params_from(const int* p) {
  OpacityParams o;
  o.use_dU = p[0] != 0;
  o.use_dm_eff = p[1] != 0;
  o.use_WM_ab = p[2] != 0;
  o.use_WM_sc = p[3] != 0;
  o.use_decay = p[4] != 0;
  o.brem_implementation = p[5] == 1 ? BREM_BRT06 : (p[5] == 2 ? BREM_GP19 : BREM_HR98);
  o.use_NN_medium_corr = p[6] != 0;
  o.neglect_blocking = p[7] != 0;
  return o;
}
void distr_load(const double* d, int64_t np, int64_t i, NuDistributionParams& o) {
  for (int s = 0; s < 4; ++s) {
    o.w_t[s] = d[(0 + s) * np + i];
    o.temp_t[s] = d[(4 + s) * np + i];
    o.eta_t[s] = d[(8 + s) * np + i];
    o.w_f[s] = d[(12 + s) * np + i];
    o.temp_f[s] = d[(16 + s) * np + i];
    o.c_f[s] = d[(20 + s) * np + i];
    o.beta_f[s] = d[(24 + s) * np + i];
  }
}
void distr_store(const NuDistributionParams& o, double* d, int64_t np, int64_t i) {
  for (int s = 0; s < 4; ++s) {
    d[(0 + s) * np + i] = o.w_t[s];
    d[(4 + s) * np + i] = o.temp_t[s];
    d[(8 + s) * np + i] = o.eta_t[s];
    d[(12 + s) * np + i] = o.w_f[s];
    d[(16 + s) * np + i] = o.temp_f[s];
    d[(20 + s) * np + i] = o.c_f[s];
    d[(24 + s) * np + i] = o.beta_f[s];
  }
}
}  // namespace

extern "C" {

int ref_bs_n_max(void) { return BS_N_MAX; }

void ref_gauss_legendre(int n, double x1, double x2, double* pts, double* w) {
  MyQuadrature q = quadrature_default;
  q.nx = n;
  q.x1 = x1;
  q.x2 = x2;
  GaussLegendre(&q);
  for (int i = 0; i < n; ++i) {
    pts[i] = q.points[i];
    w[i] = q.w[i];
  }
}

void ref_fdi(int kord, int64_t n, const double* y, double* out) {
  for (int64_t i = 0; i < n; ++i) {
    switch (kord) {
      case 0: out[i] = FDI_0(y[i]); break;
      case 1: out[i] = FDI_p1(y[i]); break;
      case 2: out[i] = FDI_p2(y[i]); break;
      case 3: out[i] = FDI_p3(y[i]); break;
      case 4: out[i] = FDI_p4(y[i]); break;
      default: out[i] = FDI_p5(y[i]); break;
    }
  }
}

void ref_fermi_distr(int64_t n, const double* e, const double* T, const double* mu, double* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = FermiDistr(e[i], T[i], mu[i]);
}

void ref_bessk1(int64_t n, const double* x, double* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = bessk1(x[i]);
}

void ref_pair_kernels(int64_t n, const double* om, const double* omp, const double* T,
                      const double* mu_e, double* em, double* ab) {
  for (int64_t i = 0; i < n; ++i) {
    MyEOSParams e;
    std::memset(&e, 0, sizeof(e));
    e.temp = T[i];
    e.mu_e = mu_e[i];
    PairKernelParams kp;
    std::memset(&kp, 0, sizeof(kp));
    kp.omega = om[i];
    kp.omega_prime = omp[i];
    kp.cos_theta = 1.;
    kp.mu = 1.;
    kp.mu_prime = 1.;
    MyKernelOutput o = PairKernels(&e, &kp);
    for (int s = 0; s < 4; ++s) {
      em[i * 4 + s] = o.em[s];
      ab[i * 4 + s] = o.abs[s];
    }
  }
}

void ref_brem_brt06(int64_t n, const double* om, const double* omp, const double* nb,
                    const double* T, const double* yn, const double* yp, double* em,
                    double* ab) {
  for (int64_t i = 0; i < n; ++i) {
    MyEOSParams e;
    std::memset(&e, 0, sizeof(e));
    e.nb = nb[i];
    e.temp = T[i];
    e.yn = yn[i];
    e.yp = yp[i];
    BremKernelParams kp;
    std::memset(&kp, 0, sizeof(kp));
    kp.omega = om[i];
    kp.omega_prime = omp[i];
    MyKernelOutput o = BremKernelsBRT06(&kp, &e);
    em[i] = o.em[0];
    ab[i] = o.abs[0];
  }
}

void ref_brem_hr98(int64_t n, const double* om, const double* omp, const double* nb,
                   const double* T, const double* yn, const double* yp, int medium,
                   double* em, double* ab) {
  for (int64_t i = 0; i < n; ++i) {
    MyEOSParams e;
    std::memset(&e, 0, sizeof(e));
    e.nb = nb[i];
    e.temp = T[i];
    e.yn = yn[i];
    e.yp = yp[i];
    BremKernelParams kp;
    std::memset(&kp, 0, sizeof(kp));
    kp.omega = om[i];
    kp.omega_prime = omp[i];
    kp.l = 0;
    kp.use_NN_medium_corr = medium != 0;
    MyKernelOutput o = BremKernelsLegCoeff(&kp, &e);
    em[i] = o.em[0];
    ab[i] = o.abs[0];
  }
}

void ref_inel_kernels(int64_t n, const double* om, const double* omp, const double* T,
                      const double* mu_e, double* em, double* ab) {
  for (int64_t i = 0; i < n; ++i) {
    MyEOSParams e;
    std::memset(&e, 0, sizeof(e));
    e.temp = T[i];
    e.mu_e = mu_e[i];
    InelasticScattKernelParams kp;
    kp.omega = om[i];
    kp.omega_prime = omp[i];
    MyKernelOutput o = InelasticScattKernels(&kp, &e);
    for (int s = 0; s < 4; ++s) {
      em[i * 4 + s] = o.em[s];
      ab[i * 4 + s] = o.abs[s];
    }
  }
}

void ref_abs_opacity(int64_t n, const double* om, const double* eos9, const int* params,
                     int stim, double* em, double* ab) {
  MyEOSParams e = eos_at(eos9, 1, 0);
  OpacityParams p = params_from(params);
  for (int64_t i = 0; i < n; ++i) {
    MyOpacity o = stim ? StimAbsOpacity(om[i], &p, &e) : AbsOpacity(om[i], &p, &e);
    em[i * 2] = o.em[0];
    em[i * 2 + 1] = o.em[1];
    ab[i * 2] = o.abs[0];
    ab[i * 2 + 1] = o.abs[1];
  }
}

void ref_iso_scatt(int64_t n, const double* om, const double* eos9, const int* params,
                   double* out) {
  MyEOSParams e = eos_at(eos9, 1, 0);
  OpacityParams p = params_from(params);
  for (int64_t i = 0; i < n; ++i) out[i] = IsoScattTotal(om[i], &p, &e);
}

void ref_iso_leg0(int64_t n, const double* om, const double* eos9, const int* params,
                  double* out) {
  MyEOSParams e = eos_at(eos9, 1, 0);
  OpacityParams p = params_from(params);
  for (int64_t i = 0; i < n; ++i) out[i] = IsoScattLegCoeff(om[i], &p, &e, 0);
}

void ref_distr_from_m1(int64_t np, const double* pts, double* distr) {
  for (int64_t i = 0; i < np; ++i) {
    M1Quantities m;
    std::memset(&m, 0, sizeof(m));
    for (int s = 0; s < 4; ++s) {
      m.n[s] = pts[(9 + s) * np + i];
      m.J[s] = pts[(13 + s) * np + i];
      m.chi[s] = pts[(17 + s) * np + i];
    }
    MyEOSParams e = eos_at(pts, np, i);
    NuDistributionParams d = CalculateDistrParamsFromM1(&m, &e);
    distr_store(d, distr, np, i);
  }
}

void ref_distr_equilibrium(int64_t np, const double* pts, double* distr) {
  for (int64_t i = 0; i < np; ++i) {
    MyEOSParams e = eos_at(pts, np, i);
    NuDistributionParams d = NuEquilibriumParams(&e);
    distr_store(d, distr, np, i);
  }
}

void ref_m1_densities_eq(int64_t np, double* pts, const double* distr) {
  for (int64_t i = 0; i < np; ++i) {
    MyEOSParams e = eos_at(pts, np, i);
    NuDistributionParams d;
    distr_load(distr, np, i, d);
    M1Quantities m;
    std::memset(&m, 0, sizeof(m));
    ComputeM1DensitiesEq(&e, &d, &m);
    for (int s = 0; s < 4; ++s) {
      pts[(9 + s) * np + i] = m.n[s];
      pts[(13 + s) * np + i] = m.J[s];
    }
  }
}

void ref_total_nuf(int64_t n, const double* om, const double* distr28, int s, double* out) {
  NuDistributionParams d;
  distr_load(distr28, 1, 0, d);
  for (int64_t i = 0; i < n; ++i) out[i] = TotalNuF(om[i], &d, s);
}

int ref_m1_opacities(int64_t np, int n1, const double* x1, const double* w1, int n2,
                     const double* x2, const double* w2, const double* pts,
                     const double* distr, const int* flags, const int* params, int stim,
                     double* out, int nthreads) {
  if (2 * n1 > BS_N_MAX || 2 * n2 > BS_N_MAX || n1 < 1 || n2 < 1) return 2;
  MyQuadrature q1 = quadrature_default, q2 = quadrature_default;
  q1.nx = n1;
  q2.nx = n2;
  for (int i = 0; i < n1; ++i) {
    q1.points[i] = x1[i];
    q1.w[i] = w1[i];
  }
  for (int i = 0; i < n2; ++i) {
    q2.points[i] = x2[i];
    q2.w[i] = w2[i];
  }
  (void)nthreads;
#pragma omp parallel for schedule(dynamic, 8) if (nthreads != 1)
  for (int64_t i = 0; i < np; ++i) {
    GreyOpacityParams g;
    std::memset(&g, 0, sizeof(g));
    g.opacity_pars = params_from(params);
    g.eos_pars = eos_at(pts, np, i);
    for (int s = 0; s < 4; ++s) {
      g.m1_pars.n[s] = pts[(9 + s) * np + i];
      g.m1_pars.J[s] = pts[(13 + s) * np + i];
      g.m1_pars.chi[s] = pts[(17 + s) * np + i];
    }
    if (distr) distr_load(distr, np, i, g.distr_pars);
    else g.distr_pars = CalculateDistrParamsFromM1(&g.m1_pars, &g.eos_pars);
    g.opacity_flags.use_abs_em = flags[0];
    g.opacity_flags.use_pair = flags[1];
    g.opacity_flags.use_brem = flags[2];
    g.opacity_flags.use_inelastic_scatt = flags[3];
    g.opacity_flags.use_iso = flags[4];
    MyQuadrature lq1 = q1, lq2 = q2;
    M1Opacities o = stim ? ComputeM1Opacities(&lq1, &lq2, &g)
                         : ComputeM1OpacitiesNotStimulated(&lq1, &lq2, &g);
    for (int s = 0; s < 4; ++s) {
      out[(0 + s) * np + i] = o.eta_0[s];
      out[(4 + s) * np + i] = o.kappa_0_a[s];
      out[(8 + s) * np + i] = o.eta[s];
      out[(12 + s) * np + i] = o.kappa_a[s];
      out[(16 + s) * np + i] = o.kappa_s[s];
    }
  }
  return 0;
}

// ComputeSpectralOpacities{Stimulated,NotStimulated}Abs over np points x
// nnu energies; arguments and output layout of orc_spectral_opacities.
int ref_spectral_opacities(int64_t np, int n, const double* x, const double* w,
                           const double* pts, const double* distr, int64_t nnu,
                           const double* nu, int64_t nu_ld, const int* flags,
                           const int* params, int stim, double* out, int nthreads) {
  if (n > BS_N_MAX || n < 1) return 2;
  MyQuadrature q = quadrature_default;
  q.nx = n;
  for (int i = 0; i < n; ++i) {
    q.points[i] = x[i];
    q.w[i] = w[i];
  }
  (void)nthreads;
#pragma omp parallel for schedule(dynamic, 4) collapse(2) if (nthreads != 1)
  for (int64_t i = 0; i < np; ++i) {
    for (int64_t e = 0; e < nnu; ++e) {
      GreyOpacityParams g;
      std::memset(&g, 0, sizeof(g));
      g.opacity_pars = params_from(params);
      g.eos_pars = eos_at(pts, np, i);
      for (int s = 0; s < 4; ++s) {
        g.m1_pars.n[s] = pts[(9 + s) * np + i];
        g.m1_pars.J[s] = pts[(13 + s) * np + i];
        g.m1_pars.chi[s] = pts[(17 + s) * np + i];
      }
      if (distr) distr_load(distr, np, i, g.distr_pars);
      else g.distr_pars = CalculateDistrParamsFromM1(&g.m1_pars, &g.eos_pars);
      g.opacity_flags.use_abs_em = flags[0];
      g.opacity_flags.use_pair = flags[1];
      g.opacity_flags.use_brem = flags[2];
      g.opacity_flags.use_inelastic_scatt = flags[3];
      g.opacity_flags.use_iso = flags[4];
      MyQuadrature lq = q;
      const double nue = nu[nu_ld ? e * nu_ld + i : e];
      SpectralOpacities o = stim ? ComputeSpectralOpacitiesStimulatedAbs(nue, &lq, &g)
                                 : ComputeSpectralOpacitiesNotStimulatedAbs(nue, &lq, &g);
      for (int s = 0; s < 4; ++s) {
        out[((0 + s) * nnu + e) * np + i] = o.j[s];
        out[((4 + s) * nnu + e) * np + i] = o.kappa[s];
        out[((8 + s) * nnu + e) * np + i] = o.j_s[s];
        out[((12 + s) * nnu + e) * np + i] = o.kappa_s[s];
      }
    }
  }
  return 0;
}

}  // extern "C"
