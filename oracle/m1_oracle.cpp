// oracle/m1_oracle.cpp -- C entry points of the CPU oracle (FP64).
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
// the cpu_baseline leg of bench.py as the checker / the timed CPU baseline.
// The product (bns_nurates_amd) never links or calls it.
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off -fopenmp, no -march so no
// FMA contraction: the same arithmetic as the reference built the same way).
#include "m1_oracle.hpp"

#include <omp.h>

using namespace orc;

namespace {
Eos<double> eos_at(const double* pts, int64_t np, int64_t i) {
  Eos<double> e;
  e.nb = pts[P_NB * np + i];
  e.T = pts[P_T * np + i];
  e.yp = pts[P_YP * np + i];
  e.yn = pts[P_YN * np + i];
  e.mu_p = pts[P_MUP * np + i];
  e.mu_n = pts[P_MUN * np + i];
  e.mu_e = pts[P_MUE * np + i];
  e.dU = pts[P_DU * np + i];
  e.dm_eff = pts[P_DMEFF * np + i];
  return e;
}
M1<double> m1_at(const double* pts, int64_t np, int64_t i) {
  M1<double> m;
  for (int s = 0; s < 4; ++s) {
    m.n[s] = pts[(P_N0 + s) * np + i];
    m.J[s] = pts[(P_J0 + s) * np + i];
    m.chi[s] = pts[(P_CHI0 + s) * np + i];
  }
  return m;
}
void distr_load(const double* d, int64_t np, int64_t i, Distr<double>& o) {
  for (int s = 0; s < 4; ++s) {
    o.w_t[s] = d[(D_WT + s) * np + i];
    o.temp_t[s] = d[(D_TT + s) * np + i];
    o.eta_t[s] = d[(D_ETAT + s) * np + i];
    o.w_f[s] = d[(D_WF + s) * np + i];
    o.temp_f[s] = d[(D_TF + s) * np + i];
    o.c_f[s] = d[(D_CF + s) * np + i];
    o.beta_f[s] = d[(D_BF + s) * np + i];
  }
}
void distr_store(const Distr<double>& o, double* d, int64_t np, int64_t i) {
  for (int s = 0; s < 4; ++s) {
    d[(D_WT + s) * np + i] = o.w_t[s];
    d[(D_TT + s) * np + i] = o.temp_t[s];
    d[(D_ETAT + s) * np + i] = o.eta_t[s];
    d[(D_WF + s) * np + i] = o.w_f[s];
    d[(D_TF + s) * np + i] = o.temp_f[s];
    d[(D_CF + s) * np + i] = o.c_f[s];
    d[(D_BF + s) * np + i] = o.beta_f[s];
  }
}
Flags flags_from(const int* f) { return Flags{f[0], f[1], f[2], f[3], f[4]}; }
Params params_from(const int* p) {
  return Params{p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7]};
}
}  // namespace

extern "C" {

int orc_version(void) { return 1; }

// GaussLegendre on [x1, x2] (integration.hpp:28-87)
void orc_gauss_legendre(int n, double x1, double x2, double* pts, double* w) {
  gauss_legendre(n, x1, x2, pts, w);
}

// F_k(y) elementwise, k = 0..5
void orc_fdi(int kord, int64_t n, const double* y, double* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = fdi(kord, y[i]);
}

void orc_fermi_distr(int64_t n, const double* e, const double* T, const double* mu, double* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = fermi_distr(e[i], T[i], mu[i]);
}

void orc_set_gp19_table(int nnb, const double* nb, const double* lognb, int nye, const double* ye,
                        int nT, const double* T, int nw, const double* w, const double* logw,
                        const double* data) {
  orc::gp19_assign(nnb, nb, lognb, nye, ye, nT, T, nw, w, logw, data);
}

// gp19_interpolator (nb [fm^-3], Ye, T, w [MeV]), elementwise
void orc_gp19_interp(int64_t n, const double* nb, const double* ye, const double* T,
                     const double* w, double* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = orc::gp19_interp<double>(nb[i], ye[i], T[i], w[i]);
}

// BremKernelAbsGP19 (omega, omega', nb [nm^-3], T, Ye = yp), elementwise
void orc_brem_gp19(int64_t n, const double* om, const double* omp, const double* nb,
                   const double* T, const double* yp, double* em, double* ab) {
  for (int64_t i = 0; i < n; ++i) {
    orc::Eos<double> e{};
    e.nb = nb[i];
    e.T = T[i];
    e.yp = yp[i];
    e.yn = 1. - yp[i];
    orc::brem_gp19<double>(om[i], omp[i], e, em[i], ab[i]);
  }
}

void orc_bessk1(int64_t n, const double* x, double* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = bessk1(x[i]);
}

// PairKernels forward (kernel_pair.hpp:165-198): em/abs [i*4+s]
void orc_pair_kernels(int64_t n, const double* om, const double* omp, const double* T,
                      const double* mu_e, double* em, double* ab) {
  for (int64_t i = 0; i < n; ++i) {
    double eta = mu_e[i] / T[i];
    double y = om[i] / T[i], z = omp[i] / T[i];
    double F[6][7];
    double a[6], b[6], c[6], d[6], e[6], f[6], g[6];
    fdi_bundle(1, eta - y, a);
    fdi_bundle(1, eta - z, b);
    fdi_bundle(1, eta + y, c);
    fdi_bundle(1, eta + z, d);
    fdi_bundle(3, eta, e);
    fdi_bundle(3, eta - y - z, f);
    fdi_bundle(3, eta + y + z, g);
    for (int kk = 1; kk <= 5; ++kk) {
      F[kk][A_EMY] = a[kk];
      F[kk][A_EMZ] = b[kk];
      F[kk][A_EPY] = c[kk];
      F[kk][A_EPZ] = d[kk];
      if (kk >= 3) {
        F[kk][A_E] = e[kk];
        F[kk][A_EMYZ] = f[kk];
        F[kk][A_EPYZ] = g[kk];
      }
    }
    KOut<double> o;
    pair_kernel_fwd(om[i], omp[i], T[i], eta, y, z, F, o);
    for (int s = 0; s < 4; ++s) {
      em[i * 4 + s] = o.em[s];
      ab[i * 4 + s] = o.abs[s];
    }
  }
}

// BremKernelsBRT06 (kernel_brem_BRT06.hpp:42-69)
void orc_brem_brt06(int64_t n, const double* om, const double* omp, const double* nb,
                    const double* T, const double* yn, const double* yp, double* em,
                    double* ab) {
  for (int64_t i = 0; i < n; ++i) {
    Eos<double> e{};
    e.nb = nb[i];
    e.T = T[i];
    e.yn = yn[i];
    e.yp = yp[i];
    brem_brt06(om[i], omp[i], e, em[i], ab[i]);
  }
}

// BremKernelsLegCoeff, l = 0 (kernel_brem_HR98.hpp:342-393)
void orc_brem_hr98(int64_t n, const double* om, const double* omp, const double* nb,
                   const double* T, const double* yn, const double* yp, int medium,
                   double* em, double* ab) {
  for (int64_t i = 0; i < n; ++i) {
    Eos<double> e{};
    e.nb = nb[i];
    e.T = T[i];
    e.yn = yn[i];
    e.yp = yp[i];
    brem_hr98(om[i], omp[i], e, medium != 0, em[i], ab[i]);
  }
}

// InelasticScattKernels (kernel_nes.hpp:343-359): em/abs [i*4+s]
void orc_inel_kernels(int64_t n, const double* om, const double* omp, const double* T,
                      const double* mu_e, double* em, double* ab) {
  for (int64_t i = 0; i < n; ++i) {
    NepsPoint<double> P;
    P.T = T[i];
    P.eta[0] = mu_e[i] / T[i];
    P.eta[1] = -mu_e[i] / T[i];
    for (int lep = 0; lep < 2; ++lep) {
      double b[6];
      fdi_bundle(0, P.eta[lep], b);
      for (int kk = 0; kk < 6; ++kk) P.Fe[lep][kk] = b[kk];
      P.fd0[lep] = fermi_distr(0., 1., P.eta[lep]);
    }
    double G[2][6][3];
    neps_fdi_table(om[i] / T[i], omp[i] / T[i], P, G);
    KOut<double> o;
    inel_kernels(om[i], omp[i], P, G, false, o);
    for (int s = 0; s < 4; ++s) {
      em[i * 4 + s] = o.em[s];
      ab[i * 4 + s] = o.abs[s];
    }
  }
}

// AbsOpacity / StimAbsOpacity (formulas_beta_iso.hpp:291-341) for n energies
// at one EOS point given as the 9 eos fields (nb,T,yp,yn,mu_p,mu_n,mu_e,dU,
// dm_eff); em/abs [i*2+s] for s = nue, anue.
void orc_abs_opacity(int64_t n, const double* om, const double* eos9, const int* params,
                     int stim, double* em, double* ab) {
  Eos<double> e{eos9[0], eos9[1], eos9[2], eos9[3], eos9[4], eos9[5], eos9[6], eos9[7], eos9[8]};
  Params p = params_from(params);
  for (int64_t i = 0; i < n; ++i) {
    double m[2], a[2];
    abs_opacity(om[i], p, e, m, a);
    if (stim) {
      a[0] = a[0] + m[0];
      a[1] = a[1] + m[1];
    }
    em[i * 2] = m[0];
    em[i * 2 + 1] = m[1];
    ab[i * 2] = a[0];
    ab[i * 2 + 1] = a[1];
  }
}

// IsoScattTotal (formulas_beta_iso.hpp:497-507)
void orc_iso_scatt(int64_t n, const double* om, const double* eos9, const int* params,
                   double* out) {
  Eos<double> e{eos9[0], eos9[1], eos9[2], eos9[3], eos9[4], eos9[5], eos9[6], eos9[7], eos9[8]};
  Params p = params_from(params);
  for (int64_t i = 0; i < n; ++i) out[i] = iso_scatt_total(om[i], p, e);
}

// IsoScattLegCoeff, l = 0 (formulas_beta_iso.hpp:524-560)
void orc_iso_leg0(int64_t n, const double* om, const double* eos9, const int* params,
                  double* out) {
  Eos<double> e{eos9[0], eos9[1], eos9[2], eos9[3], eos9[4], eos9[5], eos9[6], eos9[7], eos9[8]};
  Params p = params_from(params);
  for (int64_t i = 0; i < n; ++i) out[i] = iso_leg0(om[i], p, e);
}

// CalculateDistrParamsFromM1 (distribution.hpp:323-333) over the SoA point
// block; distr receives the 28-field block.
void orc_distr_from_m1(int64_t np, const double* pts, double* distr) {
  for (int64_t i = 0; i < np; ++i) {
    Distr<double> d;
    distr_from_m1(m1_at(pts, np, i), d);
    distr_store(d, distr, np, i);
  }
}

// NuEquilibriumParams (distribution.hpp:253-290)
void orc_distr_equilibrium(int64_t np, const double* pts, double* distr) {
  for (int64_t i = 0; i < np; ++i) {
    Distr<double> d;
    distr_equilibrium(eos_at(pts, np, i), d);
    distr_store(d, distr, np, i);
  }
}

// ComputeM1DensitiesEq (distribution.hpp:447-473): writes n, J into the
// point block (fields P_N0.., P_J0..).
void orc_m1_densities_eq(int64_t np, double* pts, const double* distr) {
  for (int64_t i = 0; i < np; ++i) {
    Distr<double> d;
    distr_load(distr, np, i, d);
    M1<double> m;
    m1_densities_eq(eos_at(pts, np, i), d, m);
    for (int s = 0; s < 4; ++s) {
      pts[(P_N0 + s) * np + i] = m.n[s];
      pts[(P_J0 + s) * np + i] = m.J[s];
    }
  }
}

// TotalNuF (distribution.hpp:298-315) for n energies of one point's
// distribution (28 values, one per field) and species s.
void orc_total_nuf(int64_t n, const double* om, const double* distr28, int s, double* out) {
  Distr<double> d;
  distr_load(distr28, 1, 0, d);
  for (int64_t i = 0; i < n; ++i) out[i] = total_nuf(om[i], d, s);
}

// ComputeM1Opacities (stim = 1) / ComputeM1OpacitiesNotStimulated (stim = 0)
// over npts points (m1_opacities.hpp:1297-1313).  When distr is NULL the
// distribution parameters are reconstructed from the M1 moments
// (CalculateDistrParamsFromM1), as the reference's callers do.
int orc_m1_opacities(int64_t np, int n1, const double* x1, const double* w1, int n2,
                     const double* x2, const double* w2, const double* pts,
                     const double* distr, const int* flags, const int* params, int stim,
                     double* out, int nthreads) {
  if (np < 0 || n1 < 1 || n2 < 1) return 1;
  Quad q1{n1, x1, w1}, q2{n2, x2, w2};
  Flags fl = flags_from(flags);
  Params pr = params_from(params);
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
  {
    Scratch<double> M;
#pragma omp for schedule(dynamic, 16)
    for (int64_t i = 0; i < np; ++i) {
      Eos<double> e = eos_at(pts, np, i);
      M1<double> m = m1_at(pts, np, i);
      Distr<double> d;
      if (distr) distr_load(distr, np, i, d);
      else distr_from_m1(m, d);
      double o[20];
      m1_opacities_point(q1, q2, e, m, d, fl, pr, stim, M, o);
      for (int f = 0; f < 20; ++f) out[f * np + i] = o[f];
    }
  }
  return 0;
}

// ComputeSpectralOpacitiesStimulatedAbs (stim = 1) / ...NotStimulatedAbs
// (stim = 0), m1_opacities.hpp:1505-1660, for np points x nnu energies with
// the 1D quadrature (n, x, w).  Energies: nu[e] for every point when
// nu_ld == 0, else nu[e * nu_ld + i].  out[(f * nnu + e) * np + i], f in
// S_J.. S_KAPPAS + 3.  distr NULL: reconstructed from the M1 moments.
int orc_spectral_opacities(int64_t np, int n, const double* x, const double* w,
                           const double* pts, const double* distr, int64_t nnu,
                           const double* nu, int64_t nu_ld, const int* flags,
                           const int* params, int stim, double* out, int nthreads) {
  if (np < 0 || nnu < 0 || n < 1) return 1;
  Quad q{n, x, w};
  Flags fl = flags_from(flags);
  Params pr = params_from(params);
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  if (n > kSpecMaxNodes / 2) return 1;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
  for (int64_t i = 0; i < np; ++i) {
    Eos<double> eos = eos_at(pts, np, i);
    Distr<double> d;
    if (distr) distr_load(distr, np, i, d);
    else distr_from_m1(m1_at(pts, np, i), d);
    spectral_point(
        q, eos, d, fl, pr, nnu, [&](int64_t e) { return nu[nu_ld ? e * nu_ld + i : e]; }, stim,
        [&](int64_t e, const double* o) {
          for (int f = 0; f < 16; ++f) out[(f * nnu + e) * np + i] = o[f];
        });
  }
  return 0;
}

}  // extern "C"
