// bns_nurates_amd.hpp -- header-only C++ host API of the MI355X gray M1
// opacity library, over the C ABI of bns_nurates_amd.h.
//
// Source-level drop-in for the gray path of bns_nurates: the types a caller
// fills and reads keep the reference's names and field names
// (bns_nurates.hpp:76-474: MyQuadrature, OpacityParams, OpacityFlags,
// MyEOSParams, NuDistributionParams, M1Quantities, GreyOpacityParams,
// M1Opacities, BremImpl, the species ids and the default parameter sets), and
// the per-point functions keep their signatures:
//
//   GaussLegendre                    integration.hpp:28
//   CalculateDistrParamsFromM1       distribution.hpp:324
//   NuEquilibriumParams              distribution.hpp:254
//   ComputeM1DensitiesEq             distribution.hpp:448
//   ComputeM1Opacities               m1_opacities.hpp:1306
//   ComputeM1OpacitiesNotStimulated  m1_opacities.hpp:1297
//   ComputeSpectralOpacitiesNotStimulatedAbs  m1_opacities.hpp:1505
//   ComputeSpectralOpacitiesStimulatedAbs     m1_opacities.hpp:1644
//
// Every one of them runs on the GPU through a process-wide default context
// (device 0, or $BNSAMD_DEVICE).  A per-point call is one small launch; the
// batched overloads (arrays of GreyOpacityParams) and bns_amd::Context (host
// arrays or device SoA buffers) are the forms meant for many points.
// Differences from the reference's structs: GreyOpacityParams has no
// kernel_pars member (scratch of the reference's implementation, not an
// input); BS_N_MAX defaults to 32, the largest quadrature the library takes.
// Errors throw bns_amd::Error (status code + message); nothing aborts.
#ifndef BNS_NURATES_AMD_HPP_
#define BNS_NURATES_AMD_HPP_

#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "bns_nurates_amd.h"

typedef double BS_REAL;

#ifndef id_nue
#define id_nue 0
#define id_anue 1
#define id_nux 2
#define id_anux 3
#endif
#ifndef total_num_species
#define total_num_species 4
#endif
#ifndef BS_N_MAX
#define BS_N_MAX 32
#endif
static_assert(BS_N_MAX >= 1 && BS_N_MAX <= 32, "the library takes quadratures of 1..32 nodes");

// ---- types (bns_nurates.hpp:76-474) --------------------------------------------
enum Quadrature { kGauleg, kGaulag };

struct MyQuadrature {
  enum Quadrature type;
  BS_REAL alpha;
  int dim;
  int nx, ny, nz;
  BS_REAL x1, x2, y1, y2, z1, z2;
  BS_REAL points[BS_N_MAX];
  BS_REAL w[BS_N_MAX];
};
static const MyQuadrature quadrature_default = {kGauleg, 0., 1, 20, 1, 1, 0., 1.,
                                                -42., -42., -42., -42., {0.}, {0.}};

enum BremImpl { BREM_HR98 = 0, BREM_BRT06 = 1, BREM_GP19 = 2 };

struct OpacityParams {
  bool use_dU, use_dm_eff, use_WM_ab, use_WM_sc, use_decay;
  BremImpl brem_implementation;
  bool use_NN_medium_corr, neglect_blocking;
};
static const OpacityParams opacity_params_default_all = {true, true, true, true, true,
                                                         BREM_HR98, true, true};
static const OpacityParams opacity_params_default_none = {false, false, false, false, false,
                                                          BREM_HR98, false, false};

struct MyEOSParams {
  BS_REAL nb, temp, ye, yp, yn, mu_p, mu_n, mu_e, mu_mu, dU, dm_eff;
};

struct NuDistributionParams {
  BS_REAL w_t[total_num_species], temp_t[total_num_species], eta_t[total_num_species];
  BS_REAL w_f[total_num_species], temp_f[total_num_species], c_f[total_num_species],
      beta_f[total_num_species];
};

struct M1Quantities {
  BS_REAL n[total_num_species];
  BS_REAL J[total_num_species];
  BS_REAL H[total_num_species][4];
  BS_REAL chi[total_num_species];
};

struct OpacityFlags {
  int use_abs_em, use_pair, use_brem, use_inelastic_scatt, use_iso;
};
static const OpacityFlags opacity_flags_default_all = {1, 1, 1, 1, 1};
static const OpacityFlags opacity_flags_default_none = {0, 0, 0, 0, 0};

struct GreyOpacityParams {
  OpacityParams opacity_pars;
  MyEOSParams eos_pars;
  NuDistributionParams distr_pars;
  M1Quantities m1_pars;
  OpacityFlags opacity_flags;
};

struct M1Opacities {
  BS_REAL eta_0[total_num_species];
  BS_REAL kappa_0_a[total_num_species];
  BS_REAL eta[total_num_species];
  BS_REAL kappa_a[total_num_species];
  BS_REAL kappa_s[total_num_species];
};

struct SpectralOpacities {  // bns_nurates.hpp:517-523
  BS_REAL j[total_num_species];        // emissivity
  BS_REAL kappa[total_num_species];    // absorptivity / inverse mean free path
  BS_REAL j_s[total_num_species];      // "emissivity" for scattering
  BS_REAL kappa_s[total_num_species];  // inverse mean free path for scattering
};

namespace bns_amd {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

namespace detail {

inline bnsamd_flags to_c(const OpacityFlags& f) {
  return bnsamd_flags{f.use_abs_em, f.use_pair, f.use_brem, f.use_inelastic_scatt, f.use_iso};
}

inline bnsamd_params to_c(const OpacityParams& p) {
  return bnsamd_params{p.use_dU,    p.use_dm_eff,          p.use_WM_ab,
                       p.use_WM_sc, p.use_decay,           static_cast<int>(p.brem_implementation),
                       p.use_NN_medium_corr, p.neglect_blocking};
}

// batches of points that share flags and parameters go through one launch
inline std::tuple<int, int, int, int, int, int, int, int, int, int, int, int, int> key_of(
    const GreyOpacityParams& g) {
  const OpacityFlags& f = g.opacity_flags;
  const OpacityParams& p = g.opacity_pars;
  return std::make_tuple(f.use_abs_em, f.use_pair, f.use_brem, f.use_inelastic_scatt, f.use_iso,
                         int(p.use_dU), int(p.use_dm_eff), int(p.use_WM_ab), int(p.use_WM_sc),
                         int(p.use_decay), int(p.brem_implementation),
                         int(p.use_NN_medium_corr), int(p.neglect_blocking));
}

// AoS <-> SoA (field f of point i at base[f * n + i], layouts of the C ABI)
inline void put_eos(const MyEOSParams& e, size_t i, size_t n, double* pts) {
  pts[BNSAMD_PT_NB * n + i] = e.nb;
  pts[BNSAMD_PT_TEMP * n + i] = e.temp;
  pts[BNSAMD_PT_YP * n + i] = e.yp;
  pts[BNSAMD_PT_YN * n + i] = e.yn;
  pts[BNSAMD_PT_MU_P * n + i] = e.mu_p;
  pts[BNSAMD_PT_MU_N * n + i] = e.mu_n;
  pts[BNSAMD_PT_MU_E * n + i] = e.mu_e;
  pts[BNSAMD_PT_DU * n + i] = e.dU;
  pts[BNSAMD_PT_DM_EFF * n + i] = e.dm_eff;
}

inline void put_m1(const M1Quantities& m, size_t i, size_t n, double* pts) {
  for (int s = 0; s < total_num_species; ++s) {
    pts[(BNSAMD_PT_N + s) * n + i] = m.n[s];
    pts[(BNSAMD_PT_J + s) * n + i] = m.J[s];
    pts[(BNSAMD_PT_CHI + s) * n + i] = m.chi[s];
  }
}

inline void get_nj(const double* pts, size_t i, size_t n, M1Quantities& m) {
  for (int s = 0; s < total_num_species; ++s) {
    m.n[s] = pts[(BNSAMD_PT_N + s) * n + i];
    m.J[s] = pts[(BNSAMD_PT_J + s) * n + i];
  }
}

inline void put_distr(const NuDistributionParams& d, size_t i, size_t n, double* di) {
  for (int s = 0; s < total_num_species; ++s) {
    di[(BNSAMD_DI_W_T + s) * n + i] = d.w_t[s];
    di[(BNSAMD_DI_TEMP_T + s) * n + i] = d.temp_t[s];
    di[(BNSAMD_DI_ETA_T + s) * n + i] = d.eta_t[s];
    di[(BNSAMD_DI_W_F + s) * n + i] = d.w_f[s];
    di[(BNSAMD_DI_TEMP_F + s) * n + i] = d.temp_f[s];
    di[(BNSAMD_DI_C_F + s) * n + i] = d.c_f[s];
    di[(BNSAMD_DI_BETA_F + s) * n + i] = d.beta_f[s];
  }
}

inline void get_distr(const double* di, size_t i, size_t n, NuDistributionParams& d) {
  for (int s = 0; s < total_num_species; ++s) {
    d.w_t[s] = di[(BNSAMD_DI_W_T + s) * n + i];
    d.temp_t[s] = di[(BNSAMD_DI_TEMP_T + s) * n + i];
    d.eta_t[s] = di[(BNSAMD_DI_ETA_T + s) * n + i];
    d.w_f[s] = di[(BNSAMD_DI_W_F + s) * n + i];
    d.temp_f[s] = di[(BNSAMD_DI_TEMP_F + s) * n + i];
    d.c_f[s] = di[(BNSAMD_DI_C_F + s) * n + i];
    d.beta_f[s] = di[(BNSAMD_DI_BETA_F + s) * n + i];
  }
}

inline void get_spectral(const double* o, size_t i, size_t n, SpectralOpacities& r) {
  for (int s = 0; s < 4; ++s) {
    r.j[s] = o[(BNSAMD_SP_J + s) * n + i];
    r.kappa[s] = o[(BNSAMD_SP_KAPPA + s) * n + i];
    r.j_s[s] = o[(BNSAMD_SP_J_S + s) * n + i];
    r.kappa_s[s] = o[(BNSAMD_SP_KAPPA_S + s) * n + i];
  }
}
inline void get_out(const double* o, size_t i, size_t n, M1Opacities& r) {
  for (int s = 0; s < total_num_species; ++s) {
    r.eta_0[s] = o[(BNSAMD_OUT_ETA_0 + s) * n + i];
    r.kappa_0_a[s] = o[(BNSAMD_OUT_KAPPA_0_A + s) * n + i];
    r.eta[s] = o[(BNSAMD_OUT_ETA + s) * n + i];
    r.kappa_a[s] = o[(BNSAMD_OUT_KAPPA_A + s) * n + i];
    r.kappa_s[s] = o[(BNSAMD_OUT_KAPPA_S + s) * n + i];
  }
}

}  // namespace detail

// One device, one stream, the resident quadrature tables of that device.
// Thread-safe (the C library serialises calls on a context).
class Context {
 public:
  explicit Context(int device = 0) {
    int rc = bnsamd_create(device, &ctx_);
    if (rc != BNSAMD_OK) throw Error(rc, "bnsamd_create failed on device " +
                                             std::to_string(device));
  }
  ~Context() { bnsamd_destroy(ctx_); }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;

  bnsamd_context* get() const { return ctx_; }
  void set_stream(void* hip_stream) { check(bnsamd_set_stream(ctx_, hip_stream)); }
  void* stream() const { return bnsamd_get_stream(ctx_); }
  void synchronize() { check(bnsamd_synchronize(ctx_)); }

  // Selects (building once) the resident tables of quad_1d and quad_2d.
  void SetQuadrature(const MyQuadrature* quad_1d, const MyQuadrature* quad_2d) {
    if (!quad_1d || !quad_2d) throw Error(BNSAMD_ERR_INVALID, "null quadrature");
    if (quad_1d->nx < 1 || quad_1d->nx > BS_N_MAX || quad_2d->nx < 1 || quad_2d->nx > BS_N_MAX)
      throw Error(BNSAMD_ERR_INVALID, "quadrature nx out of range");
    check(bnsamd_set_quadrature(ctx_, quad_1d->nx, quad_1d->points, quad_1d->w, quad_2d->nx,
                                quad_2d->points, quad_2d->w));
  }

  // The GP19 bremsstrahlung table that BREM_GP19 then interpolates
  // (bnsamd_set_gp19_table; the reference's GP19Table.hpp arrays, which its
  // tree does not ship); n_nb = 0 clears it (GP19 adds zero).
  void SetGP19Table(int n_nb, const double* nb, const double* lognb, int n_ye, const double* ye,
                    int n_T, const double* T, int n_w, const double* w, const double* logw,
                    const double* data) {
    check(bnsamd_set_gp19_table(ctx_, n_nb, nb, lognb, n_ye, ye, n_T, T, n_w, w, logw, data));
  }

  // Device SoA batch, asynchronous on the context stream.
  void ComputeM1OpacitiesDevice(int64_t npts, int64_t ld, const double* d_pts,
                                const double* d_distr, const OpacityFlags& flags,
                                const OpacityParams& params, bool stimulated, int precision,
                                double* d_out) {
    bnsamd_flags f = detail::to_c(flags);
    bnsamd_params p = detail::to_c(params);
    check(bnsamd_m1_opacities(ctx_, npts, ld, d_pts, d_distr, &f, &p, stimulated ? 1 : 0,
                              precision, d_out));
  }

  // Host batch in the reference's types.  Each point uses its own
  // distr_pars, opacity_flags and opacity_pars, as ComputeM1Opacities does.
  void ComputeM1Opacities(const MyQuadrature* quad_1d, const MyQuadrature* quad_2d,
                          const GreyOpacityParams* pts, size_t npts, M1Opacities* out,
                          bool stimulated = true, int precision = BNSAMD_FP64) {
    if (npts == 0) return;
    if (!pts || !out) throw Error(BNSAMD_ERR_INVALID, "null point or output array");
    SetQuadrature(quad_1d, quad_2d);
    std::map<decltype(detail::key_of(pts[0])), std::vector<size_t>> groups;
    for (size_t i = 0; i < npts; ++i) groups[detail::key_of(pts[i])].push_back(i);
    for (const auto& kv : groups) {
      const std::vector<size_t>& idx = kv.second;
      const size_t n = idx.size();
      std::vector<double> hp(n * BNSAMD_PT_NFIELDS), hd(n * BNSAMD_DI_NFIELDS),
          ho(n * BNSAMD_OUT_NFIELDS);
      for (size_t k = 0; k < n; ++k) {
        const GreyOpacityParams& g = pts[idx[k]];
        detail::put_eos(g.eos_pars, k, n, hp.data());
        detail::put_m1(g.m1_pars, k, n, hp.data());
        detail::put_distr(g.distr_pars, k, n, hd.data());
      }
      const GreyOpacityParams& g0 = pts[idx[0]];
      bnsamd_flags f = detail::to_c(g0.opacity_flags);
      bnsamd_params p = detail::to_c(g0.opacity_pars);
      check(bnsamd_m1_opacities_host(ctx_, int64_t(n), hp.data(), hd.data(), &f, &p,
                                     stimulated ? 1 : 0, precision, ho.data()));
      for (size_t k = 0; k < n; ++k) detail::get_out(ho.data(), k, n, out[idx[k]]);
    }
  }

  // Spectral opacities of npts points at nnu energies, device SoA buffers,
  // asynchronous on the context stream (bnsamd_spectral_opacities).
  void ComputeSpectralOpacitiesDevice(int64_t npts, int64_t ld, const double* d_pts,
                                      const double* d_distr, int64_t nnu, const double* d_nu,
                                      int64_t nu_ld, const OpacityFlags& flags,
                                      const OpacityParams& params, bool stimulated,
                                      double* d_out) {
    bnsamd_flags f = detail::to_c(flags);
    bnsamd_params p = detail::to_c(params);
    check(bnsamd_spectral_opacities(ctx_, npts, ld, d_pts, d_distr, nnu, d_nu, nu_ld, &f, &p,
                                    stimulated ? 1 : 0, d_out));
  }

  // The same into a device array of npts * nnu SpectralOpacities records,
  // d_out[i * nnu + e] for point i at energy e (bnsamd_spectral_opacities_aos).
  void ComputeSpectralOpacitiesDevice(int64_t npts, int64_t ld, const double* d_pts,
                                      const double* d_distr, int64_t nnu, const double* d_nu,
                                      int64_t nu_ld, const OpacityFlags& flags,
                                      const OpacityParams& params, bool stimulated,
                                      SpectralOpacities* d_out) {
    static_assert(sizeof(SpectralOpacities) == BNSAMD_SP_NFIELDS * sizeof(double),
                  "SpectralOpacities must be 16 packed doubles");
    bnsamd_flags f = detail::to_c(flags);
    bnsamd_params p = detail::to_c(params);
    check(bnsamd_spectral_opacities_aos(ctx_, npts, ld, d_pts, d_distr, nnu, d_nu, nu_ld, &f,
                                        &p, stimulated ? 1 : 0,
                                        reinterpret_cast<double*>(d_out)));
  }

  // Host batch in the reference's types: entry i is point pts[i] at energy
  // nu[i] [MeV], with its own distr_pars, opacity_flags and opacity_pars, as
  // the reference's per-(energy, point) calls (tests/tests.hpp:1174-1176).
  void ComputeSpectralOpacities(const MyQuadrature* quad_1d, const BS_REAL* nu,
                                const GreyOpacityParams* pts, size_t npts, SpectralOpacities* out,
                                bool stimulated) {
    if (npts == 0) return;
    if (!nu || !pts || !out) throw Error(BNSAMD_ERR_INVALID, "null energy, point or output array");
    SetQuadrature(quad_1d, quad_1d);
    std::map<decltype(detail::key_of(pts[0])), std::vector<size_t>> groups;
    for (size_t i = 0; i < npts; ++i) groups[detail::key_of(pts[i])].push_back(i);
    for (const auto& kv : groups) {
      const std::vector<size_t>& idx = kv.second;
      const size_t n = idx.size();
      std::vector<double> hp(n * BNSAMD_PT_NFIELDS), hd(n * BNSAMD_DI_NFIELDS), hn(n),
          ho(n * BNSAMD_SP_NFIELDS);
      for (size_t k = 0; k < n; ++k) {
        const GreyOpacityParams& g = pts[idx[k]];
        detail::put_eos(g.eos_pars, k, n, hp.data());
        detail::put_m1(g.m1_pars, k, n, hp.data());
        detail::put_distr(g.distr_pars, k, n, hd.data());
        hn[k] = nu[idx[k]];
      }
      const GreyOpacityParams& g0 = pts[idx[0]];
      bnsamd_flags f = detail::to_c(g0.opacity_flags);
      bnsamd_params p = detail::to_c(g0.opacity_pars);
      // one energy per point: nnu = 1, energies per point
      check(bnsamd_spectral_opacities_host(ctx_, int64_t(n), hp.data(), hd.data(), 1, hn.data(),
                                           1, &f, &p, stimulated ? 1 : 0, ho.data()));
      for (size_t k = 0; k < n; ++k) detail::get_spectral(ho.data(), k, n, out[idx[k]]);
    }
  }

  // CalculateDistrParamsFromM1 for a batch (reads m1_pars, writes distr_pars).
  void CalculateDistrParamsFromM1(GreyOpacityParams* pts, size_t npts) {
    if (npts == 0) return;
    std::vector<double> hp(npts * BNSAMD_PT_NFIELDS, 0.), hd(npts * BNSAMD_DI_NFIELDS);
    for (size_t i = 0; i < npts; ++i) detail::put_m1(pts[i].m1_pars, i, npts, hp.data());
    check(bnsamd_distr_from_m1_host(ctx_, int64_t(npts), hp.data(), hd.data()));
    for (size_t i = 0; i < npts; ++i) detail::get_distr(hd.data(), i, npts, pts[i].distr_pars);
  }

  // NuEquilibriumParams + ComputeM1DensitiesEq for a batch (reads eos_pars,
  // writes distr_pars and m1_pars.n, m1_pars.J).
  void SetEquilibrium(GreyOpacityParams* pts, size_t npts) {
    if (npts == 0) return;
    std::vector<double> hp(npts * BNSAMD_PT_NFIELDS, 0.), hd(npts * BNSAMD_DI_NFIELDS);
    for (size_t i = 0; i < npts; ++i) detail::put_eos(pts[i].eos_pars, i, npts, hp.data());
    check(bnsamd_equilibrium_host(ctx_, int64_t(npts), hp.data(), hd.data()));
    for (size_t i = 0; i < npts; ++i) {
      detail::get_distr(hd.data(), i, npts, pts[i].distr_pars);
      detail::get_nj(hp.data(), i, npts, pts[i].m1_pars);
    }
  }

  // ComputeM1DensitiesEq for one point's given distribution.
  void M1DensitiesEq(const MyEOSParams& eos, const NuDistributionParams& d, M1Quantities& m1) {
    std::vector<double> hp(BNSAMD_PT_NFIELDS, 0.), hd(BNSAMD_DI_NFIELDS);
    detail::put_eos(eos, 0, 1, hp.data());
    detail::put_distr(d, 0, 1, hd.data());
    check(bnsamd_m1_densities_eq_host(ctx_, 1, hd.data(), hp.data()));
    detail::get_nj(hp.data(), 0, 1, m1);
  }

 private:
  void check(int rc) const {
    if (rc != BNSAMD_OK) throw Error(rc, bnsamd_last_error(ctx_));
  }
  bnsamd_context* ctx_ = nullptr;
};

// The process-wide context of the per-point functions: device $BNSAMD_DEVICE
// (default 0), created on first use and never destroyed (its device memory
// goes with the process; no destructor runs after the HIP runtime's own).
inline Context& DefaultContext() {
  static Context* ctx = new Context([] {
    const char* e = std::getenv("BNSAMD_DEVICE");
    return e ? std::atoi(e) : 0;
  }());
  return *ctx;
}

}  // namespace bns_amd

// ---- per-point functions with the reference's signatures -----------------------

// GaussLegendre (integration.hpp:28-87): nodes and weights of quad->nx
// points on [x1, x2], ascending.
inline void GaussLegendre(MyQuadrature* quad) {
  if (!quad || quad->dim != 1 || quad->type != kGauleg || quad->nx < 1 || quad->nx > BS_N_MAX)
    throw bns_amd::Error(BNSAMD_ERR_INVALID, "GaussLegendre: 1D Gauss-Legendre, 1..BS_N_MAX");
  int rc = bnsamd_gauss_legendre(quad->nx, quad->x1, quad->x2, quad->points, quad->w);
  if (rc != BNSAMD_OK) throw bns_amd::Error(rc, "GaussLegendre failed");
}

// CalculateDistrParamsFromM1 (distribution.hpp:323-333)
inline NuDistributionParams CalculateDistrParamsFromM1(const M1Quantities* M1_pars,
                                                       const MyEOSParams* eos_pars) {
  (void)eos_pars;  // unused by the reconstruction, as in the reference
  GreyOpacityParams g{};
  g.m1_pars = *M1_pars;
  bns_amd::DefaultContext().CalculateDistrParamsFromM1(&g, 1);
  return g.distr_pars;
}

// NuEquilibriumParams (distribution.hpp:253-290)
inline NuDistributionParams NuEquilibriumParams(const MyEOSParams* eos_pars) {
  GreyOpacityParams g{};
  g.eos_pars = *eos_pars;
  bns_amd::DefaultContext().SetEquilibrium(&g, 1);
  return g.distr_pars;
}

// ComputeM1DensitiesEq (distribution.hpp:447-473)
inline void ComputeM1DensitiesEq(const MyEOSParams* eos_pars,
                                 const NuDistributionParams* nu_distribution_params,
                                 M1Quantities* m1_pars) {
  bns_amd::DefaultContext().M1DensitiesEq(*eos_pars, *nu_distribution_params, *m1_pars);
}

// ComputeM1Opacities (m1_opacities.hpp:1306-1313): one point
inline M1Opacities ComputeM1Opacities(const MyQuadrature* quad_1d, const MyQuadrature* quad_2d,
                                      GreyOpacityParams* my_grey_opacity_params) {
  M1Opacities out;
  bns_amd::DefaultContext().ComputeM1Opacities(quad_1d, quad_2d, my_grey_opacity_params, 1,
                                                &out, true);
  return out;
}

// ComputeM1OpacitiesNotStimulated (m1_opacities.hpp:1297-1304): one point
inline M1Opacities ComputeM1OpacitiesNotStimulated(MyQuadrature* quad_1d, MyQuadrature* quad_2d,
                                                   GreyOpacityParams* my_grey_opacity_params) {
  M1Opacities out;
  bns_amd::DefaultContext().ComputeM1Opacities(quad_1d, quad_2d, my_grey_opacity_params, 1,
                                                &out, false);
  return out;
}

// Batched forms: npts points, each with its own inputs, flags and parameters.
inline void ComputeM1Opacities(const MyQuadrature* quad_1d, const MyQuadrature* quad_2d,
                               const GreyOpacityParams* pts, size_t npts, M1Opacities* out) {
  bns_amd::DefaultContext().ComputeM1Opacities(quad_1d, quad_2d, pts, npts, out, true);
}

inline void ComputeM1OpacitiesNotStimulated(const MyQuadrature* quad_1d,
                                            const MyQuadrature* quad_2d,
                                            const GreyOpacityParams* pts, size_t npts,
                                            M1Opacities* out) {
  bns_amd::DefaultContext().ComputeM1Opacities(quad_1d, quad_2d, pts, npts, out, false);
}

// ComputeSpectralOpacitiesNotStimulatedAbs (m1_opacities.hpp:1505-1641): one
// point at energy nu [MeV], the distribution as given in distr_pars
inline SpectralOpacities ComputeSpectralOpacitiesNotStimulatedAbs(
    const BS_REAL nu, MyQuadrature* quad_1d, GreyOpacityParams* my_grey_opacity_params) {
  SpectralOpacities r;
  bns_amd::DefaultContext().ComputeSpectralOpacities(quad_1d, &nu, my_grey_opacity_params, 1,
                                                     &r, false);
  return r;
}

// ComputeSpectralOpacitiesStimulatedAbs (m1_opacities.hpp:1644-1660): one point
inline SpectralOpacities ComputeSpectralOpacitiesStimulatedAbs(
    const BS_REAL nu, MyQuadrature* quad_1d, GreyOpacityParams* my_grey_opacity_params) {
  SpectralOpacities r;
  bns_amd::DefaultContext().ComputeSpectralOpacities(quad_1d, &nu, my_grey_opacity_params, 1,
                                                     &r, true);
  return r;
}

// Batched forms: entry i is point pts[i] at energy nu[i].
inline void ComputeSpectralOpacitiesNotStimulatedAbs(const BS_REAL* nu,
                                                     const MyQuadrature* quad_1d,
                                                     const GreyOpacityParams* pts, size_t npts,
                                                     SpectralOpacities* out) {
  bns_amd::DefaultContext().ComputeSpectralOpacities(quad_1d, nu, pts, npts, out, false);
}

inline void ComputeSpectralOpacitiesStimulatedAbs(const BS_REAL* nu, const MyQuadrature* quad_1d,
                                                  const GreyOpacityParams* pts, size_t npts,
                                                  SpectralOpacities* out) {
  bns_amd::DefaultContext().ComputeSpectralOpacities(quad_1d, nu, pts, npts, out, true);
}

#endif  // BNS_NURATES_AMD_HPP_
