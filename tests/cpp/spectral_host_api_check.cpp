// spectral_host_api_check.cpp -- drives the header-only C++ API the way the
// reference's spectral test drives ComputeSpectralOpacitiesNotStimulatedAbs
// (TestSpectralOpacities, tests/tests.hpp:861-1261): per point the EOS, the
// equilibrium distribution (NuEquilibriumParams), its M1 densities
// (ComputeM1DensitiesEq) and chi = 1/3; then one call per (energy, point)
// with the quadrature from GaussLegendre.  Then the same (energy, point)
// pairs once more through the batched overload.
//
//   spectral_host_api_check <in.bin> <out.bin>
//
// in.bin : int64 npts, int64 nnu, int32 nquad, int32 stim, int32 flags[5],
//          int32 params[8], double pts[21][npts] (the C ABI point block; only
//          its EOS fields are read), double nu[nnu]
// out.bin: double single[16][nnu][npts] (per-point calls),
//          double batch[16][nnu][npts], double distr[28][npts]
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "bns_nurates_amd.hpp"

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]);
    return 2;
  }
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  int64_t npts, nnu;
  int32_t nquad, stim, fl[5], pa[8];
  bool ok = std::fread(&npts, 8, 1, f) == 1 && std::fread(&nnu, 8, 1, f) == 1 &&
            std::fread(&nquad, 4, 1, f) == 1 && std::fread(&stim, 4, 1, f) == 1 &&
            std::fread(fl, 4, 5, f) == 5 && std::fread(pa, 4, 8, f) == 8;
  std::vector<double> pts(ok ? size_t(npts) * 21 : 0), nu(ok ? size_t(nnu) : 0);
  ok = ok && std::fread(pts.data(), 8, pts.size(), f) == pts.size() &&
       std::fread(nu.data(), 8, nu.size(), f) == nu.size();
  std::fclose(f);
  if (!ok) {
    std::fprintf(stderr, "bad input file\n");
    return 2;
  }
  const size_t n = size_t(npts), ne = size_t(nnu);

  try {
    MyQuadrature quad = quadrature_default;
    quad.nx = nquad;
    GaussLegendre(&quad);

    std::vector<GreyOpacityParams> gp(n);
    for (size_t i = 0; i < n; ++i) {
      GreyOpacityParams& g = gp[i];
      g.opacity_flags = {fl[0], fl[1], fl[2], fl[3], fl[4]};
      g.opacity_pars = {pa[0] != 0, pa[1] != 0, pa[2] != 0, pa[3] != 0, pa[4] != 0,
                        static_cast<BremImpl>(pa[5]), pa[6] != 0, pa[7] != 0};
      auto P = [&](int fld) { return pts[size_t(fld) * n + i]; };
      g.eos_pars.nb = P(BNSAMD_PT_NB);
      g.eos_pars.temp = P(BNSAMD_PT_TEMP);
      g.eos_pars.yp = P(BNSAMD_PT_YP);
      g.eos_pars.yn = P(BNSAMD_PT_YN);
      g.eos_pars.ye = P(BNSAMD_PT_YP);
      g.eos_pars.mu_p = P(BNSAMD_PT_MU_P);
      g.eos_pars.mu_n = P(BNSAMD_PT_MU_N);
      g.eos_pars.mu_e = P(BNSAMD_PT_MU_E);
      g.eos_pars.mu_mu = 0.;
      g.eos_pars.dU = P(BNSAMD_PT_DU);
      g.eos_pars.dm_eff = P(BNSAMD_PT_DM_EFF);
      // tests/tests.hpp:1160-1172
      g.distr_pars = NuEquilibriumParams(&g.eos_pars);
      ComputeM1DensitiesEq(&g.eos_pars, &g.distr_pars, &g.m1_pars);
      for (int s = 0; s < total_num_species; ++s) g.m1_pars.chi[s] = 1. / 3.;
    }

    std::vector<double> out(size_t(32) * ne * n + size_t(28) * n);
    double* single = out.data();
    double* batch = single + size_t(16) * ne * n;
    double* distr = batch + size_t(16) * ne * n;
    auto put = [&](double* base, size_t e, size_t i, const SpectralOpacities& r) {
      for (int s = 0; s < 4; ++s) {
        base[((BNSAMD_SP_J + s) * ne + e) * n + i] = r.j[s];
        base[((BNSAMD_SP_KAPPA + s) * ne + e) * n + i] = r.kappa[s];
        base[((BNSAMD_SP_J_S + s) * ne + e) * n + i] = r.j_s[s];
        base[((BNSAMD_SP_KAPPA_S + s) * ne + e) * n + i] = r.kappa_s[s];
      }
    };
    // per (energy, point) calls, as the reference's driver makes them
    for (size_t e = 0; e < ne; ++e) {
      for (size_t i = 0; i < n; ++i) {
        GreyOpacityParams g = gp[i];
        SpectralOpacities r = stim ? ComputeSpectralOpacitiesStimulatedAbs(nu[e], &quad, &g)
                                   : ComputeSpectralOpacitiesNotStimulatedAbs(nu[e], &quad, &g);
        put(single, e, i, r);
      }
    }
    // batched: all (energy, point) pairs in one call
    std::vector<GreyOpacityParams> pairs;
    std::vector<double> pnu;
    for (size_t e = 0; e < ne; ++e)
      for (size_t i = 0; i < n; ++i) {
        pairs.push_back(gp[i]);
        pnu.push_back(nu[e]);
      }
    std::vector<SpectralOpacities> ob(pairs.size());
    if (stim)
      ComputeSpectralOpacitiesStimulatedAbs(pnu.data(), &quad, pairs.data(), pairs.size(),
                                            ob.data());
    else
      ComputeSpectralOpacitiesNotStimulatedAbs(pnu.data(), &quad, pairs.data(), pairs.size(),
                                               ob.data());
    for (size_t e = 0; e < ne; ++e)
      for (size_t i = 0; i < n; ++i) put(batch, e, i, ob[e * n + i]);
    for (size_t i = 0; i < n; ++i) bns_amd::detail::put_distr(gp[i].distr_pars, i, n, distr);
    FILE* g = std::fopen(argv[2], "wb");
    if (!g || std::fwrite(out.data(), 8, out.size(), g) != out.size()) return 2;
    std::fclose(g);
  } catch (const bns_amd::Error& e) {
    std::fprintf(stderr, "bns_amd::Error %d: %s\n", e.code(), e.what());
    return 1;
  }
  std::printf("spectral_host_api_check: %zu points x %zu energies (%zu per-point calls), n=%d\n",
              n, ne, n * ne, nquad);
  return 0;
}
