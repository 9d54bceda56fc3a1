// m1_host_api_check.cpp -- drives the header-only C++ API the way the
// reference's test driver drives ComputeM1Opacities
// (TestM1OpacitiesSelectedPoints, tests/tests.hpp:700-760): per point, fill
// GreyOpacityParams, reconstruct the distribution with
// CalculateDistrParamsFromM1, generate the quadrature with GaussLegendre and
// call ComputeM1Opacities(&quad, &quad, &params).  Then the same points once
// more through the batched overload.
//
//   m1_host_api_check <in.bin> <out.bin> <n_single>
//
// in.bin : int64 npts, int32 nquad, int32 stim, int32 flags[5], int32 params[8],
//          double pts[21][npts] (the C ABI point block)
// out.bin: double single[20][n_single] (per-point calls, first n_single
//          points), double batch[20][npts], double distr[28][npts]
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "bns_nurates_amd.hpp"

int main(int argc, char** argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: %s in.bin out.bin n_single\n", argv[0]);
    return 2;
  }
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  int64_t npts;
  int32_t nquad, stim, fl[5], pa[8];
  bool ok = std::fread(&npts, 8, 1, f) == 1 && std::fread(&nquad, 4, 1, f) == 1 &&
            std::fread(&stim, 4, 1, f) == 1 && std::fread(fl, 4, 5, f) == 5 &&
            std::fread(pa, 4, 8, f) == 8;
  std::vector<double> pts(ok ? size_t(npts) * 21 : 0);
  ok = ok && std::fread(pts.data(), 8, pts.size(), f) == pts.size();
  std::fclose(f);
  if (!ok) {
    std::fprintf(stderr, "bad input file\n");
    return 2;
  }
  const size_t n = size_t(npts);
  size_t n_single = size_t(std::atol(argv[3]));
  if (n_single > n) n_single = n;

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
      for (int s = 0; s < total_num_species; ++s) {
        g.m1_pars.n[s] = P(BNSAMD_PT_N + s);
        g.m1_pars.J[s] = P(BNSAMD_PT_J + s);
        g.m1_pars.chi[s] = P(BNSAMD_PT_CHI + s);
      }
    }

    std::vector<double> out(size_t(20) * (n_single + n) + size_t(28) * n);
    // per-point calls, as the reference's drivers make them
    for (size_t i = 0; i < n_single; ++i) {
      GreyOpacityParams g = gp[i];
      g.distr_pars = CalculateDistrParamsFromM1(&g.m1_pars, &g.eos_pars);
      M1Opacities o = stim ? ComputeM1Opacities(&quad, &quad, &g)
                           : ComputeM1OpacitiesNotStimulated(&quad, &quad, &g);
      for (int s = 0; s < 4; ++s) {
        out[(0 + s) * n_single + i] = o.eta_0[s];
        out[(4 + s) * n_single + i] = o.kappa_0_a[s];
        out[(8 + s) * n_single + i] = o.eta[s];
        out[(12 + s) * n_single + i] = o.kappa_a[s];
        out[(16 + s) * n_single + i] = o.kappa_s[s];
      }
    }
    // batched: reconstruct all, then one call
    bns_amd::DefaultContext().CalculateDistrParamsFromM1(gp.data(), n);
    std::vector<M1Opacities> ob(n);
    if (stim) ComputeM1Opacities(&quad, &quad, gp.data(), n, ob.data());
    else ComputeM1OpacitiesNotStimulated(&quad, &quad, gp.data(), n, ob.data());
    double* batch = out.data() + size_t(20) * n_single;
    double* distr = batch + size_t(20) * n;
    for (size_t i = 0; i < n; ++i) {
      for (int s = 0; s < 4; ++s) {
        batch[(0 + s) * n + i] = ob[i].eta_0[s];
        batch[(4 + s) * n + i] = ob[i].kappa_0_a[s];
        batch[(8 + s) * n + i] = ob[i].eta[s];
        batch[(12 + s) * n + i] = ob[i].kappa_a[s];
        batch[(16 + s) * n + i] = ob[i].kappa_s[s];
      }
      bns_amd::detail::put_distr(gp[i].distr_pars, i, n, distr);
    }
    FILE* g = std::fopen(argv[2], "wb");
    if (!g || std::fwrite(out.data(), 8, out.size(), g) != out.size()) return 2;
    std::fclose(g);
  } catch (const bns_amd::Error& e) {
    std::fprintf(stderr, "bns_amd::Error %d: %s\n", e.code(), e.what());
    return 1;
  }
  std::printf("m1_host_api_check: %zu points (%zu per-point calls), n=%d\n", n, n_single, nquad);
  return 0;
}
