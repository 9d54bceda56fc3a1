// spectral_kernel.hip -- spectral (energy-dependent) opacity kernel for CDNA4
// (gfx950), FP64.
//
// ComputeSpectralOpacitiesNotStimulatedAbs / ...StimulatedAbs
// (m1_opacities.hpp:1505-1660) evaluate, per thermodynamic point and
// neutrino energy nu, two Gauss-Legendre integrals over the partner energy
// nubar of SpectralIntegrand (:1316-1501): pair + bremsstrahlung with nodes
// t x_i and t / x_i at t = 2.182 T, and inelastic scattering on e-+ (NEPS)
// at t = nu; plus beta emission/absorption and isoenergetic scattering at nu.
//
// One wavefront computes one point for a chunk of E energies.  Everything
// that depends on the point alone (distribution parameters, F_k(+-eta), the
// HR98 channel constants, the pair nodes with their Fermi integrals
// F_1..F_5(eta -+ z) and distribution values) and on one energy alone
// (F_1..F_5(eta -+ y), F_1..F_5(-eta - y), TotalNuF(nu), beta, iso) is
// evaluated once into LDS.  Then the (energy, node) cells of each integral
// run one per lane per round -- 64 / 2n energies per round -- and the
// quadrature sums of each energy are formed from LDS in the reference's node
// order.  Each value is computed once where the reference recomputes the
// whole integrand for every one of the eight integrals
// (integration.hpp:511-546).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "gray_math.h"
#include "gray_tables.h"
#include "point_physics.h"
#include "spectral_kernel.h"

namespace bns {

namespace {

// Offsets (in doubles) of the dynamic LDS block of one workgroup:
//   quadrature   x_i, w_i, 1 / x_i (correctly rounded), log x_i (i < n)
//   node block   (pair/brem partner nodes a < N = 2n)
//     z = nubar / T, nubar, F_1..F_5(eta - z), F_1..F_5(eta + z),
//     TotalNuF(nubar, s) (the blocking factor 1 - f is formed from it)
//   energy block (e < E)
//     nu, y = nu / T, F_1..F_5(eta - y), F_1..F_5(eta + y), F_1..F_5(-eta - y),
//     TotalNuF(nu, s), beta em/ab (nue, anue), iso, the 8 pair/brem integrals,
//     the 8 NEPS integrals
//   cell block   8 integrand values of the 64 lanes of a round
struct SpLds {
  int qx, qw, qrx, qlx;
  int z, nub, fzm, fzp, gz;
  int nu, lnu, y, fym, fyp, fyq, gnu, bem, bab, iso, ip, in;
  int sf;
  int total;
};

__host__ __device__ inline SpLds sp_lds(int n, int E) {
  const int N = 2 * n;
  SpLds L;
  int o = 0;
  L.qx = o;  o += n;
  L.qw = o;  o += n;
  L.qrx = o; o += n;
  L.qlx = o; o += n;
  L.z = o;   o += N;
  L.nub = o; o += N;
  L.fzm = o; o += 5 * N;
  L.fzp = o; o += 5 * N;
  L.gz = o;  o += 4 * N;
  L.nu = o;  o += E;
  L.lnu = o; o += E;
  L.y = o;   o += E;
  L.fym = o; o += 5 * E;
  L.fyp = o; o += 5 * E;
  L.fyq = o; o += 5 * E;
  L.gnu = o; o += 4 * E;
  L.bem = o; o += 2 * E;
  L.bab = o; o += 2 * E;
  L.iso = o; o += E;
  L.ip = o;  o += 8 * E;
  L.in = o;  o += 8 * E;
  L.sf = o;  o += 8 * 64;
  L.total = o;
  return L;
}

// Absorption kernel of BremKernelsLegCoeff (l = 0), summed over the three
// channels, at x = (omega + omega') / T (kernel_brem_HR98.hpp:33-393), given
// the per-point scalars hp and channel constants hc
__device__ __forceinline__ double hr98_abs(double x, const HrPoint& hp,
                                           lds_cv<HrChan>* hc) {
  // the logarithms' arguments are positive and normal and the exponents
  // logarithms of O(1e-30..1e30) quantities scaled by O(1) powers: the
  // bounded forms, without the library's special-case handling
  const double xc = x > c::BremMin ? x : c::BremMin;  // BremKernelS clamps x
  const double y = hp.y > c::BremMin ? hp.y : c::BremMin;
  const double snum = 2. * c::SqrtPi * exp_bounded(1.5 * log_bounded(xc + 2. - hp.emy)) *
                      (xc * xc + 2. * xc * y + 5. * (y * y) / 3. + 1.);
  const double bb = c::Pi2OneEighth + xc + y;
  const double b2 = bb * bb;
  const double snd = div_normal(snum, c::SqrtPi + b2 * b2);
  const double lx = log_bounded(xc);
  const double px11 = exp_bounded(1.1 * lx), pxm4 = exp_bounded(-4. * lx);
  const double px093 = exp_bounded(0.93 * lx);
  const double px12 = exp_bounded(1.2 * lx), px28 = exp_bounded(2.8 * lx);
  const double lsdx = log_bounded((xc * xc + c::FourPiSquared) * xc /
                                  (c::FourPiSquared * (1. - safe_exp(-xc))));
  const double pnd = exp_bounded(-hp.p * log_bounded(snd));
  const double xm = xc - 1.2;
  const double fx = 3. + xm * xm + pxm4;
  // the factors the three channels share (as in the gray kernel)
  const double sdx = exp_bounded(-hp.p * lsdx);
  const double rfx = div_normal(1., fx);
  const double r28 = div_normal(30., 30. + 0.005 * px28);
  double tot = 0.;
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) {
    const double sd_pt = hc[ch].sd_pt, rf_pt = hc[ch].rf_pt, g0 = hc[ch].g0, h = hc[ch].h;
    const double gx = hc[ch].gamma_half_gb, got = hc[ch].gamma_over_T, wt = hc[ch].weight;
    double sd = sd_pt * sdx;
    double sbase = exp_bounded(hp.pinv * log_bounded(pnd + sd));
    double fb = 1. + rfx * rf_pt;
    double gbm = 1. - px11 * g0;
    double cb = div_normal(1.1 * px11 * h, 2.3 + h * px093 + 0.0001 * px12) * r28;
    double sb = sbase * fb * (1. + cb * gbm);
    tot += wt * (div_normal(got, x * x + gx * gx) * sb);
  }
  return 3. * c::BremConst * tot * hp.medium;
}

// WAVES: waves per SIMD the registers are budgeted for; NQ, ECH > 0: the
// quadrature order and energies per chunk fixed at compile time (the LDS
// offsets become constants instead of values held in registers), 0: taken
// from the arguments.
// GP19: the instance for GP19 bremsstrahlung over a table.
template <int WAVES, int NQ, int ECH, bool GP19 = false>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WAVES, WAVES))) void
k_spectral_f64(SpectralKernelArgs A) {
  extern __shared__ double sm[];
  __shared__ double s_in[64];  // the 21 inputs (+ 28 distribution values), 1 / T
  __shared__ Spec s_D[4];
  __shared__ double s_Fe[2][6], s_fd0[2];
  __shared__ HrChan s_hc[3];
  __shared__ double s_brem[6];  // BRT06 prefactor; HR98 y, e^-y/12, p, -1/p, medium
  // per-point constants of the beta and isoenergetic terms: EtaNNAbs n -> p,
  // p -> n, dQ + dU, mu_p + mu_e - mu_n; EtaNNSc of protons and neutrons
  __shared__ double s_bpt[4], s_etasc[2];

  const int lane = threadIdx.x;
  const int n = NQ > 0 ? NQ : A.n, N = 2 * n, E = ECH > 0 ? ECH : A.ech;
  // workgroup -> (point, group of cpw energy chunks); consecutive points of
  // one group share an XCD (and its L2), as in the gray kernel
  const int64_t b = xcd_point(blockIdx.x, gridDim.x);
  const int64_t p = b % A.npts, grp = b / A.npts;
  const int64_t c_begin = grp * A.cpw;
  const int64_t c_end = c_begin + A.cpw < A.nchunk ? c_begin + A.cpw : A.nchunk;
  const int64_t ld = A.ld;
  const SpLds L = sp_lds(n, E);
  lds_cv<double>* vin = lds_volatile(s_in);
  // the LDS block: written and read in phases separated by barriers
  double* V = sm;
  double* s_x = V + L.qx;
  double* s_w = V + L.qw;
  double* s_rx = V + L.qrx;  // 1 / x_i correctly rounded
  double* s_lx = V + L.qlx;  // log(x_i)

  // ---- inputs -----------------------------------------------------------------
  if (lane < 21) s_in[lane] = A.pts[lane * ld + p];
  else if (A.distr && lane < 49) s_in[lane] = A.distr[(lane - 21) * ld + p];
  if (lane < n) {
    s_x[lane] = A.q1[Q1_REC * lane + Q1_X];
    s_w[lane] = A.q1[Q1_REC * lane + Q1_W];
    s_rx[lane] = A.q1[Q1_REC * lane + Q1_RX];
    s_lx[lane] = log(A.q1[Q1_REC * lane + Q1_X]);
  }
  __syncthreads();
  const double T = vin[1], mu_e = vin[6];
  // 1 / T: energies / T are formed as div_by_recip(nu, T, 1 / T), which
  // rounds as nu / T does
  if (lane == 63) s_in[kInvT] = 1. / T;
  const double eta = mu_e / T;
  const bool do_pair = A.flags.use_pair != 0;
  // GP19 without a table adds zero, as the reference's table-free
  // configuration does (kernel_brem_GP19.hpp:483-493)
  const bool do_brem = A.flags.use_brem != 0 && (A.params.brem_impl != 2 || GP19);
  const bool do_pb = A.flags.use_pair != 0 || A.flags.use_brem != 0;
  const bool do_neps = A.flags.use_inelastic_scatt == 1;
  const bool blocking = A.params.neglect_blocking == 0;

  // ---- per-point constants ---------------------------------------------------------
  if (lane < 4) {
    const int s = lane;
    s_D[s] = A.distr ? spec_from_fields(vin[21 + 0 + s], vin[21 + 4 + s], vin[21 + 8 + s],
                                        vin[21 + 12 + s], vin[21 + 16 + s], vin[21 + 20 + s],
                                        vin[21 + 24 + s])
                     : spec_from_m1(vin[9 + s], vin[13 + s], vin[17 + s], A.gamma36);
  }
  {
    // F_0..F_5 at +eta (electrons) and -eta (positrons), FermiDistr(0, 1, +-eta)
    double F[6];
    fdi_bundle<0, 5>(lane == 0 ? eta : -eta, F);
    if (lane < 2) {
#pragma unroll
      for (int k = 0; k < 6; ++k) s_Fe[lane][k] = F[k];
      s_fd0[lane] = fermi_of_arg(lane == 0 ? -eta : eta);
    }
  }
  double brt_pre = 0.;
  HrPoint hp{0., 0., 1., -1., 1.};
  if (do_brem && A.params.brem_impl == 1) {
    // QBrem_BRT06 and the prefactor of BremKernelsBRT06 (kernel_brem_BRT06.hpp:24-59)
    const double nb = vin[0], yp = vin[2], yn = vin[3];
    double rho = nb * c::Mb;
    double q = 2.0778e+02 * 0.5 * c::MeV * (yn * yn + yp * yp + (28. / 3.) * yn * yp) *
               (rho * rho) * exp_any(5.5 * log_pos(T));
    double T2 = T * T, T7 = T2 * T2 * T2 * T;
    brt_pre = c::HC6FourPiSq * c::BremC4BRT06 * (q / T7);
  } else if (do_brem && A.params.brem_impl == 0) {
    hp = hr98_point(T, vin[0], A.params.use_NN_medium_corr != 0);
    const HrChan hc = hr98_channel(lane < 3 ? lane : 0, T, vin[0], vin[2], vin[3], hp.y, hp.p);
    if (lane < 3) s_hc[lane] = hc;
  }
  // the bremsstrahlung point constants stay in LDS and are read in the cell
  // rounds where used (held in registers through the chunk loop they spill)
  // beta_point / eta_nn_sc depend on the point only: once here, not per chunk
  if (lane == 1) {
    const PointParams& pp = A.params;
    const BetaPoint bp = beta_point(vin[0], T, vin[2], vin[3], vin[4], vin[5], mu_e, vin[7],
                                    vin[8], pp.use_dU != 0, pp.use_dm_eff != 0, pp.use_WM_ab != 0,
                                    pp.use_decay != 0);
    s_bpt[0] = bp.etanp;
    s_bpt[1] = bp.etapn;
    s_bpt[2] = bp.Qp;
    s_bpt[3] = bp.munu;
  } else if (lane == 2) {
    s_etasc[0] = eta_nn_sc(vin[0], T, vin[2]);
    s_etasc[1] = eta_nn_sc(vin[0], T, vin[3]);
  }
  if (lane == 0) {
    s_brem[0] = brt_pre;
    s_brem[1] = hp.y;
    s_brem[2] = hp.emy;
    s_brem[3] = hp.p;
    s_brem[4] = hp.pinv;
    s_brem[5] = hp.medium;
  }
  lds_cv<double>* vbrem = lds_volatile(s_brem);
  // read where used, each chunk: loads the compiler may not hoist out of the
  // chunk loop (held in registers through it they spill)
  lds_cv<double>* vbpt = lds_volatile(s_bpt);
  lds_cv<double>* vetasc = lds_volatile(s_etasc);
  __syncthreads();
  lds_cv<Spec>* vD = lds_volatile(s_D);
  auto Dget = [&](int s) {
    Spec r;
    r.wt = vD[s].wt;
    r.mut = vD[s].mut;
    r.invTt = vD[s].invTt;
    r.wf = vD[s].wf;
    r.bf = vD[s].bf;
    r.cf = vD[s].cf;
    r.invTf = vD[s].invTf;
    r.lbf = vD[s].lbf;
    return r;
  };
  lds_cv<double> (*vFe)[6] = (lds_cv<double> (*)[6])s_Fe;
  // anux with the distribution of nux: its TotalNuF values are copied
  const bool same23 = __builtin_amdgcn_readfirstlane(spec_same(Dget(2), Dget(3)) ? 1 : 0) != 0;
  const double tpair = c::PairScale * T;  // s_pair = 0.5 * 4.364 T (m1_opacities.hpp:1546)
  // partner node a of the pair/brem integral: t x_a or t / x_{a-n}
  // (integration.hpp:525-534), formed with the reference's rounding
  auto pair_node = [&](int a) {
    return a < n ? mul_rn(tpair, s_x[a]) : div_by_recip(tpair, s_x[a - n], s_rx[a - n]);
  };
  auto over_T = [&](double e) { return div_by_recip(e, T, vin[kInvT]); };
  // ---- pair/brem node tables (once per point): F_1..F_5 at eta -+ z_a,
  //      TotalNuF at the node and the blocking factor --------------------------
  {
    const int nbund = do_pair ? 2 * N : 0;
    const int ntask = do_pb ? nbund + N : 0;
    for (int t = lane; t < ntask; t += kWave) {
      if (t < nbund) {
        const int a = t < N ? t : t - N;
        const double z = over_T(pair_node(a));
        double F[6];
        fdi_bundle<1, 5>(t < N ? eta - z : eta + z, F);
        const int dst = (t < N ? L.fzm : L.fzp) + a;
#pragma unroll
        for (int k = 1; k <= 5; ++k) V[dst + (k - 1) * N] = F[k];
      } else {
        const int a = t - nbund;
        const double nub = pair_node(a);
        V[L.nub + a] = nub;
        V[L.z + a] = over_T(nub);
        const double lw = log_bounded(nub);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const double g = nuf(nub, lw, Dget(s));
          V[L.gz + s * N + a] = g;
        }
      }
    }
  }

  const double a1e = c::PairA1E * c::PairA1E, a2e = c::PairA2E * c::PairA2E;
  const double a1x = c::PairA1X * c::PairA1X, a2x = c::PairA2X * c::PairA2X;
  // energies per round of cells: a round takes one half of the partner nodes
  // (x_a or 1 / x_a) for Gh energies, so that its lanes' partner energies,
  // and with them the Fermi-integral forms and branches, stay closer
  const int Gh = kWave / n;

  for (int64_t chunk = c_begin; chunk < c_end; ++chunk) {
    const int64_t e0 = chunk * E;
    // the lane index as a value of this iteration: keeps the compiler from
    // hoisting the per-lane task indices and addresses of the energy phase
    // out of the chunk loop, where they would be held (and spilled) through it
    int cl = lane;
    __asm__ volatile("" : "+v"(cl));
    const int ne = static_cast<int>(A.nnu - e0 < E ? A.nnu - e0 : E);
    auto energy = [&](int e) {
      return A.nu_ld ? A.nu[(e0 + e) * A.nu_ld + p] : A.nu[e0 + e];
    };

    // ---- Fermi-integral bundles F_1..F_5 at eta - y, eta + y, -eta - y ---------
    for (int t = cl; t < 3 * ne; t += kWave) {
      const int e = t / 3, kind = t - 3 * e;
      const double y = over_T(energy(e));
      double F[6];
      fdi_bundle<1, 5>(kind == 0 ? eta - y : kind == 1 ? eta + y : -eta - y, F);
      const int dst = (kind == 0 ? L.fym : kind == 1 ? L.fyp : L.fyq) + e;
  #pragma unroll
      for (int k = 1; k <= 5; ++k) V[dst + (k - 1) * E] = F[k];
    }
    // ---- distribution values, beta and iso at the energies ----------------------
    // one task per (energy, species) for the distribution values and the
    // beta terms, one per energy for the isoenergetic term: the three kinds
    // diverge within a round, so each kind's cost is one task's
    {
      const PointParams& pp = A.params;
      for (int t = cl; t < 4 * ne; t += kWave) {
        // TotalNuF(nu) of species s (m1_opacities.hpp:1533-1536)
        const int e = t >> 2, sp = t & 3;
        const double nu = energy(e);
        // the caller's energies: the library logarithm for a zero, negative
        // or non-finite one
        double lw;
        if (nu >= 0x1p-1022 && nu <= 0x1.fffffffffffffp+1023) lw = log_bounded(nu);
        else lw = log(nu);
        if (sp == 0) {
          V[L.nu + e] = nu;
          V[L.y + e] = over_T(nu);
          V[L.lnu + e] = lw;
        }
        V[L.gnu + sp * E + e] = nuf(nu, lw, Dget(sp));
      }
      const int nbeta = 2 * ne;
      for (int t = cl; t < nbeta + ne; t += kWave) {
        if (t < nbeta) {
          // AbsOpacity (formulas_beta_iso.hpp:291-319) of nue (sp 0) or anue (1)
          const int e = t >> 1, sp = t & 1;
          double em = 0., ab = 0.;
          if (A.flags.use_abs_em != 0) {
            BetaPoint bp;
            bp.etanp = vbpt[0];
            bp.etapn = vbpt[1];
            bp.Qp = vbpt[2];
            bp.munu = vbpt[3];
            bp.mu_e = mu_e;
            bp.T = T;
            bp.rT = vin[kInvT];
            bp.wm = pp.use_WM_ab != 0;
            bp.decay = pp.use_decay != 0;
            beta_em_ab(energy(e), sp, bp, em, ab);
          }
          V[L.bem + sp * E + e] = em;
          V[L.bab + sp * E + e] = ab;
        } else {
          // IsoScattLegCoeff, l = 0 (formulas_beta_iso.hpp:524-560)
          const int e = t - nbeta;
          double iso = 0.;
          if (A.flags.use_iso != 0) {
            const double nu = energy(e);
            double r0p = 1., r0n = 1.;
            if (pp.use_WM_sc) {
              r0p = wm_scatt_r0(nu, 1);
              r0n = wm_scatt_r0(nu, 2);
            }
            iso = vetasc[0] * c::IsoC0p * r0p + vetasc[1] * c::IsoC0n * r0n;
          }
          V[L.iso + e] = iso;
        }
      }
    }
    __syncthreads();

    // per-energy quadrature sums over one half h of the nodes
    // (GaussLegendreIntegrate1D, integration.hpp:511-546): lane (g, k) sums
    // integrand k of energy r0 + g over the half in node order (DoIntegration,
    // integration.hpp:95-108), loads in batches of four so that their LDS
    // latencies overlap; half 0 parks its sum s1 in the destination, half 1
    // stores t (s1 + s2)
    auto reduce_half = [&](int r0, int dst, bool neps, int h) {
      for (int j0 = 0; j0 < 8 * Gh; j0 += kWave) {
        const int j = j0 + lane;
        const int g = j >> 3, k = j & 7, e = r0 + g;
        if (j < 8 * Gh && e < ne) {
          const double* f = V + L.sf + k * kWave + g * n;
          double s = 0.;
          int i = 0;
          for (; i + 4 <= n; i += 4) {
            const double a0 = f[i], a1 = f[i + 1], a2 = f[i + 2], a3 = f[i + 3];
            s = fma(s_w[i], a0, s);
            s = fma(s_w[i + 1], a1, s);
            s = fma(s_w[i + 2], a2, s);
            s = fma(s_w[i + 3], a3, s);
          }
          for (; i < n; ++i) s = fma(s_w[i], f[i], s);
          double* d = V + dst + k * E + e;
          if (h == 0) {
            *d = s;
          } else {
            const double t = neps ? V[L.nu + e] : tpair;
            *d = t * (*d + s);
          }
        }
      }
    };

    // ====================== pair + bremsstrahlung integral ========================
    if (do_pb) {
      for (int r0 = 0; r0 < ne; r0 += Gh) for (int h = 0; h < 2; ++h) {
        const int g = lane / n, a = lane - g * n + h * n, e = r0 + g;
        if (g < Gh && e < ne) {
          const double nu = V[L.nu + e], y = V[L.y + e];
          const double nub = V[L.nub + a], z = V[L.z + a];
          double em[4] = {0., 0., 0., 0.}, ab[4] = {0., 0., 0., 0.};
          if (do_pair) {
            // PairKernels with cos_theta = 1, lmax = 0 (kernel_pair.hpp:135-198)
            double psi0 = 0., psi1 = 0.;
            const double emyz = (eta - y) - z;
            if (!(emyz > 200.)) {
              // PairPsi (kernel_pair.hpp:16-120)
              double Fm3[6], Fp3[6];
              fdi_bundle<3, 5>(emyz, Fm3);
              fdi_bundle<3, 5>((eta + y) + z, Fp3);
              const int ey = e, az = a;
              auto fym = [&](int k) { return V[L.fym + (k - 1) * E + ey]; };
              auto fyp = [&](int k) { return V[L.fyp + (k - 1) * E + ey]; };
              auto fzm = [&](int k) { return V[L.fzm + (k - 1) * N + az]; };
              auto fzp = [&](int k) { return V[L.fzp + (k - 1) * N + az]; };
              double x0 = 20. * fym(4), x1 = 20. * fzp(4);
              double x2 = 120. * (fym(2) - fzp(2));
              double x3 = 40. * (fzp(3) - fym(3));
              double x4 = 40. * vFe[0][3];
              double x5 = 40. * Fm3[3] - x4;
              double x6 = -20. * vFe[0][4];
              double x7 = 20. * Fm3[4] + x6;
              double x8 = x4 - 40. * Fp3[3];
              double x9 = 20. * Fp3[4] + x6;
              double x10 = 4. * (((((Fm3[5] - fym(5)) - fzm(5)) + fyp(5)) - Fp3[5]) + fzp(5));
              double x11 = 20. * fyp(4), x12 = 20. * fzm(4);
              double x13 = 40. * (fyp(3) - fzm(3));
              double x14 = 120. * (fzm(2) - fyp(2));
              double yz = y * z;
              double aux15 = 15. * (yz * yz);
              psi0 = x10 +
                     y * (-x0 + x1 + x7 +
                          y * (x3 + x5 + z * (x2 + z * (120. * (fzp(1) - fym(1))))) +
                          z * (80. * (fym(3) - fzp(3)) - x2 * z)) +
                     z * (x0 - x1 + x9 + z * (x3 + x8));
              psi1 = x10 + y * (-x11 + x12 + x9 + y * (x13 + x8)) +
                     z * (x11 - x12 + x7 + y * (80. * (fzm(3) - fyp(3)) - x14 * y) +
                          z * (x13 + x5 + y * (x14 + y * (120. * (fyp(1) - fzm(1))))));
              const double ia = div_normal(1., aux15);
              psi0 *= ia;
              psi1 *= ia;
            }
            // PairPhi: Phi_0 = (a1^2 psi0 + a2^2 psi1) C T^2 / (1 - e^{y+z}),
            // em = Phi_0 / 2, abs = e^{(w+w')/T} em
            const double aux = c::PairPhiC * (T * T) / (1. - safe_exp(y + z));
            const double db = safe_exp(over_T(nu + nub));
            em[0] = 0.5 * ((a1e * psi0 + a2e * psi1) * aux);
            em[1] = 0.5 * ((a1e * psi1 + a2e * psi0) * aux);
            em[2] = 0.5 * ((a1x * psi0 + a2x * psi1) * aux);
            em[3] = 0.5 * ((a1x * psi1 + a2x * psi0) * aux);
  #pragma unroll
            for (int s = 0; s < 4; ++s) ab[s] = db * em[s];
          }
          double se = 0., sa = 0.;
          if (do_brem) {
            if constexpr (GP19) {
              // BremKernelAbsGP19 (m1_opacities.hpp:1386-1390), Ye taken as yp
              gp19_brem(*A.gp19, nu, nub, vin[0], vin[2], T, se, sa);
            } else if (A.params.brem_impl == 1) {
              // BremKernelsBRT06 at x = (w + w') / (2T) (kernel_brem_BRT06.hpp:42-69)
              const double xb = over_T(0.5 * (nu + nub));
              const double k1x = bessk1_fit(xb) / xb;
              const double pre = vbrem[0];
              se = pre * k1x * safe_exp(-xb);
              sa = pre * k1x * safe_exp(xb);
            } else {
              const double x = over_T(nu + nub);
              const HrPoint hpl{vbrem[1], vbrem[2], vbrem[3], vbrem[4], vbrem[5]};
              sa = hr98_abs(x, hpl, lds_volatile(s_hc));
              se = sa * safe_exp(-x);
            }
          }
          // production: partner blocking; annihilation: partner occupation
          // (SpectralIntegrand, m1_opacities.hpp:1421-1496)
          const double nb2 = nub * nub;
          const double jac = a < n ? 1. : s_rx[a - n] * s_rx[a - n];
  #pragma unroll
          for (int s = 0; s < 4; ++s) {
            const int ps = s ^ 1;
            const double pro = (em[s] + se) * (blocking ? 1. - V[L.gz + ps * N + a] : 1.);
            const double ann = (ab[s] + sa) * V[L.gz + ps * N + a];
            V[L.sf + s * kWave + lane] = (nb2 * pro) * jac;
            V[L.sf + (4 + s) * kWave + lane] = (nb2 * ann) * jac;
          }
        }
        __syncthreads();
        reduce_half(r0, L.ip, false, h);
        __syncthreads();
      }
    } else {
      for (int j = lane; j < 8 * E; j += kWave) V[L.ip + j] = 0.;
    }

    // ================= inelastic scattering on e-+ (NES + NPS) =====================
    if (do_neps) {
      const double pre = c::NEPSConst * (T * T);
      const double bpl = c::NEPSBPlus, bze = c::NEPSBZero, bmi = c::NEPSBMinus;
      for (int r0 = 0; r0 < ne; r0 += Gh) for (int h = 0; h < 2; ++h) {
        const int g = lane / n, a = lane - g * n + h * n, e = r0 + g;
        if (g < Gh && e < ne) {
          const double nu = V[L.nu + e], w = V[L.y + e];
          const double xq = s_x[a < n ? a : a - n];
          // partner energy nu x or nu / x (t = nu, m1_opacities.hpp:1547)
          const double nub = a < n ? mul_rn(nu, xq) : div_by_recip(nu, xq, s_rx[a - n]);
          const double wp = over_T(nub);
          const double ad = fabs(w - wp);
          const double xm = w > wp ? w : wp, ym = w < wp ? w : wp;
          const double sgn = std::signbit(wp - w) ? 1. : -1.;
          double out[4] = {0., 0., 0., 0.};
  #pragma unroll
          for (int lep = 0; lep < 2; ++lep) {
            // NESKernels (lep 0, target e-) / NPSKernels (lep 1, target e+),
            // kernel_nes.hpp:139-340; couplings (b1, b2) per species, NPS
            // swaps them
            const double el = lep == 0 ? eta : -eta;
            double b1[4], b2[4];
            b1[0] = lep == 0 ? bpl : bze;  b2[0] = lep == 0 ? bze : bpl;
            b1[1] = lep == 0 ? bze : bpl;  b2[1] = lep == 0 ? bpl : bze;
            b1[2] = lep == 0 ? bmi : bze;  b2[2] = lep == 0 ? bze : bmi;
            b1[3] = lep == 0 ? bze : bmi;  b2[3] = lep == 0 ? bmi : bze;
            if (ym > el * 1e-3) {
              // MezzacappaIntOut (kernel_nes.hpp:50-78) with the differences of
              // ComputeFDIForInelastic (:29-44); F_k(el - w) per energy from LDS
              double A5[6], C3[6];
              fdi_bundle<1, 5>(el - wp, A5);
              fdi_bundle<3, 5>(el - ad, C3);
              const int fb = (lep == 0 ? L.fym : L.fyq) + e;
              double fw0 = A5[1] - V[fb], fw1 = A5[2] - V[fb + E], fw2 = A5[3] - V[fb + 2 * E];
              double fw3 = A5[4] - V[fb + 3 * E], fw4 = A5[5] - V[fb + 4 * E];
              double fa0 = vFe[lep][3] - C3[3], fa1 = vFe[lep][4] - C3[4];
              double fa2 = vFe[lep][5] - C3[5];
              double ws = w + wp;
              double inv = div_normal(1., w * w * wp * wp);
              double q4 = fw3 + 2. * (ws * fw2 + 3. * w * wp * fw1);
              // the out-kernel is linear in the couplings: b1 U + b2 V
              double t1 = (sgn * fa2 - fw4) * 0.2;
              double U = (t1 - ws * q4 + sgn * (xm * fa1 + 2. * (xm * xm) * fa0)) * inv - 6. * fw0;
              double Vv = (t1 + sgn * (2. * (ym * ym) * fa0 - ym * fa1)) * inv;
              U *= pre;
              Vv *= pre;
  #pragma unroll
              for (int s = 0; s < 4; ++s) out[s] = fma(b1[s], U, fma(b2[s], Vv, out[s]));
            } else if (xm > el * 1e-3) {
              // MezzacappaIntOneEnergy (kernel_nes.hpp:82-107)
              const double xq = opaque(xm), yq = opaque(ym);
              double D3[6];
              fdi_bundle<0, 2>(el - xq, D3);
              double f0 = D3[2] - vFe[lep][2], f1 = D3[1], f2 = vFe[lep][1];
              double f3 = D3[0] + yq * fermi_of_arg(xq - el) / 6.;
              double f4 = vFe[lep][0] - yq * s_fd0[lep] / 6.;
              double ix2 = 1. / (xq * xq);
  #pragma unroll
              for (int s = 0; s < 4; ++s) {
                double B1 = b1[s], B2 = b2[s];
                double S =
                    2. * (B1 + B2) * f0 + (B1 * (yq + 4. * xq) + 3. * B2 * yq) * f1 +
                    (B1 * (3. * yq - 4. * xq) + B2 * yq) * f2 +
                    ((B1 + 6. * B2) * yq * yq * 0.2 + B1 * xq * yq + 2. * B1 * xq * xq) * f3 -
                    ((6. * B1 + B2) * yq * yq * 0.2 - 3. * B1 * xq * yq + 2. * B1 * xq * xq) * f4;
                out[s] = fma(pre, -sgn * yq * S * ix2, out[s]);
              }
            } else {
              // MezzacappaIntTwoEnergies (kernel_nes.hpp:111-132)
              const double xq = opaque(xm), yq = opaque(ym);
              double f0 = s_fd0[lep], f1 = vFe[lep][0], f2 = vFe[lep][1];
              double ix2 = 1. / (xq * xq);
              double pol = (11. * yq * yq - 25. * xq * yq + 20. * xq * xq) / 30.;
  #pragma unroll
              for (int s = 0; s < 4; ++s) {
                double S = (b1[s] + b2[s]) * (4. * f2 + f0 * pol) +
                           (b1[s] - b2[s]) * (2. * xq - yq) * f1;
                out[s] = fma(pre, yq * (w - wp) * S * ix2, out[s]);
              }
            }
          }
          // detailed balance of the in/out kernels (kernel_nes.hpp:229-234)
          double ef, efx;
          neps_exp_func_pair(wp - w, w - wp, ef, efx);
          // distribution at nubar, formed after the Fermi integrals so that it
          // is not live through them: production weighted by the occupation,
          // annihilation by the blocking factor of the same species
          // (m1_opacities.hpp:1430-1433, 1481-1484)
          // log(nu x) or log(nu / x) from the tabled logarithms
          const double lw = V[L.lnu + e] + (a < n ? s_lx[a] : -s_lx[a - n]);
          const double nb2 = nub * nub;
          const double jac = a < n ? 1. : s_rx[a - n] * s_rx[a - n];
          double gb[4];
#pragma unroll
          for (int s = 0; s < 3; ++s) gb[s] = nuf(nub, lw, Dget(s));
          gb[3] = same23 ? gb[2] : nuf(nub, lw, Dget(3));
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            const double bf = blocking ? 1. - gb[s] : 1.;
            const double em = -out[s] * efx, ab = out[s] * ef;
            V[L.sf + s * kWave + lane] = (nb2 * (em * gb[s])) * jac;
            V[L.sf + (4 + s) * kWave + lane] = (nb2 * (ab * bf)) * jac;
          }
        }
        __syncthreads();
        reduce_half(r0, L.in, true, h);
        __syncthreads();
      }
    } else {
      for (int j = lane; j < 8 * E; j += kWave) V[L.in + j] = 0.;
    }
    __syncthreads();

    // ============================ assemble + store ============================
    // ComputeSpectralOpacitiesNotStimulatedAbs (m1_opacities.hpp:1572-1641);
    // the stimulated form adds j / c to kappa (:1644-1660).  Field-major
    // output: lane t writes field t / ne of energy t % ne, one 8-byte value
    // per row of the block; struct output: the chunk's 16 ne values are one
    // contiguous run, lane t writes field t % 16 of energy t / 16.
    {
      const double F = c::FourPiHc3, cl = c::Clight, four_pi = 4. * c::Pi;
      const bool aos = A.aos != 0;
      for (int t = lane; t < 16 * ne; t += kWave) {
        const int f = aos ? t & 15 : t / ne, e = aos ? t >> 4 : t - f * ne;
        const int s = f & 3, kind = f >> 2;
        const double nu = V[L.nu + e];
        const double emb = s < 2 ? V[L.bem + s * E + e] : 0.;
        const double abb = s < 2 ? V[L.bab + s * E + e] : 0.;
        const double j = emb + F * (V[L.ip + s * E + e] + V[L.in + s * E + e]);
        double kap = (abb + F * (V[L.ip + (4 + s) * E + e] + V[L.in + (4 + s) * E + e])) / cl;
        const double g = V[L.gnu + s * E + e], iso = V[L.iso + e];
        const double js = four_pi * (nu * nu) * g * iso;
        double ks = four_pi * (nu * nu) * (1. - g) * iso / cl;
        if (A.stim == 1) {
          kap = j / cl + kap;
          ks = js / cl + ks;
        }
        const double v = kind == 0 ? j : kind == 1 ? kap : kind == 2 ? js : ks;
        if (aos) A.out[(p * A.nnu + e0 + e) * 16 + f] = v;
        else A.out[(f * A.nnu + e0 + e) * ld + p] = v;
      }
    }
    __syncthreads();  // the next chunk rewrites the energy block
  }  // chunk
}

}  // namespace

// The instance for the reference's default order, n = 10: 6 energies per
// chunk keep the workgroup's LDS (1 KB static + 9 KB) within the 10 KB that
// 4 waves per SIMD allow, and the compile-time layout keeps it within 128
// VGPRs without spills.  Other orders: 3 waves, up to four rounds per chunk.
constexpr int kFixedN = 10, kFixedE = 6;

int spectral_energies_per_chunk(int n) {
  if (n == kFixedN) return kFixedE;
  const int G = kWave / (2 * n);
  return std::min(4 * G, 16);
}

int64_t spectral_chunks_per_group(int64_t npts, int64_t nchunk) {
  // 256 CUs x 12 resident workgroups (3 waves per SIMD); ask for at least 8
  // workgroups per slot so that the tail of the grid is short
  constexpr int64_t kWanted = 256 * 12 * 8;
  const int64_t groups = npts * nchunk;
  int64_t cpw = groups / kWanted;
  if (cpw < 1) cpw = 1;
  if (cpw > nchunk) cpw = nchunk;
  return cpw;
}

size_t spectral_lds_bytes(int n, int ech) {
  return sizeof(double) * static_cast<size_t>(sp_lds(n, ech).total);
}

hipError_t launch_spectral_f64(const SpectralKernelArgs& a, hipStream_t stream) {
  if (a.npts <= 0 || a.nnu <= 0) return hipSuccess;
  const int64_t groups = a.npts * ((a.nchunk + a.cpw - 1) / a.cpw);
  dim3 grid(static_cast<unsigned>(groups)), block(kWave);
  const size_t lds = spectral_lds_bytes(a.n, a.ech);
  if (a.gp19 && a.params.brem_impl == 2 && a.flags.use_brem != 0)  // GP19 over a table
    hipLaunchKernelGGL((k_spectral_f64<3, 0, 0, true>), grid, block, lds, stream, a);
  else if (a.n == kFixedN && a.ech == kFixedE)
    hipLaunchKernelGGL((k_spectral_f64<4, kFixedN, kFixedE>), grid, block, lds, stream, a);
  else
    hipLaunchKernelGGL((k_spectral_f64<3, 0, 0>), grid, block, lds, stream, a);
  return hipGetLastError();
}

}  // namespace bns
