// gray_kernel.hip -- fused gray M1 opacity kernel for CDNA4 (gfx950), FP64.
//
// One wavefront computes one thermodynamic point
// (ComputeM1Opacities, m1_opacities.hpp:1062-1313).  Its 64 lanes share the
// work of that point:
//   1D   isoenergetic scattering and beta processes: one (species, energy)
//        item per lane, wave-reduced;
//   2D   node tables (Fermi integrals F_1..F_5 at eta -+ y_a, distribution
//        values) staged in LDS, then one pair/bremsstrahlung cell (a <= b)
//        per lane per round, both orderings (a,b) and (b,a) accumulated;
//   NEPS one half-cell (i, j, half) per lane per round;
// and the 16 accumulators of the 2D sums are reduced across the wavefront at
// the end.  The cell geometry is point-independent (gray_tables.h) and read
// from a resident table; per-point constants are computed once per wave.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "gp19.h"
#include "gray_kernel.h"
#include "gray_math.h"
#include "point_physics.h"

namespace bns {

// slot of eta = mu_e / T in the per-point LDS block (after the 21 inputs and
// 28 distribution values, before log T and 1 / T), and of e^-eta (eta >= 0)
constexpr int kEtaSlot = 61;
constexpr int kEmEtaSlot = 60;
constexpr int kNpOffSlot = 59;

// ---------------------------------------------------------------------------
// The body of both kernel variants: MIX = false is the FP64 path, MIX = true
// the mixed-precision path (FP32 distribution values and HR98 channel
// functions, everything else FP64).  STIM (stimulated absorption,
// A.stim == 1) and BLOCK (partner blocking, neglect_blocking == 0) are
// fixed per launch, so that a cell evaluates only its own weighting.
// NQ > 0: both quadrature orders fixed at NQ at compile time (table
// offsets and loop bounds are constants); 0: taken from the arguments.
// GP19: the instance for GP19 bremsstrahlung over a table (its interpolation
// and extrapolations kept out of the other instances' register budget).
template <bool MIX, bool STIM, bool BLOCK, int NQ, bool GP19 = false>
__device__ __forceinline__ void gray_point(const GrayKernelArgs& A) {
  __shared__ double s_Fm[5][64];  // F_1..F_5 (eta - y_a)
  __shared__ double s_Fp[5][64];  // F_1..F_5 (eta + y_a)
  __shared__ double s_g[4][64];   // TotalNuF(nu_a, s)
  // blocking factor 1 - g of the mixed form (formed without cancellation);
  // the FP64 form takes 1 - g from s_g, the same bits, so that its LDS stays
  // below 10 KB (four workgroups of one wave per SIMD)
  __shared__ double s_bf[MIX && BLOCK ? 4 : 1][64];
  __shared__ double s_y[64];      // node energies / T, formed as the reference does
  // wave-uniform per-point data, read back with broadcast LDS loads so that
  // it occupies neither scalar nor vector registers through the cell loops
  __shared__ double s_in[64];     // the 21 inputs (+ 28 distribution values), 1 / T, log T
  __shared__ Spec s_D[4];
  // the 1D results (iso[4]; beta n_em, e_em, n_ab, e_ab of nue, anue), parked
  // here through the 2D phase so that they hold no registers there
  __shared__ double s_r1d[12];
  __shared__ double s_brem[6];
  __shared__ double s_Fe[2][6], s_fd0[2];
  __shared__ HrChan s_hc[3];
  // the n = 16 instance holds its point's order of the NEPS half-cells here
  // (1 KB, within the 10 KB of four workgroups per SIMD); others read it
  // from the table block
  constexpr bool kPermLds = NQ == 16;
  __shared__ uint4 s_perm[kPermLds ? 2 * NQ * NQ * 2 / 16 : 1];

  const int lane = threadIdx.x;
  const int64_t p = xcd_point(blockIdx.x, gridDim.x);
  if (p >= A.npts) return;
  const int64_t ld = A.ld;
  const double* __restrict__ tb = A.tab;
  const TableLayout L = NQ > 0 ? gray_table_layout(NQ, NQ) : A.L;
  lds_cv<double>* vin = lds_volatile(s_in);

  // ---- point inputs: lane f loads field f ----------------------------------
  if (lane < 21) s_in[lane] = A.pts[lane * ld + p];
  else if (A.distr && lane < 49) s_in[lane] = A.distr[(lane - 21) * ld + p];
  __syncthreads();
  const double T = vin[1];
  const double mu_e = vin[6];

  // ---- distribution parameters: lane s < 4 builds species s ------------------
  if (lane < 4) {
    const int s = lane;
    Spec d = A.distr ? spec_from_fields(vin[21 + 0 + s], vin[21 + 4 + s], vin[21 + 8 + s],
                                        vin[21 + 12 + s], vin[21 + 16 + s], vin[21 + 20 + s],
                                        vin[21 + 24 + s])
                     : spec_from_m1(vin[9 + s], vin[13 + s], vin[17 + s], A.gamma36);
    d.lbf = MIX ? log_pos(d.bf) : 0.;
    s_D[s] = d;
  }
  // 1 / T: energies / T are formed as div_by_recip(nu, T, 1 / T); log(T):
  // the logarithms of the NEPS energies are log(T) + log(nu / T)
  if (lane == 63) s_in[kInvT] = 1. / T;
  if (lane == 62) s_in[kLogT] = log_pos(T);
  if (lane == 61) s_in[kEtaSlot] = mu_e / T;
  if (lane == 60) {
    const double e = mu_e / T;
    s_in[kEmEtaSlot] = e >= 0. ? exp_any(-e) : 0.;
  }
  if (lane == 59) {
    // this point's order of the NEPS half-cells (degeneracy bin)
    const double a = fabs(mu_e / T);
    int bin = 0;
    if (a >= kNpBinHi) bin = kNpBins - 1;
    else if (a >= kNpBinLo) bin = np_bin_of(log_bounded(a) * 0.43429448190325182765);
    s_in[kNpOffSlot] = static_cast<double>(bin * L.Q);
  }

  // ---- per-point constants --------------------------------------------------
  const double eta = mu_e / T;
  {
    // lanes 0/1 evaluate F_0..F_5 at +eta / -eta; lanes 0..3 publish one
    // species' distribution parameters each
    double F[6];
    fdi_bundle<0, 5>(lane == 0 ? eta : -eta, F);
    if (lane < 2) {
#pragma unroll
      for (int k = 0; k < 6; ++k) s_Fe[lane][k] = F[k];
      s_fd0[lane] = fermi_of_arg(lane == 0 ? -eta : eta);
    }
  }
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

  double acc[16];  // [0..3] n_em, [4..7] n_ab, [8..11] e_em, [12..15] e_ab
#pragma unroll
  for (int i = 0; i < 16; ++i) acc[i] = 0.;

  // ======================= 1D: isoenergetic scattering =======================
  lds_cv<double>* vr1d = lds_volatile(s_r1d);
  if (lane < 12) s_r1d[lane] = 0.;
  const double s_nux = 1.5 * T;
  double s_iso[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    double ns = vin[9 + s];
    s_iso[s] = (ns > c::ThresholdN) ? (vin[13 + s] / ns) : s_nux;
  }
  const int n1 = L.n1;
  if (A.flags.use_iso == 1) {
    const double etaNNp = eta_nn_sc(vin[0], T, vin[2]), etaNNn = eta_nn_sc(vin[0], T, vin[3]);
    const bool wm = A.params.use_WM_sc != 0;
    double a4[4] = {0., 0., 0., 0.};
    for (int it = lane; it < 8 * n1; it += kWave) {
      int s = it / (2 * n1), r = it - s * 2 * n1, half = r >= n1, i = r - half * n1;
      const double* q = tb + L.q1 + Q1_REC * i;
      double x = q[Q1_X], w = q[Q1_W], rx = q[Q1_RX];
      double ts = s == 0 ? s_iso[0] : s == 1 ? s_iso[1] : s == 2 ? s_iso[2] : s_iso[3];
      double nu = half ? div_by_recip(ts, x, rx) : ts * x;
      double W = half ? w * (rx * rx) : w;
      double g = dist<MIX>(nu, Dget(s)).g;
      double isc = etaNNp * iso_leg(nu, 1, wm) + etaNNn * iso_leg(nu, 2, wm);
      double nu2 = nu * nu;
      double v = W * g * (4. * c::Pi) * nu2 * nu2 * nu * isc;
      a4[0] += s == 0 ? v : 0.;
      a4[1] += s == 1 ? v : 0.;
      a4[2] += s == 2 ? v : 0.;
      a4[3] += s == 3 ? v : 0.;
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const double v = wave_sum(a4[s]) * s_iso[s];
      if (lane == 0) s_r1d[s] = v;
    }
  }

  // ========================== 1D: beta processes ============================
  double s_beta[2];
  if (eta > -30. && eta < 30.) {
    s_beta[0] = T * vFe[0][5] / vFe[0][4];
    s_beta[1] = T * vFe[1][5] / vFe[1][4];
  } else if (eta > 30.) {
    s_beta[0] = T * eta * 0.8333333333333333;
    s_beta[1] = T * 5.;
  } else {
    s_beta[0] = T * 5.;
    s_beta[1] = T * eta * 0.8333333333333333;
  }
  if (A.flags.use_abs_em == 1) {
    const PointParams& pp = A.params;
    const BetaPoint bp = beta_point(vin[0], T, vin[2], vin[3], vin[4], vin[5], mu_e, vin[7], vin[8],
                                    pp.use_dU != 0, pp.use_dm_eff != 0, pp.use_WM_ab != 0,
                                    pp.use_decay != 0);
    double a8[8] = {0., 0., 0., 0., 0., 0., 0., 0.};
    for (int it = lane; it < 4 * n1; it += kWave) {
      int s = it / (2 * n1), r = it - s * 2 * n1, half = r >= n1, i = r - half * n1;
      const double* q = tb + L.q1 + Q1_REC * i;
      double x = q[Q1_X], w = q[Q1_W], rx = q[Q1_RX];
      double ts = s == 0 ? s_beta[0] : s_beta[1];
      double nu = half ? div_by_recip(ts, x, rx) : ts * x;
      double W = half ? w * (rx * rx) : w;
      double g = dist<MIX>(nu, Dget(s)).g;
      double em, ab;
      beta_em_ab(nu, s, bp, em, ab);
      if constexpr (STIM) ab = ab + em;
      double nu2 = nu * nu;
      double ve = W * nu2 * em, va = W * nu2 * g * ab;
      a8[0] += s == 0 ? ve : 0.;
      a8[1] += s == 1 ? ve : 0.;
      a8[2] += s == 0 ? ve * nu : 0.;
      a8[3] += s == 1 ? ve * nu : 0.;
      a8[4] += s == 0 ? va : 0.;
      a8[5] += s == 1 ? va : 0.;
      a8[6] += s == 0 ? va * nu : 0.;
      a8[7] += s == 1 ? va * nu : 0.;
    }
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const double ne = wave_sum(a8[0 + s]) * s_beta[s], je = wave_sum(a8[2 + s]) * s_beta[s];
      const double na = wave_sum(a8[4 + s]) * s_beta[s], ja = wave_sum(a8[6 + s]) * s_beta[s];
      if (lane == 0) {
        s_r1d[4 + s] = ne;
        s_r1d[6 + s] = je;
        s_r1d[8 + s] = na;
        s_r1d[10 + s] = ja;
      }
    }
  }

  constexpr bool blocking = BLOCK;
  constexpr bool stim = STIM;

  // ========================= 2D: pair + bremsstrahlung =======================
  const bool do_pair = A.flags.use_pair == 1;
  // GP19 without a table adds zero, as the reference's table-free
  // configuration does (kernel_brem_GP19.hpp:483-493)
  const bool do_brem = A.flags.use_brem == 1 && (A.params.brem_impl != 2 || GP19);
  const int N = L.N;
  if (A.flags.use_pair == 1 || A.flags.use_brem == 1) {
    // node tables.  The node energy nu_a = t x_a or t / x_a with t = 2.182 T
    // and y_a = nu_a / T are formed with the reference's roundings
    // (m1_opacities.hpp:783-787, kernel_pair.hpp:142-143) so that thresholds
    // such as eta - y - z > 200 decide as they do there.
    const double tpair = c::PairScale * T;
    auto node_nu = [&](int a) {
      const double* nd = tb + L.nd + ND_REC * a;
      return nd[ND_INV] != 0. ? div_by_recip(tpair, nd[ND_X], nd[ND_RX]) : mul_rn(tpair, nd[ND_X]);
    };
    auto node_y = [&](int a) { return div_by_recip(node_nu(a), T, vin[kInvT]); };
    for (int task = lane; task < 3 * N; task += kWave) {
      if (task < 2 * N) {
        if (do_pair) {
          int a = task < N ? task : task - N;
          double ya = node_y(a);
          double F[6];
          fdi_bundle<1, 5>(task < N ? eta - ya : eta + ya, F);
          if (task < N) {
#pragma unroll
            for (int k = 1; k <= 5; ++k) s_Fm[k - 1][a] = F[k];
          } else {
#pragma unroll
            for (int k = 1; k <= 5; ++k) s_Fp[k - 1][a] = F[k];
          }
        }
      } else {
        int a = task - 2 * N;
        double nu = node_nu(a);
        s_y[a] = div_by_recip(nu, T, vin[kInvT]);
        double lw = MIX ? 0. : log_bounded(nu);
        float lw32 = MIX ? logf(static_cast<float>(nu)) : 0.f;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          GOmg go;
          if constexpr (MIX) {
            go = nuf_mixed(nu, lw32, Dget(s));
          } else {
            go.g = nuf(nu, lw, Dget(s));
            go.omg = 1. - go.g;
          }
          s_g[s][a] = go.g;
          if constexpr (MIX && BLOCK) s_bf[s][a] = go.omg;
        }
      }
    }
    __syncthreads();

    // per-point bremsstrahlung constants, kept in LDS (s_brem: BRT06
    // prefactor; HR98 y, e^-y, p, 1/p, medium factor) and read in the cell
    // loop where they are used
    double brt_pre = 0.;
    double hr_y = 0., hr_emy = 0., hr_p = 1., hr_pinv = 1., hr_medium = 1.;
    if (do_brem && A.params.brem_impl == 1) {
      // QBrem_BRT06 and the prefactor of BremKernelsBRT06
      // (kernel_brem_BRT06.hpp:24-59)
      const double nb = vin[0], yp = vin[2], yn = vin[3];
      double rho = nb * c::Mb;
      double q = 2.0778e+02 * 0.5 * c::MeV * (yn * yn + yp * yp + (28. / 3.) * yn * yp) *
                 (rho * rho) * exp_any(5.5 * log_pos(T));
      double T2 = T * T, T7 = T2 * T2 * T2 * T;
      brt_pre = c::HC6FourPiSq * c::BremC4BRT06 * (q / T7);
    } else if (do_brem && A.params.brem_impl == 0) {
      // BremKernelS / BremKernelG / single channels, the parts that do not
      // depend on x (kernel_brem_HR98.hpp:33-338)
      const HrPoint hp = hr98_point(T, vin[0], A.params.use_NN_medium_corr != 0);
      hr_y = hp.y;
      hr_emy = hp.emy;
      hr_p = hp.p;
      hr_pinv = hp.pinv;
      hr_medium = hp.medium;
      const HrChan hc = hr98_channel(lane < 3 ? lane : 0, T, vin[0], vin[2], vin[3], hp.y, hp.p);
      if (lane < 3) s_hc[lane] = hc;
    }
    if (lane == 0) {
      s_brem[0] = brt_pre;
      s_brem[1] = hr_y;
      s_brem[2] = hr_emy;
      s_brem[3] = hr_p;
      s_brem[4] = hr_pinv;
      s_brem[5] = hr_medium;
    }

    __syncthreads();
    const double a1e = c::PairA1E * c::PairA1E, a2e = c::PairA2E * c::PairA2E;
    const double a1x = c::PairA1X * c::PairA1X, a2x = c::PairA2X * c::PairA2X;
    lds_cv<double>* vbrem = lds_volatile(s_brem);
    for (int cc = lane; cc < L.P; cc += kWave) {
      const double eta = vin[kEtaSlot];
      const double pairT2 = c::PairPhiC * (vin[1] * vin[1]);
      const double* __restrict__ cr = tb + L.cl + CL_REC * cc;
      int ab = int(cr[CL_AB]);
      int a = ab >> 8, b = ab & 255;
      double ya = s_y[a], yb = s_y[b];
      double Kem[4] = {0., 0., 0., 0.}, Kab[4] = {0., 0., 0., 0.};
      if (do_pair) {
        double psi0 = 0., psi1 = 0.;
        double emyz = (eta - ya) - yb;
        if (!(emyz > 200.)) {
          // PairPsi (kernel_pair.hpp:16-120)
          double Fm3[6], Fp3[6];
          fdi_bundle<3, 5>(emyz, Fm3);
          fdi_bundle<3, 5>((eta + ya) + yb, Fp3);
          double F1emy = s_Fm[0][a], F1emz = s_Fm[0][b], F1epy = s_Fp[0][a], F1epz = s_Fp[0][b];
          double F2emy = s_Fm[1][a], F2emz = s_Fm[1][b], F2epy = s_Fp[1][a], F2epz = s_Fp[1][b];
          double F3emy = s_Fm[2][a], F3emz = s_Fm[2][b], F3epy = s_Fp[2][a], F3epz = s_Fp[2][b];
          double F4emy = s_Fm[3][a], F4emz = s_Fm[3][b], F4epy = s_Fp[3][a], F4epz = s_Fp[3][b];
          double F5emy = s_Fm[4][a], F5emz = s_Fm[4][b], F5epy = s_Fp[4][a], F5epz = s_Fp[4][b];
          double y = ya, z = yb;
          double x0 = 20. * F4emy, x1 = 20. * F4epz;
          double x2 = 120. * (F2emy - F2epz);
          double x3 = 40. * (F3epz - F3emy);
          double x4 = 40. * vFe[0][3];
          double x5 = 40. * Fm3[3] - x4;
          double x6 = -20. * vFe[0][4];
          double x7 = 20. * Fm3[4] + x6;
          double x8 = x4 - 40. * Fp3[3];
          double x9 = 20. * Fp3[4] + x6;
          double x10 = 4. * (((((Fm3[5] - F5emy) - F5emz) + F5epy) - Fp3[5]) + F5epz);
          double x11 = 20. * F4epy, x12 = 20. * F4emz;
          double x13 = 40. * (F3epy - F3emz);
          double x14 = 120. * (F2emz - F2epy);
          double yz = y * z;
          double aux = 15. * (yz * yz);
          psi0 = x10 +
                 y * (-x0 + x1 + x7 +
                      y * (x3 + x5 + z * (x2 + z * (120. * (F1epz - F1emy)))) +
                      z * (80. * (F3emy - F3epz) - x2 * z)) +
                 z * (x0 - x1 + x9 + z * (x3 + x8));
          psi1 = x10 + y * (-x11 + x12 + x9 + y * (x13 + x8)) +
                 z * (x11 - x12 + x7 + y * (80. * (F3emz - F3epy) - x14 * y) +
                      z * (x13 + x5 + y * (x14 + y * (120. * (F1epy - F1emz)))));
          double ia = div_normal(1., aux);
          psi0 *= ia;
          psi1 *= ia;
        }
        // PairPhi + PairKernels (kernel_pair.hpp:135-198): em = phi/2.  Past
        // the SafeExp clamp (y + z > 700) phi is subnormal; it is rounded and
        // then halved in the reference's order, because the absorption kernel
        // multiplies it by e^700 afterwards.
        double aux = pairT2 * cr[CL_PAUX];
        double db = cr[CL_DB];
        Kem[0] = 0.5 * ((a1e * psi0 + a2e * psi1) * aux);
        Kem[1] = 0.5 * ((a1e * psi1 + a2e * psi0) * aux);
        Kem[2] = 0.5 * ((a1x * psi0 + a2x * psi1) * aux);
        Kem[3] = 0.5 * ((a1x * psi1 + a2x * psi0) * aux);
#pragma unroll
        for (int s = 0; s < 4; ++s) Kab[s] = db * Kem[s];
      }
      double bem = 0., bab = 0.;
      if (do_brem) {
        const double brt_pre = vbrem[0], hr_y = vbrem[1], hr_emy = vbrem[2];
        const double hr_p = vbrem[3], hr_pinv = vbrem[4], hr_medium = vbrem[5];
        if constexpr (GP19) {
          // BremKernelAbsGP19 at omega = nu_a, omega' = nu_b (m1_opacities.hpp:433-470),
          // the electron fraction Ye taken as yp (the point block has no ye)
          gp19_brem(*A.gp19, node_nu(a), node_nu(b), vin[0], vin[2], vin[1], bem, bab);
        } else if (A.params.brem_impl == 1) {
          bem = brt_pre * cr[CL_K1EM];
          bab = brt_pre * cr[CL_K1AB];
        } else if constexpr (MIX) {
          // the same channel sum in FP32: smooth fits, no cancellation
          float x = static_cast<float>(cr[CL_S]);
          x = x > static_cast<float>(c::BremMin) ? x : static_cast<float>(c::BremMin);
          float y = static_cast<float>(hr_y > c::BremMin ? hr_y : c::BremMin);
          const float pf = static_cast<float>(hr_p), pinvf = static_cast<float>(hr_pinv);
          float snum = static_cast<float>(2. * c::SqrtPi) *
                       expf(1.5f * logf(x + 2.f - static_cast<float>(hr_emy))) *
                       (x * x + 2.f * x * y + 5.f * (y * y) / 3.f + 1.f);
          float bb = static_cast<float>(c::Pi2OneEighth) + x + y;
          float b2 = bb * bb;
          float snd = snum / (static_cast<float>(c::SqrtPi) + b2 * b2);
          float lsdx = static_cast<float>(cr[CL_LSD]);
          float pnd = expf(-pf * logf(snd));
          float px11 = static_cast<float>(cr[CL_X11]), px093 = static_cast<float>(cr[CL_X093]);
          float px12 = static_cast<float>(cr[CL_X12]), px28 = static_cast<float>(cr[CL_X28]);
          float xm = x - 1.2f;
          float fx = 3.f + xm * xm + static_cast<float>(cr[CL_XM4]);
          float cbx = 30.f / (30.f + 0.005f * px28);
          float tot = 0.f;
#pragma unroll
          for (int ch = 0; ch < 3; ++ch) {
            lds_cv<HrChan>& h = lds_volatile(s_hc)[ch];
            float hh = static_cast<float>(h.h), gx = static_cast<float>(h.gamma_half_gb);
            float sd = expf(-pf * (static_cast<float>(h.lsd_pt) + lsdx));
            float sbase = expf(pinvf * logf(pnd + sd));
            float fb = 1.f + 1.f / (fx * static_cast<float>(h.f_pt));
            float gbm = 1.f - px11 * static_cast<float>(h.g0);
            float cb = 1.1f * px11 * hh / (2.3f + hh * px093 + 0.0001f * px12) * cbx;
            float sb = sbase * fb * (1.f + cb * gbm);
            float sabs = static_cast<float>(h.gamma_over_T) / (x * x + gx * gx) * sb;
            tot += static_cast<float>(h.weight) * sabs;
          }
          bab = 3. * c::BremConst * static_cast<double>(tot) * hr_medium;
          bem = bab * cr[CL_EMX];
        } else {
          // BremKernelS at x = (w+w')/T for the three channels
          double x = cr[CL_S];
          x = x > c::BremMin ? x : c::BremMin;
          double y = hr_y > c::BremMin ? hr_y : c::BremMin;
          // the exponents below are logarithms of O(1e-30..1e30) quantities
          // scaled by O(1) powers: the bounded exponential (within an ulp of
          // e^x) without the library's range handling
          double snum = 2. * c::SqrtPi * exp_bounded(1.5 * log_bounded(x + 2. - hr_emy)) *
                        (x * x + 2. * x * y + 5. * (y * y) / 3. + 1.);
          double bb = c::Pi2OneEighth + x + y;
          double b2 = bb * bb;
          double snd = div_normal(snum, c::SqrtPi + b2 * b2);
          double lsnd = log_bounded(snd), lsdx = cr[CL_LSD];
          double pnd = exp_bounded(-hr_p * lsnd);
          double px11 = cr[CL_X11], px093 = cr[CL_X093];
          double px12 = cr[CL_X12], px28 = cr[CL_X28];
          double xm = x - 1.2;
          double fx = 3. + xm * xm + cr[CL_XM4];
          // the factors the three channels share: s_d^-p = e^{-p lsd_pt}
          // e^{-p lsdx} with the first per point, 1 / (fx f_pt) and the
          // second division of C with their cell reciprocals
          // (operands normal, quotients O(1e-30..1e30))
          const double sdx = exp_bounded(-hr_p * lsdx);
          const double rfx = div_normal(1., fx);
          const double r28 = div_normal(30., 30. + 0.005 * px28);
          double tot = 0.;
#pragma unroll
          for (int ch = 0; ch < 3; ++ch) {
            lds_cv<HrChan>& h = lds_volatile(s_hc)[ch];
            double sd = h.sd_pt * sdx;
            double sbase = exp_bounded(hr_pinv * log_bounded(pnd + sd));
            double fb = 1. + rfx * h.rf_pt;
            double gbm = 1. - px11 * h.g0;
            double cb = div_normal(1.1 * px11 * h.h, 2.3 + h.h * px093 + 0.0001 * px12) * r28;
            double sb = sbase * fb * (1. + cb * gbm);
            double gx = h.gamma_half_gb;
            double sabs = div_normal(h.gamma_over_T, x * x + gx * gx) * sb;
            tot += h.weight * sabs;
          }
          bab = 3. * c::BremConst * tot * hr_medium;
          bem = bab * cr[CL_EMX];
        }
      }
      // weights: blocking / stimulated absorption (m1_opacities.hpp:577-758)
      double wab = tb[L.nd + ND_REC * a + ND_WQ] * tb[L.nd + ND_REC * b + ND_WQ];
#pragma unroll
      for (int ord = 0; ord < 2; ++ord) {
        if (ord == 1 && a == b) break;
        int u = ord == 0 ? a : b, v = ord == 0 ? b : a;  // nu at u, nubar at v
        double yu = ord == 0 ? ya : yb;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          int ps = s ^ 1;
          double ke = (ord == 0 ? Kem[s] : Kem[ps]) + bem;
          double ka = (ord == 0 ? Kab[s] : Kab[ps]) + bab;
          double gu = s_g[s][u];
          double bfv = 1.;
          if constexpr (BLOCK) bfv = MIX ? s_bf[ps][v] : 1. - s_g[ps][v];
          double eb = ke * bfv;
          double e_ent, a_ent;
          if constexpr (stim) {
            e_ent = eb;
            a_ent = gu * fma(ka, s_g[ps][v], eb);
          } else {
            e_ent = (1. - gu) * eb;
            a_ent = gu * ka * s_g[ps][v];
          }
          acc[s] = fma(wab, e_ent, acc[s]);
          acc[4 + s] = fma(wab, a_ent, acc[4 + s]);
          acc[8 + s] = fma(wab * yu, e_ent, acc[8 + s]);
          acc[12 + s] = fma(wab * yu, a_ent, acc[12 + s]);
        }
      }
    }
  }

  // ====================== 2D: inelastic scattering on e-+ ===================
  if (A.flags.use_inelastic_scatt == 1) {
    const double bpl = c::NEPSBPlus, bze = c::NEPSBZero, bmi = c::NEPSBMinus;
    const uint16_t* const order = reinterpret_cast<const uint16_t*>(tb + L.npp);
    if constexpr (kPermLds) {
      // 512 indices, 16 bytes per lane (L.npp is even and a bin's block is
      // 1 KB: 16-byte aligned)
      const int npo = __builtin_amdgcn_readfirstlane(static_cast<int>(vin[kNpOffSlot]));
      s_perm[lane] = reinterpret_cast<const uint4*>(order + npo)[lane];
      __syncthreads();
    }
    const uint16_t* const s_order = reinterpret_cast<const uint16_t*>(s_perm);
    for (int h = lane; h < L.Q; h += kWave) {
      // T and eta from LDS in every round, and the products formed from them
      // here: hoisted out of the loop they would hold registers through it
      // (spilled at four waves per SIMD)
      const double T = vin[1], eta = vin[kEtaSlot];
      const double pre = c::NEPSConst * (T * T);
      // t = 4 s_neps with s_neps = 2.182 T (m1_opacities.hpp:1092, 1167)
      const double ht = 0.5 * (4. * (c::PairScale * T));
      // this point's order of visit (its degeneracy bin)
      int hk;
      if constexpr (kPermLds) {
        hk = s_order[h];
      } else {
        const int npo = __builtin_amdgcn_readfirstlane(static_cast<int>(vin[kNpOffSlot]));
        hk = order[npo + h];
      }
      const double* __restrict__ hr = tb + L.np + NP_REC * hk;
      // nu, nubar and w = nu / T with the reference's roundings
      // (m1_opacities.hpp:910-912, 982-984; kernel_nes.hpp:144-145) so that
      // the Taylor-branch tests of NESKernels / NPSKernels decide identically
      double xi = hr[NP_XI];
      double nu, nub;
      if (hr[NP_HALF] == 0.) {
        double hx = mul_rn(ht, xi);
        nu = mul_rn(hx, hr[NP_OMX]);
        nub = mul_rn(hx, hr[NP_OPX]);
      } else {
        nu = div_by_recip(mul_rn(ht, hr[NP_OMX]), xi, hr[NP_RXI]);
        nub = div_by_recip(mul_rn(ht, hr[NP_OPX]), xi, hr[NP_RXI]);
      }
      const double rT = vin[kInvT];
      double w = div_by_recip(nu, T, rT), wp = div_by_recip(nub, T, rT);
      double ad = fabs(w - wp);
      double xm = wp, ym = w;  // w < w' on the whole grid
      if (w > wp) { xm = w; ym = wp; }
      double sgn = hr[NP_SIGN];
      double out[4] = {0., 0., 0., 0.};
#pragma unroll
      for (int lep = 0; lep < 2; ++lep) {
        const double el = lep == 0 ? eta : -eta;
        // couplings (b1, b2): NES (B+,B0),(B0,B+),(B-,B0),(B0,B-); NPS swaps
        double b1[4], b2[4];
        b1[0] = lep == 0 ? bpl : bze;  b2[0] = lep == 0 ? bze : bpl;
        b1[1] = lep == 0 ? bze : bpl;  b2[1] = lep == 0 ? bpl : bze;
        b1[2] = lep == 0 ? bmi : bze;  b2[2] = lep == 0 ? bze : bmi;
        b1[3] = lep == 0 ? bze : bmi;  b2[3] = lep == 0 ? bmi : bze;
        if (ym > el * 1e-3) {
          // MezzacappaIntOut (kernel_nes.hpp:50-78) with the differences of
          // ComputeFDIForInelastic (:29-44)
          double A5[6], B5[6], C3[6];
          // positron targets (el = -eta): at eta > 40 every argument is in
          // the deep tail of the exponential form, so the whole wavefront
          // skips the rationals (the same bits); elsewhere the wave-uniform
          // test costs more than it saves
          if (lep == 1 && el <= 0.) {
            // every argument -eta - v (v = w', w, |w - w'|) is <= 0: its
            // exponential is e^-eta e^-v, e^-v from the cell table at the
            // table's v and corrected to first order for this point's
            // rounding of v (|v - v_tab| is a few ulps of v): three
            // operations instead of an exponential per bundle
            const double em = vin[kEmEtaSlot];
            const double wt = hr[NP_W], wpt = hr[NP_WP];
            const double ew = hr[NP_EW], ewp = hr[NP_EWP], ead = hr[NP_EAD];
            fdi_bundle_ex<1, 5, true>(el - wp, em * fma(wpt - wp, ewp, ewp), A5);
            fdi_bundle_ex<1, 5, true>(el - w, em * fma(wt - w, ew, ew), B5);
            fdi_bundle_ex<3, 5, true>(el - ad, em * fma(fabs(wt - wpt) - ad, ead, ead), C3);
          } else if (lep == 1) {
            fdi_bundle<1, 5, true>(el - wp, A5);
            fdi_bundle<1, 5, true>(el - w, B5);
            fdi_bundle<3, 5, true>(el - ad, C3);
          } else {
            // electron targets: from eta = 40 on whole rounds of arguments
            // eta - v commonly lie beyond 40 on either side (the deep tail
            // of the exponential form, reflected where positive)
            const bool tail = el > 40.;
            fdi_bundle<1, 5, true>(el - wp, A5, tail);
            fdi_bundle<1, 5, true>(el - w, B5, tail);
            fdi_bundle<3, 5, true>(el - ad, C3, tail);
          }
          double fw0 = A5[1] - B5[1], fw1 = A5[2] - B5[2], fw2 = A5[3] - B5[3];
          double fw3 = A5[4] - B5[4], fw4 = A5[5] - B5[5];
          double fa0 = vFe[lep][3] - C3[3], fa1 = vFe[lep][4] - C3[4];
          double fa2 = vFe[lep][5] - C3[5];
          double ws = w + wp;
          double inv = hr[NP_INVDEN];
          double q4 = fw3 + 2. * (ws * fw2 + 3. * w * wp * fw1);
          // the out-kernel is linear in the couplings: b1 U + b2 V
          double t1 = (sgn * fa2 - fw4) * 0.2;
          double U = (t1 - ws * q4 + sgn * (xm * fa1 + 2. * (xm * xm) * fa0)) * inv - 6. * fw0;
          double V = (t1 + sgn * (2. * (ym * ym) * fa0 - ym * fa1)) * inv;
          U *= pre;
          V *= pre;
#pragma unroll
          for (int s = 0; s < 4; ++s) out[s] = fma(b1[s], U, fma(b2[s], V, out[s]));
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
            double S = 2. * (B1 + B2) * f0 + (B1 * (yq + 4. * xq) + 3. * B2 * yq) * f1 +
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
            double S = (b1[s] + b2[s]) * (4. * f2 + f0 * pol) + (b1[s] - b2[s]) * (2. * xq - yq) * f1;
            out[s] = fma(pre, yq * (w - wp) * S * ix2, out[s]);
          }
        }
      }
      // distribution values at nu and nubar, formed after the kernels so that
      // they are not live through the Fermi-integral evaluations
      double gn[4], gb[4], on[4], ob[4];  // f and 1 - f at nu and nubar
      const double lT = vin[kLogT];
      if constexpr (MIX) {
        float l1 = static_cast<float>(lT + hr[NP_LW]), l2 = static_cast<float>(lT + hr[NP_LWP]);
        auto species = [&](int s) {
          Spec d = Dget(s);
          GOmg a = nuf_mixed(nu, l1, d), b = nuf_mixed(nub, l2, d);
          gn[s] = a.g;
          on[s] = a.omg;
          gb[s] = b.g;
          ob[s] = b.omg;
        };
#pragma unroll
        for (int s = 0; s < 3; ++s) species(s);
        if (!same23) species(3);
      } else {
        double l1 = lT + hr[NP_LW], l2 = lT + hr[NP_LWP];
        auto species = [&](int s) {
          Spec d = Dget(s);
          gn[s] = nuf(nu, l1, d);
          gb[s] = nuf(nub, l2, d);
          on[s] = 1. - gn[s];
          ob[s] = 1. - gb[s];
        };
#pragma unroll
        for (int s = 0; s < 3; ++s) species(s);
        if (!same23) species(3);
      }
      if (same23) {
        gn[3] = gn[2];
        gb[3] = gb[2];
        on[3] = on[2];
        ob[3] = ob[2];
      }
      // in/out kernels for (nu, nubar) and the swapped pair: the swapped
      // kernel is the negated out-integral with NEPSExpFunc arguments
      // exchanged, i.e. em2 = ab1, ab2 = em1 (kernel_nes.hpp:229-234)
      double ef = hr[NP_EF], efx = hr[NP_EFX];
      double wq = hr[NP_WQ];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        double em1 = -out[s] * efx, ab1 = out[s] * ef;
        double bfn = blocking ? on[s] : 1., bfb = blocking ? ob[s] : 1.;
        double te1 = em1 * gb[s], ta1 = ab1 * bfb;
        double te2 = ab1 * gn[s], ta2 = em1 * bfn;
        double Ae, Aa, Be, Ba;
        if constexpr (stim) {
          Ae = te1;
          Aa = gn[s] * (te1 + ta1);
          Be = te2;
          Ba = gb[s] * (te2 + ta2);
        } else {
          Ae = on[s] * te1;
          Aa = gn[s] * ta1;
          Be = ob[s] * te2;
          Ba = gb[s] * ta2;
        }
        acc[s] = fma(wq, Ae + Be, acc[s]);
        acc[4 + s] = fma(wq, Aa + Ba, acc[4 + s]);
        acc[8 + s] = fma(wq, fma(w, Ae, wp * Be), acc[8 + s]);
        acc[12 + s] = fma(wq, fma(w, Aa, wp * Ba), acc[12 + s]);
      }
    }
  }

  // ============================ reduce + assemble ============================
#pragma unroll
  for (int i = 0; i < 16; ++i) acc[i] = wave_sum(acc[i]);
  if (lane == 0) {
    const double T = vin[1];
    const double tp = c::PairScale * T;
    const double T2 = T * T;
    const double sN = tp * tp * (T2 * T2);  // t^2 T^4
    const double sE = sN * T;               // energy weights are y = nu / T
    double n2d[8], e2d[8];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      n2d[s] = acc[s] * sN;
      n2d[4 + s] = acc[4 + s] * sN;
      e2d[s] = acc[8 + s] * sE;
      e2d[4 + s] = acc[12 + s] * sE;
    }
    const double F = c::FourPiHc3, F2 = c::FourPiHc3Sqr, cl = c::Clight;
    double o[20];
#pragma unroll
    for (int i = 0; i < 20; ++i) o[i] = 0.;
    // m1_opacities.hpp:1187-1292
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const double bne = vr1d[4 + s], bje = vr1d[6 + s], bna = vr1d[8 + s], bja = vr1d[10 + s];
      o[0 + s] = F * (F * n2d[s] + bne);
      o[8 + s] = F * (F * e2d[s] + bje);
      const double ns = vin[9 + s], Js = vin[13 + s];
      if (ns > c::ThresholdN) o[4 + s] = F / (cl * ns) * (F * n2d[4 + s] + bna);
      if (Js > c::ThresholdJ) {
        double ka = F / (cl * Js) * (F * e2d[4 + s] + bja);
        o[12 + s] = (s == 0 && ns == 0.) ? 0. : ka;
        o[16 + s] = F / (cl * Js) * vr1d[s];
      }
    }
#pragma unroll
    for (int s = 2; s < 4; ++s) {
      o[0 + s] = F2 * n2d[s];
      o[8 + s] = F2 * e2d[s];
      const double ns = vin[9 + s], Js = vin[13 + s];
      if (ns > c::ThresholdN) {
        double k0 = F2 / (cl * ns) * n2d[4 + s];
        o[4 + s] = (s == 3 && ns == 0.) ? 0. : k0;
      }
      if (Js > c::ThresholdJ) {
        o[12 + s] = F2 / (cl * Js) * e2d[4 + s];
        o[16 + s] = F / (cl * Js) * vr1d[s];
      }
    }
#pragma unroll
    for (int i = 0; i < 20; ++i) s_in[i] = o[i];
  }
  __syncthreads();
  if (lane < 20) A.out[lane * ld + p] = vin[lane];
}

template <int WAVES_PER_EU, bool STIM, bool BLOCK, int NQ, bool GP19 = false>
__global__ __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(WAVES_PER_EU, WAVES_PER_EU))) void k_gray_m1_f64(
    GrayKernelArgs A) {
  gray_point<false, STIM, BLOCK, NQ, GP19>(A);
}

template <int WAVES_PER_EU, bool STIM, bool BLOCK, int NQ, bool GP19 = false>
__global__ __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(WAVES_PER_EU, WAVES_PER_EU))) void k_gray_m1_mixed(
    GrayKernelArgs A) {
  gray_point<true, STIM, BLOCK, NQ, GP19>(A);
}

}  // namespace bns

namespace bns {
namespace {
template <bool MIX, bool STIM, bool BLOCK>
void launch_instance(const GrayKernelArgs& a, dim3 grid, hipStream_t stream) {
  // 4 waves per SIMD (128 VGPRs; the spills are a few values reloaded once
  // per cell round); the mixed form with blocking keeps its 2 KB blocking
  // table, and its LDS allows 3.7 workgroups per SIMD, so it stays at 3
  constexpr int W = (MIX && BLOCK) ? 3 : 4;
  // the 128^3-grid workload's order, n = 16, has its own instance (0.6%
  // faster: its table offsets and loop bounds are constants)
  const bool fixed = a.L.n1 == 16 && a.L.n2 == 16;
  // GP19 bremsstrahlung over a table: its own instances, 3 waves per SIMD
  // (163 VGPRs)
  if (a.gp19 && a.params.brem_impl == 2 && a.flags.use_brem == 1) {
    if constexpr (MIX)
      hipLaunchKernelGGL((k_gray_m1_mixed<3, STIM, BLOCK, 0, true>), grid, dim3(64), 0, stream, a);
    else
      hipLaunchKernelGGL((k_gray_m1_f64<3, STIM, BLOCK, 0, true>), grid, dim3(64), 0, stream, a);
    return;
  }
  if constexpr (MIX) {
    if (fixed)
      hipLaunchKernelGGL((k_gray_m1_mixed<W, STIM, BLOCK, 16>), grid, dim3(64), 0, stream, a);
    else
      hipLaunchKernelGGL((k_gray_m1_mixed<W, STIM, BLOCK, 0>), grid, dim3(64), 0, stream, a);
  } else {
    if (fixed)
      hipLaunchKernelGGL((k_gray_m1_f64<W, STIM, BLOCK, 16>), grid, dim3(64), 0, stream, a);
    else
      hipLaunchKernelGGL((k_gray_m1_f64<W, STIM, BLOCK, 0>), grid, dim3(64), 0, stream, a);
  }
}
// the instance for the launch's stimulated-absorption and blocking switches
template <bool MIX>
hipError_t launch_gray(const GrayKernelArgs& a, hipStream_t stream) {
  if (a.npts <= 0) return hipSuccess;
  const dim3 grid(static_cast<unsigned>(a.npts));
  const bool blocking = a.params.neglect_blocking == 0;
  if (a.stim == 1) {
    if (blocking) launch_instance<MIX, true, true>(a, grid, stream);
    else launch_instance<MIX, true, false>(a, grid, stream);
  } else {
    if (blocking) launch_instance<MIX, false, true>(a, grid, stream);
    else launch_instance<MIX, false, false>(a, grid, stream);
  }
  return hipGetLastError();
}
}  // namespace

hipError_t launch_gray_m1_f64(const GrayKernelArgs& a, hipStream_t stream) {
  return launch_gray<false>(a, stream);
}
hipError_t launch_gray_m1_mixed(const GrayKernelArgs& a, hipStream_t stream) {
  return launch_gray<true>(a, stream);
}
}  // namespace bns
