// gp19.h -- GP19 nucleon-nucleon bremsstrahlung (Guo & Martinez-Pinedo 2019)
// on the device: BremKernelAbsGP19 of kernel_brem_GP19.hpp:354-480 over a
// device-resident 4D table of the response function (log nb, Ye, T, log w)
// that the caller supplies through bnsamd_set_gp19_table; the reference
// ships its table as a generated header that is absent from its tree.
#ifndef BNS_NURATES_AMD_GP19_H_
#define BNS_NURATES_AMD_GP19_H_

#include <hip/hip_runtime.h>

#include "gray_math.h"

namespace bns {

// Axes and data of one table (device pointers); the kernels receive a
// pointer to this struct in device memory, and nullptr when there is no table.
// nb [fm^-3] and w [MeV] with their logarithms as the table states them,
// Ye, T [MeV]; data[((i nye + j) nT + k) nw + l].
struct Gp19Table {
  int nnb, nye, nT, nw;
  const double *nb, *lognb, *ye, *T, *w, *logw, *data;
};

// The interval [ax[i], ax[i+1]] that the reference's linear scan returns
// (the first one holding v, kernel_brem_GP19.hpp:34-56), found by bisection:
// the smallest k >= 1 with ax[k] >= v gives i = k - 1, also when v is a node.
// False outside the axis and for NaN.
__device__ inline bool gp19_bracket(double v, const double* __restrict__ ax, int n, int& i,
                                    double& t) {
  if (!(v >= ax[0] && v <= ax[n - 1])) return false;
  int lo = 1, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (ax[mid] >= v) hi = mid;
    else lo = mid + 1;
  }
  i = lo - 1;
  t = (v - ax[i]) / (ax[i + 1] - ax[i]);
  return true;
}

// gp19_interpolator (kernel_brem_GP19.hpp:72-138): multilinear in
// (log nb, Ye, T, log w) over the 16 vertices, summed in the reference's
// order; NaN outside the table
__device__ inline double gp19_interp(const Gp19Table& tb, double nb, double ye, double T,
                                     double w) {
  int i, j, k, l;
  double tx, ty, tz, tw;
  if (!gp19_bracket(log(nb), tb.lognb, tb.nnb, i, tx) || !gp19_bracket(ye, tb.ye, tb.nye, j, ty) ||
      !gp19_bracket(T, tb.T, tb.nT, k, tz) || !gp19_bracket(log(w), tb.logw, tb.nw, l, tw))
    return __builtin_nan("");
  const int64_t sl = 1, sk = tb.nw, sj = int64_t(tb.nT) * sk, si = int64_t(tb.nye) * sj;
  const double* __restrict__ d = tb.data + i * si + j * sj + k * sk + l;
  const double a[2] = {1. - tx, tx}, b[2] = {1. - ty, ty}, c[2] = {1. - tz, tz},
               e[2] = {1. - tw, tw};
  double s = 0.;
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const int vi = v >> 3, vj = (v >> 2) & 1, vk = (v >> 1) & 1, vl = v & 1;
    const double term = d[vi * si + vj * sj + vk * sk + vl * sl] * a[vi] * b[vj] * c[vk] * e[vl];
    s = v == 0 ? term : s + term;
  }
  return s;
}

// gp19_BremKernel_in_ranges (:158-170), nb in fm^-3: kernel [nm^3 s^-1]
__device__ inline double gp19_in_ranges(const Gp19Table& tb, double nb, double ye, double T,
                                        double w) {
  return c::BremConst * (nb * 1e18) * gp19_interp(tb, nb, ye, T, w);
}

// below the table's density: quadratic in nb from its first node (:176-189)
__device__ inline double gp19_low_density(const Gp19Table& tb, double nb, double ye, double T,
                                          double w) {
  const double nb_min = tb.nb[0], r = nb / nb_min;
  return gp19_in_ranges(tb, nb_min, ye, T, w) * (r * r);
}

// below T = 2 MeV: the power law through the kernels at 2 and 4 MeV
// (:192-212; with the density extrapolation and its 1e-30 floors, :216-258)
__device__ inline double gp19_low_temperature(const Gp19Table& tb, double nb, double ye, double T,
                                              double w, bool low_density) {
  constexpr double ilog2 = 1.4426950408889634;  // 1 / log(4 / 2)
  // the floor leaves NaN as it is, as the reference's `if (k < eps) k = eps`
  auto floor30 = [](double k) { return k < 1e-30 ? 1e-30 : k; };
  double k2, k4;
  if (low_density) {
    k2 = floor30(gp19_low_density(tb, nb, ye, 2., w));
    k4 = floor30(gp19_low_density(tb, nb, ye, 4., w));
  } else {
    k2 = gp19_in_ranges(tb, nb, ye, 2., w);
    k4 = gp19_in_ranges(tb, nb, ye, 4., w);
  }
  const double m = log(k4 / k2) * ilog2;
  const double kx = k2 * pow(T * 0.5, m);
  return low_density ? floor30(kx) : kx;
}

// above T = 50 MeV: the Lorentzian matched to the kernel and its slope at
// 48-50 MeV (:262-312)
__device__ inline double gp19_high_temperature(const Gp19Table& tb, double nb, double ye, double T,
                                               double w) {
  constexpr double sqrt50 = 7.0710678118654755;
  const double nb_nm = nb * 1e18, nb2 = nb_nm * nb_nm, w2 = w * w;
  const double r48 = gp19_in_ranges(tb, nb, ye, 48., w), r50 = gp19_in_ranges(tb, nb, ye, 50., w);
  const double K = 4. * c::BremConst * nb2;
  const double A = 1. / r50, B = 1. / (K * sqrt50);
  const double m = (r50 - r48) / (50. - 48.);
  const double beta = 8. * w2 * K * (B * B) / (2. * m * (A * A) * sqrt50 + A * K * B);
  const double gamma = (A * beta - 4. * B * w2) / (B * nb2 * 50.);
  const double inv = 1. / ((w / T) * (w / T) + (nb2 / (4. * T)) * gamma);
  return c::BremConst * nb_nm * ((beta * nb_nm) / (T * sqrt(T)) * inv);
}

// BremKernelAbsGP19 (:354-480): absorption and emission kernels for the pair
// energy omega + omega' [MeV] at nb [nm^-3], Ye, T; Ye above and nb above the
// table are clipped to its edges, w to its range; the density and the
// temperature below or above the table are extrapolated; the l = 0 term is 3
// times the response, taken in absolute value; emission by detailed balance
// with the unclipped pair energy.
__device__ inline void gp19_brem(const Gp19Table& tb, double omega, double omega_p, double nb_nm,
                                 double ye, double T, double& s_em, double& s_abs) {
  const double w_original = omega + omega_p;
  double w = w_original, nb = nb_nm * 1e-18;
  const double nb_min = tb.nb[0], nb_max = tb.nb[tb.nnb - 1], ye_max = tb.ye[tb.nye - 1];
  const double T_min = tb.T[0], T_max = tb.T[tb.nT - 1];
  const double w_min = tb.w[0], w_max = tb.w[tb.nw - 1];
  if (ye > ye_max) ye = ye_max;
  if (nb > nb_max) nb = nb_max;
  if (w > w_max) w = w_max;
  else if (w < w_min) w = w_min;
  double s;
  if (nb >= nb_min && T >= T_min && T <= T_max) {
    s = gp19_in_ranges(tb, nb, ye, T, w);
  } else if (nb < nb_min) {
    if (T < T_min) {
      s = gp19_low_temperature(tb, nb, ye, T, w, true);
    } else if (T > T_max) {
      const double r = nb / nb_min;
      s = gp19_high_temperature(tb, nb_min, ye, T, w) * (r * r);
    } else {
      s = gp19_low_density(tb, nb, ye, T, w);
    }
  } else if (T < T_min) {
    s = gp19_low_temperature(tb, nb, ye, T, w, false);
  } else {
    s = gp19_high_temperature(tb, nb, ye, T, w);
  }
  s_abs = fabs(3. * s);
  s_em = s_abs * safe_exp(-w_original / T);
}

}  // namespace bns

#endif  // BNS_NURATES_AMD_GP19_H_
