"""Seeded synthetic thermodynamic points for the gray M1 opacity workloads.

The reference ships no equation of state (``include/eos.hpp`` is empty), so
the workloads define their own, in closed form:

* baryon density ``nb = rho / m_u * 1e-21`` nm^-3 (as ``tests/tests.hpp:304``);
* electron chemical potential from an ideal, massless, relativistic e-/e+
  gas with net density ``ne = Ye nb``: ``ne = (mu^3 + pi^2 T^2 mu) /
  (3 pi^2 (hbar c)^3)``, solved exactly (hyperbolic form of Cardano);
* neutron / proton chemical potentials from non-relativistic ideal Fermi
  gases, ``n_N = 2 (m_N T / (2 pi (hbar c)^2))^{3/2} F_{1/2}(eta_N)``
  (normalised F), inverted by Newton iteration on the Bednarczyk &
  Bednarczyk (1978) closed-form approximation of F_{1/2}; ``mu_N = m_N + T eta_N``
  (rest mass included, ``bns_nurates.hpp:287``);
* neutrino moments at equilibrium (``NuEquilibriumParams`` +
  ``ComputeM1DensitiesEq``, ``distribution.hpp:253-290,447-473``) times
  independent log-uniform factors in [0.1, 10] for n and for J; chi uniform in
  [1/3, 1]; the heavy antineutrino copies the heavy neutrino
  (``tests/tests.hpp:551-553``).

Everything is float64 numpy, vectorised, and deterministic for a given seed
(numpy's PCG64).  The point block layout is the one of the C ABI
(``include/bns_nurates_amd.h``): 21 fields of ``npts`` doubles each.
"""
from __future__ import annotations

import numpy as np

# field indices of the point block (include/bns_nurates_amd.h)
P_NB, P_T, P_YP, P_YN, P_MUP, P_MUN, P_MUE, P_DU, P_DMEFF = range(9)
P_N0, P_J0, P_CHI0 = 9, 13, 17
P_NFIELDS = 21
D_NFIELDS = 28
O_NFIELDS = 20
OUT_NAMES = ([f"eta_0[{s}]" for s in range(4)] + [f"kappa_0_a[{s}]" for s in range(4)]
             + [f"eta[{s}]" for s in range(4)] + [f"kappa_a[{s}]" for s in range(4)]
             + [f"kappa_s[{s}]" for s in range(4)])

# constants (constants.hpp:32-74)
M_U = 1.66053906892e-24        # g
M_N = 939.56542194             # MeV
M_P = 938.27208943             # MeV
Q_NP = 1.29333251              # MeV
H_PLANCK = 4.135667696e-21     # MeV s
C_LIGHT = 2.99792458e+17       # nm / s
HBARC = 0.5 * H_PLANCK / np.pi * C_LIGHT          # MeV nm
FOUR_PI_HC3 = (4.0 * np.pi) / (H_PLANCK * C_LIGHT) ** 3

# OpacityFlags / OpacityParams (bns_nurates.hpp:338-438) as int arrays
FLAGS_ALL = np.array([1, 1, 1, 1, 1], dtype=np.int32)
BREM_HR98, BREM_BRT06, BREM_GP19 = 0, 1, 2


def params_default_all(brem: int = BREM_BRT06) -> np.ndarray:
    """``opacity_params_default_all`` (bns_nurates.hpp:353-361) with the
    bremsstrahlung implementation switched as the workload names it."""
    return np.array([1, 1, 1, 1, 1, brem, 1, 1], dtype=np.int32)


def params_default_none(brem: int = BREM_HR98) -> np.ndarray:
    """``opacity_params_default_none`` (bns_nurates.hpp:362-370)."""
    return np.array([0, 0, 0, 0, 0, brem, 0, 0], dtype=np.int32)


def gauss_legendre(n: int):
    """Gauss-Legendre nodes/weights on [0, 1], ascending (integration.hpp:28-87
    generates the same rule by Newton iteration; numpy's is used here only
    for workload set-up -- the C library has its own generator)."""
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (x + 1.0), 0.5 * w


# ---------------------------------------------------------------------------
def _mu_e_massless(ne: np.ndarray, T: np.ndarray) -> np.ndarray:
    """Root of mu^3 + pi^2 T^2 mu - 3 pi^2 (hbar c)^3 ne = 0 (one real root)."""
    p = np.pi ** 2 * T ** 2
    q = 3.0 * np.pi ** 2 * HBARC ** 3 * ne
    a = 2.0 * np.sqrt(p / 3.0)
    arg = (3.0 * q / (2.0 * p)) * np.sqrt(3.0 / p)
    return a * np.sinh(np.arcsinh(arg) / 3.0)


def _f12_norm(eta: np.ndarray):
    """Normalised F_{1/2} (2/sqrt(pi) int sqrt(x)/(e^{x-eta}+1)) and its
    eta-derivative, closed-form approximation of Bednarczyk & Bednarczyk,
    Phys. Lett. A 64 (1978) 409."""
    g = np.exp(-0.17 * (eta + 1.0) ** 2)
    nu = eta ** 4 + 50.0 + 33.6 * eta * (1.0 - 0.68 * g)
    dnu = 4.0 * eta ** 3 + 33.6 * (1.0 - 0.68 * g) + 33.6 * eta * 0.68 * 0.34 * (eta + 1.0) * g
    c = 0.75 * np.sqrt(np.pi)
    e = np.exp(-np.minimum(eta, 700.0))
    den = e + c * nu ** (-0.375)
    dden = -e - 0.375 * c * nu ** (-1.375) * dnu
    return 1.0 / den, -dden / den ** 2


def _eta_ideal_nucleon(n: np.ndarray, T: np.ndarray, m: float) -> np.ndarray:
    """Degeneracy of an ideal non-relativistic nucleon gas of density n:
    Newton iteration on log F_{1/2}(eta) = log(n / n_Q)."""
    nq = 2.0 * (m * T / (2.0 * np.pi * HBARC ** 2)) ** 1.5
    lu = np.log(n / nq)
    u = np.exp(np.minimum(lu, 700.0))
    eta = np.where(lu < 0.0, lu, np.maximum(lu, (0.75 * np.sqrt(np.pi) * u) ** (2.0 / 3.0)))
    for _ in range(60):
        f, df = _f12_norm(eta)
        step = np.clip((np.log(f) - lu) * f / df, -20.0, 20.0)
        eta = eta - step
        if np.max(np.abs(step) / np.maximum(1.0, np.abs(eta))) < 1e-14:
            break
    return eta


# Cohen, Rodriguez Villegas & Zagier (Exp. Math. 9 (2000) 3), algorithm 1:
# weights that sum an alternating series sum_{m>=0} (-1)^m a_m of a moment
# sequence a_m to ~5.8^-M relative accuracy with M terms.
def _crvz_weights(M: int = 24) -> np.ndarray:
    d = (3.0 + np.sqrt(8.0)) ** M
    d = 0.5 * (d + 1.0 / d)
    b, c = -1.0, -d
    w = np.zeros(M)
    for k in range(M):
        c = b - c
        w[k] = c / d          # coefficient of (-1)^k a_k
        b = (k + M) * (k - M) * b / ((k + 0.5) * (k + 1))
    return w


_CRVZ = _crvz_weights()


def _fdi_int(k: int, eta: np.ndarray) -> np.ndarray:
    """Complete Fermi integral F_k(eta) = int_0^inf x^k / (e^{x-eta}+1) dx,
    k >= 1, for all eta: for eta <= 0 the alternating series
    k! sum_{m>=1} (-1)^{m+1} e^{m eta} / m^{k+1} (accelerated); for eta > 0
    the reflection F_k(eta) = (-1)^k F_k(-eta) + P_k(eta) with the exact
    polynomial P_k of the polylogarithm inversion formula."""
    from math import factorial
    from scipy.special import zeta
    a = -np.abs(eta)
    x = np.exp(a)
    M = _CRVZ.size
    s = np.zeros_like(eta)
    xm = np.ones_like(eta)
    for m in range(M):
        xm = xm * x
        s += _CRVZ[m] * xm / float(m + 1) ** (k + 1)
    f_neg = factorial(k) * s
    p = eta ** (k + 1) / (k + 1)
    for j in range(1, (k + 1) // 2 + 1):
        p = p + 2.0 * factorial(k) / factorial(k + 1 - 2 * j) * (1.0 - 2.0 ** (1 - 2 * j)) * \
            zeta(2 * j) * eta ** (k + 1 - 2 * j)
    return np.where(eta > 0.0, (-1) ** k * f_neg + p, f_neg)


def make_points(npts: int, seed: int = 0, rho_range=(1e6, 1e15), T_range=(0.5, 50.0),
                ye_range=(0.01, 0.6), factor_range=(0.1, 10.0)) -> np.ndarray:
    """Point block [21, npts] drawn from the workload distributions."""
    rng = np.random.Generator(np.random.PCG64(seed))
    rho = np.exp(rng.uniform(np.log(rho_range[0]), np.log(rho_range[1]), npts))
    T = np.exp(rng.uniform(np.log(T_range[0]), np.log(T_range[1]), npts))
    ye = rng.uniform(ye_range[0], ye_range[1], npts)
    fac = np.exp(rng.uniform(np.log(factor_range[0]), np.log(factor_range[1]), (2, 3, npts)))
    chi = rng.uniform(1.0 / 3.0, 1.0, (3, npts))

    nb = rho / M_U * 1e-21
    yp, yn = ye, 1.0 - ye
    mu_e = _mu_e_massless(ye * nb, T)
    mu_n = M_N + T * _eta_ideal_nucleon(yn * nb, T, M_N)
    mu_p = M_P + T * _eta_ideal_nucleon(yp * nb, T, M_P)

    eta_nue = (mu_e - mu_n + mu_p) / T               # distribution.hpp:284
    etas = [eta_nue, -eta_nue, np.zeros_like(eta_nue)]
    npre = FOUR_PI_HC3 * T ** 3
    jpre = npre * T

    pts = np.empty((P_NFIELDS, npts), dtype=np.float64)
    pts[P_NB], pts[P_T], pts[P_YP], pts[P_YN] = nb, T, yp, yn
    pts[P_MUP], pts[P_MUN], pts[P_MUE] = mu_p, mu_n, mu_e
    pts[P_DU] = 0.0
    pts[P_DMEFF] = Q_NP
    for s in range(3):
        pts[P_N0 + s] = npre * _fdi_int(2, etas[s]) * fac[0, s]
        pts[P_J0 + s] = jpre * _fdi_int(3, etas[s]) * fac[1, s]
        pts[P_CHI0 + s] = chi[s]
    pts[P_N0 + 3] = pts[P_N0 + 2]
    pts[P_J0 + 3] = pts[P_J0 + 2]
    pts[P_CHI0 + 3] = pts[P_CHI0 + 2]
    return pts


def make_tail_points(npts: int, seed: int = 0) -> np.ndarray:
    """Cold, degenerate points of the same distributions (T < 1 MeV and
    eta_e = mu_e / T > 50): the tail where reduced precision is most exposed.
    Draws T in [0.5, 1] MeV and rho in [1e12, 1e15] g/cm^3 and keeps the
    first npts points with eta_e > 50."""
    out, got, k = [], 0, 0
    while got < npts:
        p = make_points(4 * npts, seed=seed + 7919 * k, rho_range=(1e12, 1e15), T_range=(0.5, 1.0))
        keep = p[:, p[P_MUE] / p[P_T] > 50.0]
        out.append(keep)
        got += keep.shape[1]
        k += 1
    return np.ascontiguousarray(np.concatenate(out, axis=1)[:, :npts])


# The named workloads (sizes, quadrature order, reactions, bremsstrahlung)
WORKLOADS = {
    "pts4096_n10_brt06": dict(npts=4096, nquad=10, brem=BREM_BRT06, seed=0),
    "grid128_n16_brt06": dict(npts=128 ** 3, nquad=16, brem=BREM_BRT06, seed=1),
    "grid128_n16_hr98": dict(npts=128 ** 3, nquad=16, brem=BREM_HR98, seed=1),
}
