"""Seeded synthetic thermodynamic points for the gray M1 opacity workloads.

The reference ships no equation of state (``include/eos.hpp`` is empty), so
the workloads define their own, in closed form:

* baryon density ``nb = rho / m_u * 1e-21`` nm^-3 (as ``tests/tests.hpp:304``);
* electron chemical potential from an ideal, massless, relativistic e-/e+
  gas with net density ``ne = Ye nb``: ``ne = (mu^3 + pi^2 T^2 mu) /
  (3 pi^2 (hbar c)^3)``, solved exactly (hyperbolic form of Cardano);
* neutron / proton chemical potentials from non-relativistic ideal Fermi
  gases, ``n_N = 2 (m_N T / (2 pi (hbar c)^2))^{3/2} F_{1/2}(eta_N)``
  (normalised F), inverted by bisection on the Bednarczyk & Bednarczyk
  (1978) closed-form approximation of F_{1/2}; ``mu_N = m_N + T eta_N``
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


def _f12_norm(eta: np.ndarray) -> np.ndarray:
    """Normalised F_{1/2} (2/sqrt(pi) int sqrt(x)/(e^{x-eta}+1)), closed-form
    approximation of Bednarczyk & Bednarczyk, Phys. Lett. A 64 (1978) 409."""
    nu = eta ** 4 + 50.0 + 33.6 * eta * (1.0 - 0.68 * np.exp(-0.17 * (eta + 1.0) ** 2))
    return 1.0 / (np.exp(-eta) + 0.75 * np.sqrt(np.pi) * nu ** (-0.375))


def _eta_ideal_nucleon(n: np.ndarray, T: np.ndarray, m: float) -> np.ndarray:
    nq = 2.0 * (m * T / (2.0 * np.pi * HBARC ** 2)) ** 1.5
    u = n / nq
    lo = np.full_like(u, -200.0)
    hi = np.maximum(np.log(np.maximum(u, 1e-300)) + 5.0, (1.5 * u) ** (2.0 / 3.0) + 5.0)
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        f = _f12_norm(mid) - u
        lo = np.where(f < 0.0, mid, lo)
        hi = np.where(f < 0.0, hi, mid)
    return 0.5 * (lo + hi)


def _fdi_quad(k: int, eta: np.ndarray, nodes: int = 64) -> np.ndarray:
    """Complete Fermi integral F_k(eta) = int_0^inf x^k / (e^{x-eta}+1) dx by
    two-segment Gauss-Legendre ([0, max(eta,0)] and 60 units beyond)."""
    x, w = np.polynomial.legendre.leggauss(nodes)
    x = 0.5 * (x + 1.0)
    w = 0.5 * w
    a = np.maximum(eta, 0.0)[:, None]
    out = np.zeros(eta.shape, dtype=np.float64)
    for lo, hi in ((0.0 * a, a), (a, a + 60.0)):
        L = hi - lo
        xs = lo + L * x[None, :]
        z = np.clip(xs - eta[:, None], -700.0, 700.0)
        f = xs ** k / (np.exp(z) + 1.0)
        out += (L * (f * w[None, :])).sum(axis=1)
    return out


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
        pts[P_N0 + s] = npre * _fdi_quad(2, etas[s]) * fac[0, s]
        pts[P_J0 + s] = jpre * _fdi_quad(3, etas[s]) * fac[1, s]
        pts[P_CHI0 + s] = chi[s]
    pts[P_N0 + 3] = pts[P_N0 + 2]
    pts[P_J0 + 3] = pts[P_J0 + 2]
    pts[P_CHI0 + 3] = pts[P_CHI0 + 2]
    return pts


# The named workloads (sizes, quadrature order, reactions, bremsstrahlung)
WORKLOADS = {
    "cfg0_n10_brt06": dict(npts=4096, nquad=10, brem=BREM_BRT06, seed=0),
    "cfg1_grid128_n16_brt06": dict(npts=128 ** 3, nquad=16, brem=BREM_BRT06, seed=1),
    "cfg1_grid128_n16_hr98": dict(npts=128 ** 3, nquad=16, brem=BREM_HR98, seed=1),
}
