"""Synthetic GP19 response tables for the tests.  The reference's own
table (GP19Table.hpp) is absent from its tree, so the GP19 physics values
are parity unpinned; these tables pin the interpolator and the
extrapolations by their mathematics instead.  The axes follow the ranges the
reference's code assumes: nb from 1e-4 to 1 fm^-3 (log), Ye up to 0.5, T from
2 to 50 MeV with the 2, 4, 48 and 50 MeV nodes its extrapolations read, w
from 0.05 to 398 MeV (log)."""
import numpy as np


def axes():
    nb = np.logspace(-4.0, 0.0, 13)
    ye = np.array([0.01, 0.05, 0.1, 0.2, 0.3, 0.4, 0.5])
    T = np.array([2.0, 3.0, 4.0, 6.0, 8.0, 12.0, 16.0, 24.0, 32.0, 40.0, 48.0, 50.0])
    w = np.logspace(np.log10(0.05), np.log10(398.0), 24)
    return nb, ye, T, w


def table(kind="smooth"):
    """A table dict (nb, lognb, ye, T, w, logw, data[nb, ye, T, w]).
    smooth: a positive response that falls with w / T; multilinear: a
    function multilinear in (log nb, Ye, T, log w), which the interpolation
    reproduces exactly."""
    nb, ye, T, w = axes()
    ln, lw = np.log(nb), np.log(w)
    L, Y, TT, LW = np.meshgrid(ln, ye, T, lw, indexing="ij")
    if kind == "multilinear":
        data = (3.0 + 0.2 * L + 1.5 * Y + 0.01 * TT + 0.3 * LW + 0.02 * L * TT
                - 0.4 * Y * LW + 0.001 * L * Y * TT * LW)
    else:
        W = np.exp(LW)
        data = (1e-3 * np.exp(0.3 * (L - np.log(1e-2))) * (1.0 + Y) * TT ** -1.2
                * np.exp(-W / (4.0 * TT)) / (1.0 + (W / TT) ** 2))
    return {"nb": nb, "lognb": ln, "ye": ye, "T": T, "w": w, "logw": lw, "data": data}


def multilinear_value(lnb, ye, T, lw):
    return (3.0 + 0.2 * lnb + 1.5 * ye + 0.01 * T + 0.3 * lw + 0.02 * lnb * T
            - 0.4 * ye * lw + 0.001 * lnb * ye * T * lw)
