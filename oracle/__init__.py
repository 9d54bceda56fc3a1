"""ctypes access to the CPU checker libraries built by ``oracle/Makefile``.

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the cpu_baseline leg of bench.py -- as the checker, never as the product.

* ``liborc.so``      -- the CPU restatement of the gray M1 path (FP64)
* ``liborc_count.so`` -- the same code over a FLOP-counting real type
* ``liborc_bound.so`` -- the same code over a magnitude-propagating type: a
  running error bound m per output (see oracle/error_bound.cpp)
* ``liborc_fast.so``  -- the same code built for speed (-O3, AVX2/FMA): the
  timed CPU baseline of bench.py (not bit-exact: contraction is allowed)
* ``_ref/libbnsref.so`` -- the genuine reference compiled from its headers
  (gray n <= 10, spectral n <= 20: BS_N_MAX = 20 in bns_nurates.hpp:68)

Both ``orc_*`` and ``ref_*`` entry points share one signature per function,
so every wrapper below takes the library and a prefix.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {
    "orc": os.path.join(HERE, "liborc.so"),
    "count": os.path.join(HERE, "liborc_count.so"),
    "bound": os.path.join(HERE, "liborc_bound.so"),
    "fast": os.path.join(HERE, "liborc_fast.so"),
    "ref": os.path.join(HERE, "_ref", "libbnsref.so"),
    # the reference's own all-float build (REAL_TYPE=float, bns_nurates.hpp:29-34):
    # an upper bound of all-FP32 error, for comparison with the mixed form
    "ref_f32": os.path.join(HERE, "_ref", "libbnsref_f32.so"),
}
_cache = {}
_gp19 = None  # the GP19 table installed in the oracle builds (set_gp19_table)

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i64 = C.c_int64


def available(kind: str) -> bool:
    return os.path.exists(_LIBS[kind])


def build() -> None:
    """Run the oracle Makefile (also builds _ref/ when the reference tree is
    present)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load(kind: str = "orc"):
    if kind in _cache:
        return _cache[kind]
    path = _LIBS[kind]
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built (run make -C oracle)")
    lib = C.CDLL(path)
    p = "ref" if kind in ("ref", "ref_f32") else "orc"
    if kind == "bound":
        f = lib.orc_bound_m1_opacities
        f.argtypes = [_i64, C.c_int, _dp, _dp, C.c_int, _dp, _dp, _dp, C.c_void_p, _ip, _ip,
                      C.c_int, _dp, _dp, C.c_int]
        f.restype = C.c_int
        _cache[kind] = (lib, p)
        _apply_gp19(kind)
        return _cache[kind]
    if kind == "count":
        f = lib.orc_count_m1_opacities
        f.argtypes = [_i64, C.c_int, _dp, _dp, C.c_int, _dp, _dp, _dp, _ip, _ip, C.c_int, _dp, _i64p]
        f.restype = C.c_int
        _cache[kind] = (lib, p)
        return _cache[kind]
    sig = {
        "gauss_legendre": ([C.c_int, C.c_double, C.c_double, _dp, _dp], None),
        "fdi": ([C.c_int, _i64, _dp, _dp], None),
        "fermi_distr": ([_i64, _dp, _dp, _dp, _dp], None),
        "bessk1": ([_i64, _dp, _dp], None),
        "pair_kernels": ([_i64, _dp, _dp, _dp, _dp, _dp, _dp], None),
        "brem_brt06": ([_i64, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp], None),
        "brem_hr98": ([_i64, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int, _dp, _dp], None),
        "inel_kernels": ([_i64, _dp, _dp, _dp, _dp, _dp, _dp], None),
        "abs_opacity": ([_i64, _dp, _dp, _ip, C.c_int, _dp, _dp], None),
        "iso_scatt": ([_i64, _dp, _dp, _ip, _dp], None),
        "distr_from_m1": ([_i64, _dp, _dp], None),
        "distr_equilibrium": ([_i64, _dp, _dp], None),
        "m1_densities_eq": ([_i64, _dp, _dp], None),
        "total_nuf": ([_i64, _dp, _dp, C.c_int, _dp], None),
        "m1_opacities": ([_i64, C.c_int, _dp, _dp, C.c_int, _dp, _dp, _dp, C.c_void_p, _ip,
                          _ip, C.c_int, _dp, C.c_int], C.c_int),
        "iso_leg0": ([_i64, _dp, _dp, _ip, _dp], None),
        "spectral_opacities": ([_i64, C.c_int, _dp, _dp, _dp, C.c_void_p, _i64, _dp, _i64,
                                _ip, _ip, C.c_int, _dp, C.c_int], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, f"{p}_{name}")
        fn.argtypes = args
        fn.restype = res
    if p == "orc":
        for name, args in (("gp19_interp", [_i64, _dp, _dp, _dp, _dp, _dp]),
                           ("brem_gp19", [_i64, _dp, _dp, _dp, _dp, _dp, _dp, _dp])):
            fn = getattr(lib, f"orc_{name}")
            fn.argtypes = args
            fn.restype = None
    _cache[kind] = (lib, p)
    _apply_gp19(kind)
    return _cache[kind]


_GP19_KEYS = ("nb", "lognb", "ye", "T", "w", "logw")


def _apply_gp19(kind):
    """Installs the current GP19 table in one loaded oracle build."""
    if kind not in ("orc", "fast", "bound"):
        return
    lib, _ = _cache[kind]
    fn = lib.orc_bound_set_gp19_table if kind == "bound" else lib.orc_set_gp19_table
    fn.restype = None
    if _gp19 is None:
        fn(0, None, None, 0, None, 0, None, 0, None, None, None)
        return
    t = _gp19
    ax = [np.ascontiguousarray(t[k], dtype=np.float64) for k in _GP19_KEYS]
    data = np.ascontiguousarray(t["data"], dtype=np.float64)
    fn.argtypes = [C.c_int, _dp, _dp, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp, _dp, _dp]
    fn(len(ax[0]), ax[0], ax[1], len(ax[2]), ax[2], len(ax[3]), ax[3], len(ax[4]), ax[4], ax[5],
       data)
    fn.argtypes = None


def set_gp19_table(table):
    """Installs (a dict with the axes nb [fm^-3], lognb, ye, T [MeV], w
    [MeV], logw and data [nnb, nye, nT, nw]) or clears (None) the GP19 table
    in every oracle build; without one, BREM_GP19 adds zero, as the
    reference's table-free configuration does."""
    global _gp19
    _gp19 = None if table is None else dict(table)
    for kind in list(_cache):
        _apply_gp19(kind)


def gp19_interp(nb, ye, T, w):
    """gp19_interpolator of the installed table, elementwise (NaN outside)."""
    a = [_f64(np.broadcast_to(v, np.broadcast(nb, ye, T, w).shape)) for v in (nb, ye, T, w)]
    out = np.empty(a[0].size)
    _fn("orc", "gp19_interp")(out.size, *[x.ravel() for x in a], out)
    return out.reshape(a[0].shape)


def brem_gp19(om, omp, nb, T, yp):
    """BremKernelAbsGP19 (em, abs) of the installed table, elementwise."""
    a = [_f64(np.broadcast_to(v, np.broadcast(om, omp, nb, T, yp).shape)).ravel()
         for v in (om, omp, nb, T, yp)]
    em, ab = np.empty(a[0].size), np.empty(a[0].size)
    _fn("orc", "brem_gp19")(em.size, *a, em, ab)
    return em, ab


def _fn(kind, name):
    lib, p = load(kind)
    return getattr(lib, f"{p}_{name}")


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def gauss_legendre(n: int, kind: str = "orc", x1: float = 0.0, x2: float = 1.0):
    x = np.zeros(n)
    w = np.zeros(n)
    _fn(kind, "gauss_legendre")(n, x1, x2, x, w)
    return x, w


def fdi(k: int, y, kind: str = "orc"):
    y = _f64(y).ravel()
    out = np.empty_like(y)
    _fn(kind, "fdi")(k, y.size, y, out)
    return out


def pair_kernels(om, omp, T, mu_e, kind="orc"):
    om, omp, T, mu_e = (_f64(a).ravel() for a in np.broadcast_arrays(om, omp, T, mu_e))
    em = np.empty((om.size, 4))
    ab = np.empty((om.size, 4))
    _fn(kind, "pair_kernels")(om.size, om, omp, T, mu_e, em, ab)
    return em, ab


def brem(impl: str, om, omp, nb, T, yn, yp, medium=1, kind="orc"):
    om, omp, nb, T, yn, yp = (_f64(a).ravel() for a in np.broadcast_arrays(om, omp, nb, T, yn, yp))
    em = np.empty(om.size)
    ab = np.empty(om.size)
    if impl == "brt06":
        _fn(kind, "brem_brt06")(om.size, om, omp, nb, T, yn, yp, em, ab)
    else:
        _fn(kind, "brem_hr98")(om.size, om, omp, nb, T, yn, yp, int(medium), em, ab)
    return em, ab


def inel_kernels(om, omp, T, mu_e, kind="orc"):
    om, omp, T, mu_e = (_f64(a).ravel() for a in np.broadcast_arrays(om, omp, T, mu_e))
    em = np.empty((om.size, 4))
    ab = np.empty((om.size, 4))
    _fn(kind, "inel_kernels")(om.size, om, omp, T, mu_e, em, ab)
    return em, ab


def abs_opacity(om, eos9, params, stim=1, kind="orc"):
    om = _f64(om).ravel()
    em = np.empty((om.size, 2))
    ab = np.empty((om.size, 2))
    _fn(kind, "abs_opacity")(om.size, om, _f64(eos9), np.ascontiguousarray(params, np.int32),
                             int(stim), em, ab)
    return em, ab


def iso_scatt(om, eos9, params, kind="orc"):
    om = _f64(om).ravel()
    out = np.empty(om.size)
    _fn(kind, "iso_scatt")(om.size, om, _f64(eos9), np.ascontiguousarray(params, np.int32), out)
    return out


def distr_from_m1(pts, kind="orc"):
    pts = _f64(pts)
    d = np.empty((28, pts.shape[1]))
    _fn(kind, "distr_from_m1")(pts.shape[1], pts, d)
    return d


def distr_equilibrium(pts, kind="orc"):
    pts = _f64(pts)
    d = np.empty((28, pts.shape[1]))
    _fn(kind, "distr_equilibrium")(pts.shape[1], pts, d)
    return d


def m1_densities_eq(pts, distr, kind="orc"):
    pts = _f64(pts).copy()
    _fn(kind, "m1_densities_eq")(pts.shape[1], pts, _f64(distr))
    return pts


def total_nuf(om, distr28, s, kind="orc"):
    om = _f64(om).ravel()
    out = np.empty(om.size)
    _fn(kind, "total_nuf")(om.size, om, _f64(distr28), int(s), out)
    return out


def m1_opacities(pts, nquad, flags, params, stim=1, distr=None, kind="orc", nthreads=0,
                 nquad_1d=None):
    """Gray M1 opacities [20, npts] for the point block pts [21, npts]."""
    pts = _f64(pts)
    npts = pts.shape[1]
    n1 = nquad if nquad_1d is None else nquad_1d
    x1, w1 = gauss_legendre(n1, kind=kind)
    x2, w2 = gauss_legendre(nquad, kind=kind)
    out = np.zeros((20, npts))
    dptr = None
    if distr is not None:
        distr = _f64(distr)
        dptr = distr.ctypes.data_as(C.c_void_p)
    rc = _fn(kind, "m1_opacities")(npts, n1, x1, w1, nquad, x2, w2, pts, dptr,
                                   np.ascontiguousarray(flags, np.int32),
                                   np.ascontiguousarray(params, np.int32), int(stim), out,
                                   int(nthreads))
    if rc != 0:
        raise ValueError(f"{kind}_m1_opacities returned {rc}")
    return out


def _nu_arg(nu, npts):
    """(nu array, nnu, nu_ld): a 1-D grid is shared by every point, a 2-D
    [nnu, npts] array gives every point its own energies."""
    nu = _f64(nu)
    if nu.ndim == 1:
        return nu, nu.shape[0], 0
    if nu.ndim != 2 or nu.shape[1] != npts:
        raise ValueError("nu must be [nnu] or [nnu, npts]")
    return nu, nu.shape[0], npts


def spectral_opacities(pts, nu, nquad, flags, params, stim=0, distr=None, kind="orc",
                       nthreads=0):
    """Spectral opacities [16, nnu, npts] (j, kappa, j_s, kappa_s per
    species) for the point block pts [21, npts] at the energies nu."""
    pts = _f64(pts)
    npts = pts.shape[1]
    nu, nnu, nu_ld = _nu_arg(nu, npts)
    x, w = gauss_legendre(nquad, kind=kind)
    out = np.zeros((16, nnu, npts))
    dptr = None
    if distr is not None:
        distr = _f64(distr)
        dptr = distr.ctypes.data_as(C.c_void_p)
    rc = _fn(kind, "spectral_opacities")(npts, nquad, x, w, pts, dptr, nnu, nu, nu_ld,
                                         np.ascontiguousarray(flags, np.int32),
                                         np.ascontiguousarray(params, np.int32), int(stim),
                                         out, int(nthreads))
    if rc != 0:
        raise ValueError(f"{kind}_spectral_opacities returned {rc}")
    return out


def iso_leg0(om, eos9, params, kind="orc"):
    om = _f64(om)
    out = np.zeros_like(om)
    _fn(kind, "iso_leg0")(om.size, om, _f64(eos9), np.ascontiguousarray(params, np.int32), out)
    return out


def count_spectral_opacities(pts, nu, nquad, flags, params, stim=0):
    """Algorithmic operation counts of the spectral path over pts x nu (a
    shared energy grid): dict as count_m1_opacities, and the outputs."""
    lib, _ = load("count")
    pts = _f64(pts)
    npts = pts.shape[1]
    nu = _f64(nu)
    x, w = gauss_legendre(nquad)
    out = np.zeros((16, nu.shape[0], npts))
    counts = np.zeros(9, dtype=np.int64)
    f = lib.orc_count_spectral_opacities
    f.argtypes = [_i64, C.c_int, _dp, _dp, _dp, _i64, _dp, _ip, _ip, C.c_int, _dp, _i64p]
    f.restype = C.c_int
    f(npts, nquad, x, w, pts, nu.shape[0], nu, np.ascontiguousarray(flags, np.int32),
      np.ascontiguousarray(params, np.int32), int(stim), out, counts)
    names = ["add", "mul", "div", "sqrt", "exp", "log", "pow", "atan", "cbrt"]
    return dict(zip(names, (int(c) for c in counts))), out


def count_m1_opacities(pts, nquad, flags, params, stim=1):
    """Algorithmic operation counts of the gray path over pts:
    dict(add, mul, div, sqrt, exp, log, pow, atan, cbrt) and the outputs."""
    lib, _ = load("count")
    pts = _f64(pts)
    npts = pts.shape[1]
    x, w = gauss_legendre(nquad)
    out = np.zeros((20, npts))
    counts = np.zeros(9, dtype=np.int64)
    lib.orc_count_m1_opacities(npts, nquad, x, w, nquad, x, w, pts,
                               np.ascontiguousarray(flags, np.int32),
                               np.ascontiguousarray(params, np.int32), int(stim), out, counts)
    names = ["add", "mul", "div", "sqrt", "exp", "log", "pow", "atan", "cbrt"]
    return dict(zip(names, (int(c) for c in counts))), out


U = 2.0 ** -53  # unit roundoff of IEEE double


def bound_m1_opacities(pts, nquad, flags, params, stim=1, distr=None, nthreads=0,
                       nquad_1d=None):
    """Outputs [20, npts] and their running error magnitudes m [20, npts]:
    any double evaluation of the path lies within ~k U m of the exact value."""
    lib, _ = load("bound")
    pts = _f64(pts)
    npts = pts.shape[1]
    n1 = nquad if nquad_1d is None else nquad_1d
    x1, w1 = gauss_legendre(n1)
    x2, w2 = gauss_legendre(nquad)
    out = np.zeros((20, npts))
    mag = np.zeros((20, npts))
    dptr = None
    if distr is not None:
        distr = _f64(distr)
        dptr = distr.ctypes.data_as(C.c_void_p)
    lib.orc_bound_m1_opacities(npts, n1, x1, w1, nquad, x2, w2, pts, dptr,
                               np.ascontiguousarray(flags, np.int32),
                               np.ascontiguousarray(params, np.int32), int(stim), out, mag,
                               int(nthreads))
    return out, mag


def bound_spectral_opacities(pts, nu, nquad, flags, params, stim=0, distr=None, nthreads=0):
    """Spectral outputs [16, nnu, npts] and their running error magnitudes."""
    lib, _ = load("bound")
    pts = _f64(pts)
    npts = pts.shape[1]
    nu, nnu, nu_ld = _nu_arg(nu, npts)
    x, w = gauss_legendre(nquad)
    out = np.zeros((16, nnu, npts))
    mag = np.zeros((16, nnu, npts))
    dptr = None
    if distr is not None:
        distr = _f64(distr)
        dptr = distr.ctypes.data_as(C.c_void_p)
    f = lib.orc_bound_spectral_opacities
    f.argtypes = [_i64, C.c_int, _dp, _dp, _dp, C.c_void_p, _i64, _dp, _i64, _ip, _ip, C.c_int,
                  _dp, _dp, C.c_int]
    f.restype = C.c_int
    f(npts, nquad, x, w, pts, dptr, nnu, nu, nu_ld, np.ascontiguousarray(flags, np.int32),
      np.ascontiguousarray(params, np.int32), int(stim), out, mag, int(nthreads))
    return out, mag


def bound_distr_from_m1(pts):
    """Distribution parameters [28, npts] reconstructed from M1 and their
    running error magnitudes m [28, npts]."""
    lib, _ = load("bound")
    pts = _f64(pts)
    npts = pts.shape[1]
    d = np.zeros((28, npts))
    mag = np.zeros((28, npts))
    f = lib.orc_bound_distr_from_m1
    f.argtypes = [_i64, _dp, _dp, _dp]
    f.restype = None
    f(npts, pts, d, mag)
    return d, mag


# k of the parity bar |got - ref| <= rtol |ref| + k U m.  FP64 (rtol <= 1e-12):
# to first order each double evaluation of the formulas lies within U m of
# their exact value (Wilkinson's running error bound, error_bound.cpp), so two
# of them differ by at most 2 U m.  Reduced-precision forms are not covered
# by that argument; their bar keeps a factor 64 on the FP64 term.
K_FP64 = 2.0
K_REDUCED = 64.0


def parity_excess(got, ref, mag, rtol=1e-12, k=None):
    """Elementwise |got - ref| / (rtol |ref| + k U m): <= 1 passes the
    condition-aware parity bar (k: K_FP64 when rtol <= 1e-12, else
    K_REDUCED).  Equal values (incl. both NaN) give 0."""
    if k is None:
        k = K_FP64 if rtol <= 1e-12 else K_REDUCED
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    same = (got == ref) | (np.isnan(got) & np.isnan(ref))
    tol = rtol * np.abs(ref) + k * U * np.abs(mag) + 1e-300
    with np.errstate(invalid="ignore"):
        ex = np.abs(got - ref) / tol
    ex = np.where(same, 0.0, ex)
    return np.where(np.isnan(ex), np.inf, ex)
