"""Python surface of the reference's SWIG module ``bnsnurates``
(bindings/bnsnurates.i) for the opacity path, over the C ABI of
libbnsamd.so.

A script written for the reference's bindings (bindings/test_bindings.py)
runs unchanged with ``from bns_nurates_amd import bnsnurates as bns`` in
place of ``import bnsnurates as bns``:

* structs as attribute objects with the reference's field names:
  ``MyEOSParams``, ``NuDistributionParams``, ``M1Quantities``,
  ``MyQuadrature``, ``GreyOpacityParams`` (bns_nurates.hpp:106-454);
  ``cvar.quadrature_default`` is the (mutable) module-level default
  quadrature, as SWIG's ``cvar`` proxy of ``quadrature_default``;
* ``OpacityFlags`` / ``OpacityParams`` passed as dicts with the typemaps'
  keys (bnsnurates.i:58-85; ``brem_implementation`` a string "BRT06",
  "GP19", anything else HR98);
* ``M1Opacities`` / ``SpectralOpacities`` returned as dicts of 4-element
  lists (bnsnurates.i:138-206);
* the functions ``GaussLegendre``, ``NuEquilibriumParams``,
  ``ComputeM1DensitiesEq``, ``CalculateDistrParamsFromM1``,
  ``ComputeM1Opacities``, ``ComputeM1OpacitiesNotStimulated``,
  ``ComputeSpectralOpacitiesNotStimulatedAbs``,
  ``ComputeSpectralOpacitiesStimulatedAbs`` with the reference's argument
  order, and the ``print_*`` helpers of the module's Python code.

Every computation runs on the GPU through one module-level context (device
``$BNSAMD_DEVICE``, default 0), created on first use: there is no CPU path.
"""
from __future__ import annotations

import os
from types import SimpleNamespace

import numpy as np

from . import api

BS_N_MAX = 32
_NSP = 4


class _Struct:
    """Attribute struct; the list fields hold 4 species."""
    _lists: tuple = ()
    _scalars: dict = {}

    def __init__(self, **kw):
        for k, v in self._scalars.items():
            setattr(self, k, v)
        for k in self._lists:
            setattr(self, k, [0.0] * _NSP)
        for k, v in kw.items():
            setattr(self, k, v)

    def __repr__(self):
        body = ", ".join(f"{k}={getattr(self, k)!r}" for k in (*self._scalars, *self._lists))
        return f"{type(self).__name__}({body})"


class MyEOSParams(_Struct):
    """MyEOSParams (bns_nurates.hpp:289-306)."""
    _scalars = dict(nb=0.0, temp=0.0, ye=0.0, yp=0.0, yn=0.0, mu_p=0.0, mu_n=0.0, mu_e=0.0,
                    mu_mu=0.0, dU=0.0, dm_eff=0.0)


class NuDistributionParams(_Struct):
    """NuDistributionParams (bns_nurates.hpp:387-401)."""
    _lists = ("w_t", "temp_t", "eta_t", "w_f", "temp_f", "c_f", "beta_f")


class M1Quantities(_Struct):
    """M1Quantities (bns_nurates.hpp:409-416); H (the flux) is carried, not used."""
    _lists = ("n", "J", "chi")

    def __init__(self, **kw):
        super().__init__(**kw)
        if "H" not in kw:
            self.H = [[0.0] * 4 for _ in range(_NSP)]


class MyQuadrature(_Struct):
    """MyQuadrature (bns_nurates.hpp:106-130); points / w hold BS_N_MAX values."""
    _scalars = dict(type=0, alpha=0.0, dim=1, nx=20, ny=1, nz=1, x1=0.0, x2=1.0, y1=-42.0,
                    y2=-42.0, z1=-42.0, z2=-42.0)

    def __init__(self, **kw):
        super().__init__(**kw)
        if "points" not in kw:
            self.points = [0.0] * BS_N_MAX
        if "w" not in kw:
            self.w = [0.0] * BS_N_MAX


class GreyOpacityParams(_Struct):
    """GreyOpacityParams (bns_nurates.hpp:444-454) without kernel_pars (the
    reference's scratch); opacity_flags / opacity_pars are dicts."""

    def __init__(self, **kw):
        self.opacity_flags = dict(OPACITY_FLAGS_DEFAULT_ALL)
        self.opacity_pars = dict(OPACITY_PARAMS_DEFAULT_NONE)
        self.eos_pars = MyEOSParams()
        self.distr_pars = NuDistributionParams()
        self.m1_pars = M1Quantities()
        for k, v in kw.items():
            setattr(self, k, v)


OPACITY_FLAGS_DEFAULT_ALL = dict(use_abs_em=1, use_pair=1, use_brem=1, use_inelastic_scatt=1,
                                 use_iso=1)
OPACITY_FLAGS_DEFAULT_NONE = {k: 0 for k in OPACITY_FLAGS_DEFAULT_ALL}
OPACITY_PARAMS_DEFAULT_ALL = dict(use_dU=1, use_dm_eff=1, use_WM_ab=1, use_WM_sc=1, use_decay=1,
                                  brem_implementation="HR98", use_NN_medium_corr=1,
                                  neglect_blocking=1)
OPACITY_PARAMS_DEFAULT_NONE = dict(use_dU=0, use_dm_eff=0, use_WM_ab=0, use_WM_sc=0, use_decay=0,
                                   brem_implementation="HR98", use_NN_medium_corr=0,
                                   neglect_blocking=0)

# SWIG exposes C globals through ``cvar``: one mutable default quadrature
cvar = SimpleNamespace(quadrature_default=MyQuadrature(),
                       opacity_flags_default_all=dict(OPACITY_FLAGS_DEFAULT_ALL),
                       opacity_flags_default_none=dict(OPACITY_FLAGS_DEFAULT_NONE),
                       opacity_params_default_all=dict(OPACITY_PARAMS_DEFAULT_ALL),
                       opacity_params_default_none=dict(OPACITY_PARAMS_DEFAULT_NONE))

_ctx = None


def _context() -> api.Context:
    global _ctx
    if _ctx is None:
        _ctx = api.Context(int(os.environ.get("BNSAMD_DEVICE", "0")))
    return _ctx


# ---- dict <-> struct conversions (the typemaps of bnsnurates.i:58-206) --------
def _flags(d) -> api.OpacityFlags:
    return api.OpacityFlags(*(int(d[k]) for k in ("use_abs_em", "use_pair", "use_brem",
                                                   "use_inelastic_scatt", "use_iso")))


def _brem(v) -> int:
    if isinstance(v, str):
        return api.BREM_BRT06 if v.startswith("BRT06") else (
            api.BREM_GP19 if v.startswith("GP19") else api.BREM_HR98)
    return int(v)  # an int BremImpl value (0 HR98, 1 BRT06, 2 GP19)


def _params(d) -> api.OpacityParams:
    return api.OpacityParams(int(d["use_dU"]), int(d["use_dm_eff"]), int(d["use_WM_ab"]),
                             int(d["use_WM_sc"]), int(d["use_decay"]),
                             _brem(d["brem_implementation"]), int(d["use_NN_medium_corr"]),
                             int(d["neglect_blocking"]))


def _point_block(eos: MyEOSParams, m1: M1Quantities | None) -> np.ndarray:
    p = np.zeros((api.PT_NFIELDS, 1))
    p[:9, 0] = [eos.nb, eos.temp, eos.yp, eos.yn, eos.mu_p, eos.mu_n, eos.mu_e, eos.dU,
                eos.dm_eff]
    if m1 is not None:
        p[9:13, 0] = m1.n
        p[13:17, 0] = m1.J
        p[17:21, 0] = m1.chi
    return p


def _distr_block(d: NuDistributionParams) -> np.ndarray:
    return np.array([[v] for f in ("w_t", "temp_t", "eta_t", "w_f", "temp_f", "c_f", "beta_f")
                     for v in getattr(d, f)], dtype=np.float64)


def _distr_struct(block: np.ndarray) -> NuDistributionParams:
    d = NuDistributionParams()
    for i, f in enumerate(("w_t", "temp_t", "eta_t", "w_f", "temp_f", "c_f", "beta_f")):
        setattr(d, f, [float(v) for v in block[4 * i:4 * i + 4, 0]])
    return d


def _set_quadrature(ctx: api.Context, quad_1d: MyQuadrature, quad_2d: MyQuadrature):
    q1 = (np.asarray(quad_1d.points[:quad_1d.nx], np.float64),
          np.asarray(quad_1d.w[:quad_1d.nx], np.float64))
    q2 = (np.asarray(quad_2d.points[:quad_2d.nx], np.float64),
          np.asarray(quad_2d.w[:quad_2d.nx], np.float64))
    ctx.set_quadrature(quad_2d.nx, n1=quad_1d.nx, quad_2d=q2, quad_1d=q1)


# ---- functions ------------------------------------------------------------------
def GaussLegendre(quad: MyQuadrature) -> None:
    """GaussLegendre (integration.hpp:28-87): nodes and weights of quad.nx
    points on [x1, x2]."""
    if not 1 <= quad.nx <= BS_N_MAX:
        raise ValueError(f"quad.nx must be in [1, {BS_N_MAX}]")
    x, w = api.gauss_legendre(quad.nx, quad.x1, quad.x2)
    quad.points = list(x) + [0.0] * (BS_N_MAX - quad.nx)
    quad.w = list(w) + [0.0] * (BS_N_MAX - quad.nx)


def NuEquilibriumParams(eos_pars: MyEOSParams) -> NuDistributionParams:
    """NuEquilibriumParams (distribution.hpp:253-290)."""
    pts = _point_block(eos_pars, None)
    distr = np.zeros((api.DI_NFIELDS, 1))
    ctx = _context()
    ctx._check(ctx.lib.bnsamd_equilibrium_host(ctx.h, 1, pts.ctypes.data, distr.ctypes.data),
               "bnsamd_equilibrium_host")
    return _distr_struct(distr)


def ComputeM1DensitiesEq(eos_pars: MyEOSParams, distr_pars: NuDistributionParams,
                         m1_pars: M1Quantities) -> None:
    """ComputeM1DensitiesEq (distribution.hpp:447-473): fills m1_pars.n, J."""
    pts = _point_block(eos_pars, None)
    distr = _distr_block(distr_pars)
    ctx = _context()
    ctx._check(ctx.lib.bnsamd_m1_densities_eq_host(ctx.h, 1, distr.ctypes.data,
                                                   pts.ctypes.data),
               "bnsamd_m1_densities_eq_host")
    m1_pars.n = [float(v) for v in pts[9:13, 0]]
    m1_pars.J = [float(v) for v in pts[13:17, 0]]


def CalculateDistrParamsFromM1(m1_pars: M1Quantities,
                               eos_pars: MyEOSParams) -> NuDistributionParams:
    """CalculateDistrParamsFromM1 (distribution.hpp:323-333)."""
    pts = _point_block(eos_pars, m1_pars)
    distr = np.zeros((api.DI_NFIELDS, 1))
    ctx = _context()
    ctx._check(ctx.lib.bnsamd_distr_from_m1_host(ctx.h, 1, pts.ctypes.data, distr.ctypes.data),
               "bnsamd_distr_from_m1_host")
    return _distr_struct(distr)


def _gray(quad_1d, quad_2d, g: GreyOpacityParams, stim: int) -> dict:
    ctx = _context()
    _set_quadrature(ctx, quad_1d, quad_2d)
    out = ctx.compute_m1_opacities(_point_block(g.eos_pars, g.m1_pars),
                                   flags=_flags(g.opacity_flags), params=_params(g.opacity_pars),
                                   stim_abs=stim, distr=_distr_block(g.distr_pars))
    return {name: [float(v) for v in out[4 * i:4 * i + 4, 0]]
            for i, name in enumerate(("eta_0", "kappa_0_a", "eta", "kappa_a", "kappa_s"))}


def ComputeM1Opacities(quad_1d: MyQuadrature, quad_2d: MyQuadrature,
                       my_grey_opacity_params: GreyOpacityParams) -> dict:
    """ComputeM1Opacities (m1_opacities.hpp:1306-1313): a dict of eta_0,
    kappa_0_a, eta, kappa_a, kappa_s (4 species each)."""
    return _gray(quad_1d, quad_2d, my_grey_opacity_params, 1)


def ComputeM1OpacitiesNotStimulated(quad_1d: MyQuadrature, quad_2d: MyQuadrature,
                                    my_grey_opacity_params: GreyOpacityParams) -> dict:
    """ComputeM1OpacitiesNotStimulated (m1_opacities.hpp:1297-1304)."""
    return _gray(quad_1d, quad_2d, my_grey_opacity_params, 0)


def _spectral(nu: float, quad_1d: MyQuadrature, g: GreyOpacityParams, stim: int) -> dict:
    ctx = _context()
    _set_quadrature(ctx, quad_1d, quad_1d)
    out = ctx.compute_spectral_opacities(_point_block(g.eos_pars, g.m1_pars),
                                         np.array([float(nu)]), flags=_flags(g.opacity_flags),
                                         params=_params(g.opacity_pars), stim_abs=stim,
                                         distr=_distr_block(g.distr_pars))
    return {name: [float(v) for v in out[4 * i:4 * i + 4, 0, 0]]
            for i, name in enumerate(("j", "kappa", "j_s", "kappa_s"))}


def ComputeSpectralOpacitiesNotStimulatedAbs(nu: float, quad_1d: MyQuadrature,
                                             my_grey_opacity_params: GreyOpacityParams) -> dict:
    """ComputeSpectralOpacitiesNotStimulatedAbs (m1_opacities.hpp:1505-1641):
    a dict of j, kappa, j_s, kappa_s (4 species each)."""
    return _spectral(nu, quad_1d, my_grey_opacity_params, 0)


def ComputeSpectralOpacitiesStimulatedAbs(nu: float, quad_1d: MyQuadrature,
                                          my_grey_opacity_params: GreyOpacityParams) -> dict:
    """ComputeSpectralOpacitiesStimulatedAbs (m1_opacities.hpp:1644-1660)."""
    return _spectral(nu, quad_1d, my_grey_opacity_params, 1)


# ---- printing helpers of the module's Python code ------------------------------
def _on(v) -> int:
    return int(bool(v))


def print_reactions(opacity_flags: dict) -> None:
    rows = [("Neutrino abs on nucleons - e+/e- capture", "use_abs_em"),
            ("Elastic scattering on nucleons", "use_iso"),
            ("Nucleon-nucleon bremmstrahlung", "use_brem"),
            ("e+ e- annihilation (and inverse)", "use_pair"),
            ("Inelastic scattering on e+/e-", "use_inelastic_scatt")]
    print("# Included reactions:")
    for label, key in rows:
        print(f"# {label:<41s} : {_on(opacity_flags[key])}")
    print("")


def print_corrections(opacity_pars: dict) -> None:
    rows = [("dU correction", "use_dU"), ("Effective mass correction", "use_dm_eff"),
            ("Weak magnetism on charged reactions", "use_WM_ab"),
            ("Weak magnetism on neutral reactions", "use_WM_sc"),
            ("Include (inverse) nucleon decays", "use_decay"),
            ("Neglect blocking factors", "neglect_blocking")]
    print("# Included corrections:")
    for label, key in rows:
        print(f"# {label:<36s} : {_on(opacity_pars[key])}")
    print(f"# {'Bremsstrahlung implementation':<36s} : {opacity_pars['brem_implementation']}")
    print(f"# {'Medium Fischer+16 correction to brem':<36s} : "
          f"{_on(opacity_pars['use_NN_medium_corr'])}")
    print("")


def print_neutrino_quantities(m1_pars: M1Quantities) -> None:
    print("    nue            anue            nux            anux")
    print("n  " + "  ".join(f"{v:13.6e}" for v in m1_pars.n) + "     (cm^-3)")
    print("J  " + "  ".join(f"{v:13.6e}" for v in m1_pars.J) + " (MeV cm^-3)")
    print("chi " + "   ".join(f"{v:12.10f}" for v in m1_pars.chi))
    print("")


_SPECIES = (" nue", "anue", " nux", "anux")


def print_spectral_rates(spec_rates: dict) -> None:
    print("      j             j_s           kappa         kappa_s")
    for s, lab in enumerate(_SPECIES):
        print(lab + "".join(f" {spec_rates[k][s]:13.6e}" for k in ("j", "j_s", "kappa", "kappa_s")))
    print("")


def print_integrated_rates(gray_rates: dict) -> None:
    print("      eta0          eta1          kappa0        kappa1        scat1")
    for s, lab in enumerate(_SPECIES):
        print(lab + "".join(f" {gray_rates[k][s]:13.6e}"
                            for k in ("eta_0", "eta", "kappa_0_a", "kappa_a", "kappa_s")))
    print("")
