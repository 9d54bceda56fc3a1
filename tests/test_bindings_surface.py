"""The Python surface of the reference's SWIG module (bindings/bnsnurates.i)
in bns_nurates_amd.bnsnurates.

CPU: struct objects and defaults, the dict typemaps (brem_implementation as a
string), GaussLegendre (host code of the library) against the reference's
nodes, the printing helpers.
GPU: the flow of the reference's bindings script (bindings/test_bindings.py:
point A of Chiesa+25, n = 6, HR98 with its corrections; part 1 equilibrium
distributions, part 2 distributions reconstructed from M1), spectral and
gray rates against the CPU oracle at the parity bar.
"""
import numpy as np
import pytest

import oracle as O
from bns_nurates_amd import api
from bns_nurates_amd import bnsnurates as bns

# point A of the reference's bindings script (bindings/test_bindings.py:14-36)
POINT_A = dict(nb=4.208366627847035e+38 * 1e-21, temp=12.406403541564941,
               ye=0.07158458232879639, mu_e=187.1814489, mu_p=1011.01797737,
               mu_n=1221.59013681, dU=18.92714728, dm=280.16495513 - 278.87162217)
N_M1 = [3.739749408027436e+33, 1.2174961961689319e+35, 2.2438496448164613e+34,
        2.2438496448164613e+34]
J_M1 = [1.246583136009145e+35, 5.360307484839323e+36, 8.726081952064015e+35,
        8.726081952064015e+35]
FLAGS = {"use_abs_em": True, "use_pair": True, "use_brem": True, "use_inelastic_scatt": True,
         "use_iso": True}
PARS = {"use_dU": True, "use_dm_eff": False, "use_WM_ab": True, "use_WM_sc": True,
        "use_decay": True, "brem_implementation": "HR98", "neglect_blocking": False,
        "use_NN_medium_corr": True}


def eos_point_a():
    e = bns.MyEOSParams()
    a = POINT_A
    e.nb, e.temp, e.ye = a["nb"], a["temp"], a["ye"]
    e.yn, e.yp = 1.0 - a["ye"], a["ye"]
    e.mu_e, e.mu_n, e.mu_p = a["mu_e"], a["mu_n"], a["mu_p"]
    e.dU, e.dm_eff = a["dU"], a["dm"]
    return e


def test_structs_and_dict_typemaps():
    q = bns.MyQuadrature()
    assert q.nx == 20 and q.x1 == 0.0 and q.x2 == 1.0 and len(q.points) == bns.BS_N_MAX
    g = bns.GreyOpacityParams()
    assert set(g.opacity_flags) == set(FLAGS) and set(g.opacity_pars) == set(PARS)
    assert bns._params(dict(PARS, brem_implementation="BRT06")).brem_implementation == \
        api.BREM_BRT06
    assert bns._params(dict(PARS, brem_implementation="GP19")).brem_implementation == \
        api.BREM_GP19
    assert bns._params(PARS).brem_implementation == api.BREM_HR98
    f = bns._flags(FLAGS)
    assert (f.use_abs_em, f.use_pair, f.use_brem, f.use_inelastic_scatt, f.use_iso) == \
        (1, 1, 1, 1, 1)
    m = bns.M1Quantities()
    assert m.n == [0.0] * 4 and len(m.H) == 4
    # the point block carries the EOS fields in the C ABI order
    p = bns._point_block(eos_point_a(), None)
    assert p.shape == (21, 1) and p[1, 0] == POINT_A["temp"] and p[6, 0] == POINT_A["mu_e"]


def test_gauss_legendre_matches_reference_nodes(golden):
    g = golden("gauss_legendre.npz")
    q = bns.cvar.quadrature_default
    saved = q.nx
    try:
        for n in (2, 6, 10, 20):
            q.nx = n
            bns.GaussLegendre(q)
            assert np.array_equal(q.points[:n], g[f"x{n}"]) and np.array_equal(q.w[:n], g[f"w{n}"])
    finally:
        q.nx = saved


def test_print_helpers(capsys):
    bns.print_reactions(FLAGS)
    bns.print_corrections(PARS)
    m = bns.M1Quantities(n=[1.0] * 4, J=[2.0] * 4, chi=[1 / 3] * 4)
    bns.print_neutrino_quantities(m)
    bns.print_spectral_rates({k: [1.0, 2.0, 3.0, 4.0] for k in ("j", "j_s", "kappa", "kappa_s")})
    bns.print_integrated_rates({k: [1.0] * 4 for k in ("eta_0", "eta", "kappa_0_a", "kappa_a",
                                                        "kappa_s")})
    out = capsys.readouterr().out
    assert "Bremsstrahlung implementation" in out and "HR98" in out
    assert out.count("anux") == 3 and "1.000000e+00" in out  # header + two tables


@pytest.mark.gpu
def test_reference_bindings_script_flow():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from tests.gpu_util import check_parity, say
    nu_energy = 10.0
    eos = eos_point_a()
    quad = bns.MyQuadrature()
    quad.nx = 6
    bns.GaussLegendre(quad)
    flags = np.array([int(FLAGS[k]) for k in ("use_abs_em", "use_pair", "use_brem",
                                              "use_inelastic_scatt", "use_iso")], np.int32)
    prm = np.array([1, 0, 1, 1, 1, 0, 1, 0], np.int32)  # PARS as the oracle's int array
    worst = 0.0
    for part in (1, 2):
        g = bns.GreyOpacityParams()
        g.eos_pars, g.opacity_flags, g.opacity_pars = eos, FLAGS, PARS
        m1 = bns.M1Quantities()
        if part == 1:
            distr = bns.NuEquilibriumParams(eos)
            bns.ComputeM1DensitiesEq(eos, distr, m1)
            m1.chi = [1.0 / 3.0] * 4
        else:
            m1.n = [x * 1e-21 for x in N_M1]
            m1.J = [x * 1e-21 for x in J_M1]
            m1.chi = [1.0 / 3.0] * 4
            distr = bns.CalculateDistrParamsFromM1(m1, eos)
        g.m1_pars, g.distr_pars = m1, distr
        say(f"bindings flow, part {part}")
        spec = bns.ComputeSpectralOpacitiesNotStimulatedAbs(nu_energy, quad, g)
        gray = bns.ComputeM1Opacities(quad, quad, g)
        # the oracle on the same inputs (the distribution the module produced)
        pts = bns._point_block(eos, m1)
        dblk = bns._distr_block(distr)
        if part == 1:
            d_ref = O.distr_equilibrium(pts)
            np.testing.assert_allclose(dblk, d_ref, rtol=1e-15, atol=0)
        else:
            d_ref, d_mag = O.bound_distr_from_m1(pts)
            assert O.parity_excess(dblk, d_ref, d_mag).max() <= 1.0
        s_ref, s_mag = O.bound_spectral_opacities(pts, np.array([nu_energy]), 6, flags, prm,
                                                  stim=0, distr=dblk)
        got_s = np.array([spec[k] for k in ("j", "kappa", "j_s", "kappa_s")]).reshape(16, 1)
        worst = max(worst, check_parity(f"spectral part {part}", got_s, s_ref.reshape(16, 1),
                                        s_mag.reshape(16, 1)))
        g_ref, g_mag = O.bound_m1_opacities(pts, 6, flags, prm, stim=1, distr=dblk)
        got_g = np.array([gray[k] for k in ("eta_0", "kappa_0_a", "eta", "kappa_a",
                                            "kappa_s")]).reshape(20, 1)
        worst = max(worst, check_parity(f"gray part {part}", got_g, g_ref, g_mag))
        assert all(v >= 0 for k in spec for v in spec[k])
    assert worst <= 1.0


def test_thc_export(tmp_path):
    """The thorn export writes the headers and the library, and a thorn
    source that includes the exported header compiles with the reference's
    names; a second export writes nothing."""
    import subprocess
    from bns_nurates_amd import export_thc
    dest = tmp_path / "src"
    dest.mkdir()
    first = export_thc.export(str(dest), log=lambda *a: None)
    assert sorted(first) == ["bns_nurates_amd.h", "bns_nurates_amd.hpp", "libbnsamd.so"]
    assert export_thc.export(str(dest), log=lambda *a: None) == []
    src = dest / "weakrates.cpp"
    src.write_text('#include "bns_nurates_amd.hpp"\n'
                   "M1Opacities f(MyQuadrature* q, GreyOpacityParams* g) {\n"
                   "  return ComputeM1Opacities(q, q, g);\n}\n")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", str(dest), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
