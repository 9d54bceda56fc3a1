"""Write the golden fixtures under tests/golden/ from the genuine reference.

Runs the reference's own functions, compiled from /root/reference/include by
oracle/Makefile into oracle/_ref/libbnsref.so, on seeded and hand-made inputs
and stores inputs + outputs as .npz (numeric arrays only, no pickles).  Run
it where the reference tree exists:

    make -C oracle && python tests/golden/make_golden.py            # all fixtures
    python tests/golden/make_golden.py spectral                     # one group

The fixtures are data: they hold no reference source.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from bns_nurates_amd import workloads as W  # noqa: E402


def save(name, **arrays):
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **arrays)
    print(f"{name}: {os.path.getsize(path) / 1024:.1f} KiB")


def edge_points() -> np.ndarray:
    """Hand-made points that reach the branches and thresholds of the path."""
    rows = []

    def pt(nb=1e-8, T=5.0, yp=0.3, mu_e=10.0, mu_hat=None, dU=0.0, n=None, J=None, chi=None):
        yn = 1.0 - yp
        mu_p = W.M_P - 2.0
        mu_n = mu_p + (W.Q_NP + 1.0 if mu_hat is None else W.Q_NP + mu_hat)
        n = n if n is not None else [1e-6, 2e-6, 3e-6, 3e-6]
        J = J if J is not None else [1e-5, 3e-5, 6e-5, 6e-5]
        chi = chi if chi is not None else [0.5, 0.6, 0.7, 0.7]
        rows.append([nb, T, yp, yn, mu_p, mu_n, mu_e, dU, W.Q_NP] + list(n) + list(J) + list(chi))

    pt()                                              # plain
    pt(T=0.5, mu_e=20.0)                              # eta_e = 40: asymptotic s_beta
    pt(T=0.5, mu_e=15.0)                              # eta_e == 30 exactly: else-branch
    pt(T=0.5, mu_e=300.0)                             # eta_e = 600: pair cutoff, NES Taylor
    pt(T=2.0, mu_e=200.0, nb=1e-4)                    # eta_e = 100
    pt(T=0.7, mu_e=1e-6)                              # eta_e ~ 0
    pt(mu_hat=0.0)                                    # |mu_hat| < 0.01 backup
    pt(mu_hat=-3.0, yp=0.1)                           # negative eta_np backup
    pt(dU=2.5)                                        # dU correction
    pt(yp=0.0)                                        # no protons
    pt(n=[0.0, 2e-6, 3e-6, 0.0])                      # n == 0 guards
    pt(n=[1e-25, 1e-22, 3e-6, 3e-6], J=[1e-30, 1e-26, 6e-5, 6e-5])  # below thresholds
    pt(chi=[1.0, 1.0, 1.0, 1.0])                      # pure thin
    pt(chi=[1.0 / 3.0] * 4)                           # pure thick
    pt(n=[1e-3, 1e-3, 1e-3, 1e-3], J=[1e-4, 1e-4, 1e-4, 1e-4])  # y > y3: eta_t = 20 branch
    pt(n=[1e-12, 1e-12, 1e-12, 1e-12], J=[1e-9, 1e-9, 1e-9, 1e-9])  # y < y1
    pt(nb=1e-4, T=30.0, mu_e=80.0, yp=0.2)            # hot dense
    pt(nb=1e-12, T=0.6, mu_e=0.05, yp=0.5)            # cold dilute
    return np.ascontiguousarray(np.array(rows, dtype=np.float64).T)


def spectral_test_points() -> np.ndarray:
    """The six thermodynamic points of the reference's spectral test
    (tests/tests.hpp:893-968) as a point block; the M1 fields are left at 0
    (the test uses equilibrium distributions)."""
    rows = [
        # nb [nm^-3], T, Yp, Yn, mu_p, mu_n, mu_e, dU, mn* - mp*
        (4.208366627847035e+38 * 1.0E-21, 12.406403541564941, 0.07158458, 0.92841542,
         1011.01797737, 1221.59013681, 187.1814489, 18.92714728, 280.16495513 - 278.87162217),
        (6.051432235073535e+37 * 1.0E-21, 16.887954711914062, 0.04352445, 0.91736565,
         843.50691102, 943.76817391, 82.3367694, 35.09393792, 733.97246814 - 732.67912955),
        (6.0146859596912e+36 * 1.0E-21, 8.919591903686523, 0.02018168, 0.88856839,
         889.20482938, 930.55123963, 36.56143955, 4.98247735, 915.90721218 - 914.61387308),
        (6.128721244299242e+35 * 1.0E-21, 6.698440074920654, 0.08023092, 0.8442462,
         903.09559414, 920.74113958, 19.49956921, 0.46637764, 937.06605368 - 935.77271381),
        (6.07728949804485e+34 * 1.0E-21, 3.586622953414917, 0.11611014, 0.84457226,
         916.18996071, 924.65692374, 8.98668385, 0.0454689, 939.30610577 - 938.01276696),
        (6.192921723010272e+33 * 1.0E-21, 2.144158124923706, 0.17166997, 0.79639333,
         922.72216948, 927.30360027, 4.20812739, 0.0007919, 939.56535 - 938.27201),
    ]
    pts = np.zeros((21, len(rows)))
    pts[:9] = np.array(rows, dtype=np.float64).T
    pts[17:21] = 1.0 / 3.0
    return np.ascontiguousarray(pts)


def spectral_fixtures():
    """Spectral opacities (m1_opacities.hpp:1505-1660) end to end."""
    # the reference's spectral test: 6 points x 100 energies 10^(-1 + 4 i / 100),
    # equilibrium distributions, all reactions, opacity_params_default_none
    # (HR98), n = 10, no stimulated absorption (tests/tests.hpp:861-1261)
    tp = spectral_test_points()
    deq = O.distr_equilibrium(tp, kind="ref")
    tp = O.m1_densities_eq(tp, deq, kind="ref")
    ebins = np.array([10.0 ** (-1.0 + i * (4.0 / 100.0)) for i in range(100)])
    prm_none = W.params_default_none(W.BREM_HR98)
    out = O.spectral_opacities(tp, ebins, 10, W.FLAGS_ALL, prm_none, stim=0, distr=deq,
                               kind="ref")
    save("spectral_test6_n10_hr98.npz", pts=tp, distr=deq, nu=ebins, out=out, nquad=10,
         flags=W.FLAGS_ALL, params=prm_none, stim=0)

    # seeded + edge points, M1-reconstructed distributions, reaction and
    # parameter variants, both absorption forms, n = 4..20
    p0 = W.make_points(96, seed=21)
    mix = np.concatenate([edge_points(), p0], axis=1)
    nu = np.exp(np.linspace(np.log(0.2), np.log(600.0), 12))
    variants = {}
    cases = [
        ("brt06_all_n10_stim", 10, W.FLAGS_ALL, W.params_default_all(W.BREM_BRT06), 1),
        ("brt06_block_n16_nostim", 16, W.FLAGS_ALL,
         np.array([1, 1, 1, 1, 1, 1, 1, 0], np.int32), 0),
        ("hr98_med_n20_nostim", 20, W.FLAGS_ALL, np.array([0, 1, 0, 1, 0, 0, 1, 0], np.int32), 0),
        ("hr98_none_n6_stim", 6, W.FLAGS_ALL, W.params_default_none(W.BREM_HR98), 1),
        ("gp19_n4_nostim", 4, W.FLAGS_ALL, W.params_default_all(W.BREM_GP19), 0),
    ]
    for i in range(5):
        f = np.zeros(5, np.int32)
        f[i] = 1
        cases.append((f"only{i}_n10_nostim", 10, f, np.array([1, 1, 1, 1, 1, 1, 1, 0], np.int32),
                      0))
    for tag, nq, fl, prm, stim in cases:
        variants[f"out_{tag}"] = O.spectral_opacities(mix, nu, nq, fl, prm, stim=stim,
                                                      kind="ref")
        variants[f"nquad_{tag}"] = np.int32(nq)
        variants[f"flags_{tag}"] = fl
        variants[f"params_{tag}"] = prm
        variants[f"stim_{tag}"] = np.int32(stim)
    # every point its own energies
    rng = np.random.Generator(np.random.PCG64(23))
    nu_pp = np.exp(rng.uniform(np.log(0.1), np.log(1000.0), size=(5, mix.shape[1])))
    variants["out_perpoint_n10_stim"] = O.spectral_opacities(
        mix, nu_pp, 10, W.FLAGS_ALL, W.params_default_all(W.BREM_BRT06), stim=1, kind="ref")
    # IsoScattLegCoeff (l = 0) on its own, with and without weak magnetism
    eos9 = np.ascontiguousarray(mix[:9, 20])
    e_iso = np.exp(np.linspace(np.log(0.05), np.log(400.0), 40))
    variants["iso_leg0_wm"] = O.iso_leg0(e_iso, eos9, W.params_default_all(1), kind="ref")
    variants["iso_leg0_nowm"] = O.iso_leg0(e_iso, eos9, W.params_default_none(1), kind="ref")
    save("spectral_variants.npz", pts=mix, nu=nu, nu_perpoint=nu_pp, iso_eos9=eos9,
         iso_omega=e_iso, **variants)


def main(groups=None):
    if not O.available("ref"):
        raise SystemExit("oracle/_ref/libbnsref.so missing: run make -C oracle where "
                         "/root/reference exists")
    if groups:
        for g in groups:
            {"spectral": spectral_fixtures}[g]()
        return

    # --- Gauss-Legendre (integration.hpp:28-87) ---------------------------
    gl = {}
    for n in range(1, 21):
        x, w = O.gauss_legendre(n, kind="ref")
        gl[f"x{n}"], gl[f"w{n}"] = x, w
    for n in (2, 3, 4):
        x, w = O.gauss_legendre(n, kind="ref", x1=-1.0, x2=1.0)
        gl[f"sym_x{n}"], gl[f"sym_w{n}"] = x, w
    save("gauss_legendre.npz", **gl)

    # --- complete Fermi-Dirac integrals F_0..F_5 ----------------------------
    y = np.concatenate([
        np.linspace(-60.0, 60.0, 2401),
        -np.logspace(np.log10(60.0), 3, 100), np.logspace(np.log10(60.0), 3, 100),
        [-2.0, np.nextafter(-2.0, 0.0), np.nextafter(-2.0, -3.0), 0.0, -0.0, 1e-300,
         -1e-300, 2.0, 700.0, -700.0, 745.0, -745.0, 1e4, -1e4]])
    save("fdi.npz", y=y, **{f"F{k}": O.fdi(k, y, kind="ref") for k in range(6)})

    # --- reaction kernels ---------------------------------------------------
    rng = np.random.Generator(np.random.PCG64(7))
    m = 2000
    om = np.exp(rng.uniform(np.log(1e-2), np.log(1e3), m))
    omp = np.exp(rng.uniform(np.log(1e-2), np.log(1e3), m))
    T = np.exp(rng.uniform(np.log(0.5), np.log(50.0), m))
    mu_e = T * np.concatenate([rng.uniform(-5, 5, m // 4), np.exp(rng.uniform(-5, np.log(800.), m - m // 4))])
    om[:8] = omp[:8]                                   # diagonal cells (w == w')
    nb = np.exp(rng.uniform(np.log(1e-7), np.log(1.0), m))
    yp = rng.uniform(0.01, 0.6, m)
    yn = 1.0 - yp
    pe, pa = O.pair_kernels(om, omp, T, mu_e, kind="ref")
    ie, ia = O.inel_kernels(om, omp, T, mu_e, kind="ref")
    be, ba = O.brem("brt06", om, omp, nb, T, yn, yp, kind="ref")
    he, ha = O.brem("hr98", om, omp, nb, T, yn, yp, medium=1, kind="ref")
    he0, ha0 = O.brem("hr98", om, omp, nb, T, yn, yp, medium=0, kind="ref")
    k1x = np.exp(rng.uniform(np.log(1e-3), np.log(50.0), 500))
    save("kernels.npz", om=om, omp=omp, T=T, mu_e=mu_e, nb=nb, yp=yp, yn=yn,
         pair_em=pe, pair_abs=pa, inel_em=ie, inel_abs=ia, brt06_em=be, brt06_abs=ba,
         hr98_em=he, hr98_abs=ha, hr98_nomed_em=he0, hr98_nomed_abs=ha0,
         k1_x=k1x, k1=_bessk1_ref(k1x))

    # --- beta processes, isoenergetic scattering ----------------------------
    pts = W.make_points(24, seed=11)
    ebeta = np.exp(np.linspace(np.log(0.05), np.log(400.0), 60))
    beta = {"omega": ebeta, "pts": pts}
    for tag, prm in (("all", W.params_default_all(1)), ("none", W.params_default_none(1)),
                     ("mixed", np.array([1, 0, 1, 0, 0, 1, 0, 0], np.int32))):
        ems, abs_, emn, absn, iso = [], [], [], [], []
        for i in range(pts.shape[1]):
            eos9 = np.ascontiguousarray(pts[:9, i])
            if tag == "mixed":
                eos9[7] = 1.7  # dU
            e1, a1 = O.abs_opacity(ebeta, eos9, prm, stim=1, kind="ref")
            e0, a0 = O.abs_opacity(ebeta, eos9, prm, stim=0, kind="ref")
            ems.append(e1); abs_.append(a1); emn.append(e0); absn.append(a0)
            iso.append(O.iso_scatt(ebeta, eos9, prm, kind="ref"))
        beta[f"params_{tag}"] = prm
        beta[f"stim_em_{tag}"], beta[f"stim_abs_{tag}"] = np.array(ems), np.array(abs_)
        beta[f"em_{tag}"], beta[f"abs_{tag}"] = np.array(emn), np.array(absn)
        beta[f"iso_{tag}"] = np.array(iso)
    save("beta_iso.npz", **beta)

    # --- distribution function --------------------------------------------
    dp = np.concatenate([W.make_points(256, seed=12), edge_points()], axis=1)
    dfm = O.distr_from_m1(dp, kind="ref")
    deq = O.distr_equilibrium(dp, kind="ref")
    meq = O.m1_densities_eq(dp, deq, kind="ref")
    enu = np.exp(np.linspace(np.log(1e-3), np.log(2e3), 60))
    tnuf = np.array([[O.total_nuf(enu, np.ascontiguousarray(dfm[:, i]), s, kind="ref")
                      for s in range(4)] for i in range(0, dp.shape[1], 8)])
    save("distribution.npz", pts=dp, distr_m1=dfm, distr_eq=deq, m1_eq=meq, omega=enu,
         total_nuf=tnuf, total_nuf_points=np.arange(0, dp.shape[1], 8))

    # --- gray M1 opacities end to end ---------------------------------------
    cfg = W.WORKLOADS["pts4096_n10_brt06"]
    p0 = W.make_points(cfg["npts"], seed=cfg["seed"])
    out = O.m1_opacities(p0, 10, W.FLAGS_ALL, W.params_default_all(W.BREM_BRT06), kind="ref")
    save("m1_pts4096_n10_brt06.npz", pts=p0, out=out, nquad=10, flags=W.FLAGS_ALL,
         params=W.params_default_all(W.BREM_BRT06), stim=1)

    sub = np.ascontiguousarray(p0[:, :512])
    out = O.m1_opacities(sub, 10, W.FLAGS_ALL, W.params_default_all(W.BREM_HR98), kind="ref")
    save("m1_pts512_n10_hr98.npz", pts=sub, out=out, nquad=10, flags=W.FLAGS_ALL,
         params=W.params_default_all(W.BREM_HR98), stim=1)

    ep = edge_points()
    variants = {}
    cases = [
        ("brt06_n10", 10, W.FLAGS_ALL, W.params_default_all(1), 1),
        ("hr98_blocking_n8", 8, W.FLAGS_ALL, np.array([1, 1, 1, 1, 1, 0, 1, 0], np.int32), 1),
        ("hr98_nomed_n6", 6, W.FLAGS_ALL, np.array([0, 1, 0, 1, 0, 0, 0, 0], np.int32), 1),
        ("nostim_n10", 10, W.FLAGS_ALL, np.array([1, 1, 1, 1, 1, 1, 1, 0], np.int32), 0),
        ("gp19_n4", 4, W.FLAGS_ALL, W.params_default_all(2), 1),
    ]
    for i in range(5):
        f = np.zeros(5, np.int32)
        f[i] = 1
        cases.append((f"only{i}_n10", 10, f, np.array([1, 1, 1, 1, 1, 1, 1, 0], np.int32), 1))
    mix = np.concatenate([ep, np.ascontiguousarray(p0[:, 512:640])], axis=1)
    for tag, nq, fl, prm, stim in cases:
        variants[f"out_{tag}"] = O.m1_opacities(mix, nq, fl, prm, stim=stim, kind="ref")
        variants[f"nquad_{tag}"] = np.int32(nq)
        variants[f"flags_{tag}"] = fl
        variants[f"params_{tag}"] = prm
        variants[f"stim_{tag}"] = np.int32(stim)
    deq = O.distr_equilibrium(mix, kind="ref")
    meq = O.m1_densities_eq(mix, deq, kind="ref")
    prm = np.array([0, 0, 1, 1, 1, 1, 1, 0], np.int32)
    variants["out_equilibrium_n10"] = O.m1_opacities(meq, 10, W.FLAGS_ALL, prm, distr=deq,
                                                      kind="ref")
    save("m1_variants.npz", pts=mix, pts_eq=meq, distr_eq=deq, eq_params=prm, **variants)

    spectral_fixtures()


def _bessk1_ref(x):
    lib, _ = O.load("ref")
    out = np.empty_like(x)
    lib.ref_bessk1(x.size, np.ascontiguousarray(x), out)
    return out


if __name__ == "__main__":
    main(sys.argv[1:])
