"""The oracle's GP19 bremsstrahlung (oracle/m1_oracle.hpp, restating
kernel_brem_GP19.hpp:35-480) over synthetic tables (tests/gp19_synth.py):
the reference's table blob is absent, so the physics values are parity
unpinned, and the interpolation and extrapolations are pinned by their
mathematics:

* the interpolation returns the table at its nodes and reproduces any
  function multilinear in (log nb, Ye, T, log w); NaN outside the axes;
* the kernel is 3 |S| times the density factor, with emission by detailed
  balance at the unclipped pair energy; Ye and nb above the table are
  clipped;
* every extrapolation meets the in-range kernel at the table's edge
  (density, low and high temperature);
* in the gray path, GP19 obeys Kirchhoff's law at equilibrium and adds to
  the other reactions; without a table it adds zero.
"""
import numpy as np
import pytest

import oracle as O
from bns_nurates_amd import workloads as W
from tests import gp19_synth as G



@pytest.fixture()
def smooth():
    O.set_gp19_table(G.table("smooth"))
    yield G.table("smooth")
    O.set_gp19_table(None)


def test_interpolation_at_nodes_and_on_multilinear_data():
    t = G.table("multilinear")
    O.set_gp19_table(t)
    try:
        rng = np.random.default_rng(0)
        i, j, k, l = (rng.integers(0, n, 200) for n in t["data"].shape)
        got = O.gp19_interp(t["nb"][i], t["ye"][j], t["T"][k], t["w"][l])
        assert np.array_equal(got, t["data"][i, j, k, l])
        nb = np.exp(rng.uniform(t["lognb"][0], t["lognb"][-1], 2000))
        ye = rng.uniform(t["ye"][0], t["ye"][-1], 2000)
        T = rng.uniform(t["T"][0], t["T"][-1], 2000)
        w = np.exp(rng.uniform(t["logw"][0], t["logw"][-1], 2000))
        got = O.gp19_interp(nb, ye, T, w)
        ref = G.multilinear_value(np.log(nb), ye, T, np.log(w))
        assert np.max(np.abs(got - ref) / np.abs(ref)) < 1e-13
        # outside any axis: NaN
        out = O.gp19_interp(np.array([1e-5, 0.5, 0.5, 0.5, 0.5]),
                            np.array([0.3, 0.001, 0.3, 0.3, 0.3]),
                            np.array([10., 10., 1.0, 60., 10.]),
                            np.array([1., 1., 1., 1., 500.]))
        assert np.isnan(out).all()
    finally:
        O.set_gp19_table(None)


def test_kernel_in_range_and_detailed_balance(smooth):
    t = smooth
    rng = np.random.default_rng(1)
    n = 500
    nb_fm = np.exp(rng.uniform(np.log(2e-4), np.log(0.5), n))
    T = rng.uniform(2.5, 45.0, n)
    yp = rng.uniform(0.02, 0.45, n)
    om, omp = rng.uniform(0.1, 100.0, n), rng.uniform(0.1, 100.0, n)
    em, ab = O.brem_gp19(om, omp, nb_fm * 1e18, T, yp)
    S = O.gp19_interp(nb_fm, yp, T, om + omp)
    brem_const = ab / (3.0 * nb_fm * 1e18 * S)
    assert np.allclose(brem_const, brem_const[0], rtol=1e-14, atol=0)
    assert np.allclose(em, ab * np.exp(-(om + omp) / T), rtol=1e-14, atol=0)
    # Ye and nb above the table are clipped to its edges; w to its range
    e1, a1 = O.brem_gp19(om[:5], omp[:5], np.full(5, 2.0e18), T[:5], np.full(5, 0.6))
    e2, a2 = O.brem_gp19(om[:5], omp[:5], np.full(5, 1.0e18), T[:5], np.full(5, 0.5))
    assert np.array_equal(a1, a2)
    e3, a3 = O.brem_gp19(np.full(5, 300.), np.full(5, 300.), nb_fm[:5] * 1e18, T[:5], yp[:5])
    e4, a4 = O.brem_gp19(np.full(5, 199.), np.full(5, 199.), nb_fm[:5] * 1e18, T[:5], yp[:5])
    assert np.array_equal(a3, a4)  # both above w_max = 398 MeV
    assert np.allclose(e3, a3 * np.exp(-600. / T[:5]), rtol=1e-14, atol=0)


@pytest.mark.parametrize("edge", ["low density", "low T", "high T", "low density, low T",
                                  "low density, high T"])
def test_extrapolations_meet_the_table_at_its_edges(smooth, edge):
    """Each extrapolation, evaluated just past the table's edge, equals the
    in-range kernel at the edge to first order (relative step 1e-9)."""
    t = smooth
    rng = np.random.default_rng(2)
    n = 200
    om, omp = rng.uniform(0.2, 60.0, n), rng.uniform(0.2, 60.0, n)
    yp = rng.uniform(0.02, 0.45, n)
    nb_in = np.exp(rng.uniform(np.log(2e-4), np.log(0.5), n))
    T_in = rng.uniform(2.5, 45.0, n)
    eps = 1e-9
    nb_edge, T_edge = nb_in.copy(), T_in.copy()
    nb_out, T_out = nb_in.copy(), T_in.copy()
    if "low density" in edge:
        nb_edge[:] = t["nb"][0]
        nb_out[:] = t["nb"][0] * (1 - eps)
    if "low T" in edge:
        T_edge[:] = t["T"][0]
        T_out[:] = t["T"][0] * (1 - eps)
    if "high T" in edge:
        T_edge[:] = t["T"][-1]
        T_out[:] = t["T"][-1] * (1 + eps)
    _, a_edge = O.brem_gp19(om, omp, nb_edge * 1e18, T_edge, yp)
    _, a_out = O.brem_gp19(om, omp, nb_out * 1e18, T_out, yp)
    assert np.all(np.isfinite(a_out)) and np.all(a_out > 0)
    assert np.max(np.abs(a_out / a_edge - 1)) < 1e-6, edge


def test_gray_path_kirchhoff_and_additivity(smooth):
    """GP19 in ComputeM1Opacities: with equilibrium distributions and
    blocking on, eta = c kappa_a J holds with GP19 as the only reaction
    (its kernels obey detailed balance); all reactions = the sum of each."""
    pts = W.make_points(200, seed=5)
    eta_nu = np.abs(pts[W.P_MUE] - pts[W.P_MUN] + pts[W.P_MUP]) / pts[W.P_T]
    pts = np.ascontiguousarray(pts[:, eta_nu < 10.0])
    deq = O.distr_equilibrium(pts)
    meq = O.m1_densities_eq(pts, deq)
    prm = np.array([0, 0, 1, 1, 1, W.BREM_GP19, 1, 0], np.int32)
    only_brem = np.array([0, 0, 1, 0, 0], np.int32)
    out = O.m1_opacities(meq, 10, only_brem, prm, distr=deq)
    c = 2.99792458e+17
    for s in range(4):
        J = meq[W.P_J0 + s]
        ok = (J > 1e-25) & (out[8 + s] > 1e-250)
        assert ok.sum() > 10
        assert np.allclose(c * out[12 + s][ok] * J[ok], out[8 + s][ok], rtol=1e-9)
    tot = O.m1_opacities(pts, 10, W.FLAGS_ALL, prm)
    acc = np.zeros_like(tot)
    for i in range(5):
        f = np.zeros(5, np.int32)
        f[i] = 1
        acc += O.m1_opacities(pts, 10, f, prm)
    big = np.abs(tot) > 1e-200
    assert np.max(np.abs(acc - tot)[big] / np.abs(tot[big])) < 1e-10
    brem = O.m1_opacities(pts, 10, only_brem, prm)
    assert np.abs(brem).max() > 0


def test_without_a_table_gp19_adds_zero():
    O.set_gp19_table(None)
    pts = W.make_points(16, seed=7)
    prm = W.params_default_all(W.BREM_GP19)
    assert np.array_equal(O.m1_opacities(pts, 6, np.array([0, 0, 1, 0, 0], np.int32), prm),
                          np.zeros((20, 16)))
