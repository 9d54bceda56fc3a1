"""Parity of the HIP gray M1 path with the reference / the CPU oracle.

Every test calls the product through its C ABI (bns_nurates_amd.api, ctypes
over libbnsamd.so) on a MI355X and compares with:
* the reference's own outputs stored in tests/golden (n <= 10), and
* the CPU oracle (oracle/, bit-exact with the reference) on the same seeded
  inputs, which also covers n = 16 where the reference cannot run.
FP64 bar, per output: |gpu - ref| <= 1e-12 |ref| + 2 u m, where m is the
running error magnitude of that output (oracle/error_bound.cpp, u = 2^-53):
to first order each double evaluation lies within u m of the exact value of
the formulas, so two of them differ by at most 2 u m.
For well-conditioned outputs this is the plain 1e-12 relative bar; outputs
formed by cancellation (pair and NES rates at large degeneracy, where the
reference itself yields rounding-noise and even negative rates, see
kernel_pair.hpp:35-41) are held to the bound any double evaluation obeys.
"""
import numpy as np
import pytest

import oracle as O
from bns_nurates_amd import workloads as W
from tests.gpu_util import check_parity, distinct_heavy_species, rel_err, report, say

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL64 = 1e-12


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from bns_nurates_amd import api
    c = api.Context(0)
    yield c
    c.close()


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def run(ctx, pts, nquad, flags, params, stim=1, distr=None, precision=0, n1=None):
    ctx.set_quadrature(nquad, n1=n1)
    out = ctx.compute_m1_opacities(dev(pts), flags=flags, params=params, stim_abs=stim,
                                   precision=precision,
                                   distr=None if distr is None else dev(distr))
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_fdi_device_vs_reference(ctx, golden):
    g = golden("fdi.npz")
    y = dev(g["y"])
    for k in range(6):
        say(f"fdi k={k} n={y.numel()}")
        got = ctx.fdi(k, y)
        torch.cuda.synchronize()
        r = report(f"F{k}", got.cpu().numpy(), g[f"F{k}"])
        assert r.max() <= 1e-13


def test_distribution_device_vs_reference(ctx, golden):
    g = golden("distribution.npz")
    say("distr_from_m1 / equilibrium", g["pts"].shape)
    d = ctx.distr_from_m1(dev(g["pts"]))
    torch.cuda.synchronize()
    # eta_t = log(...) - 30-type reconstructions cancel: condition-aware bar
    _, m = O.bound_distr_from_m1(g["pts"])
    names = [f"{f}[{s}]" for f in ("w_t", "temp_t", "eta_t", "w_f", "temp_f", "c_f", "beta_f")
             for s in range(4)]
    assert check_parity("distr_from_m1", d.cpu().numpy(), g["distr_m1"], m, names) <= 1.0
    p = dev(g["pts"])
    deq = ctx.equilibrium(p)
    torch.cuda.synchronize()
    assert rel_err(deq.cpu().numpy(), g["distr_eq"]).max() <= 1e-15
    r = report("m1_densities_eq", p.cpu().numpy(), g["m1_eq"])
    assert r.max() <= 1e-13


def test_pts4096_n10_brt06_vs_reference(ctx, golden):
    g = golden("m1_pts4096_n10_brt06.npz")
    say("4096 points, n=10, BRT06 vs reference outputs")
    out = run(ctx, g["pts"], 10, g["flags"], g["params"], int(g["stim"]))
    v, m = O.bound_m1_opacities(g["pts"], 10, g["flags"], g["params"], int(g["stim"]))
    # the host's oracle reproduces the reference (bit for bit where its libm
    # rounds as the one the fixtures were made with)
    assert O.parity_excess(v, g["out"], m).max() <= 1.0
    assert check_parity("4096 pts brt06", out, g["out"], m, W.OUT_NAMES) <= 1.0


def test_pts512_n10_hr98_vs_reference(ctx, golden):
    g = golden("m1_pts512_n10_hr98.npz")
    say("512-point subset, n=10, HR98 vs reference outputs")
    out = run(ctx, g["pts"], 10, g["flags"], g["params"], int(g["stim"]))
    _, m = O.bound_m1_opacities(g["pts"], 10, g["flags"], g["params"], int(g["stim"]))
    assert check_parity("512 pts hr98", out, g["out"], m, W.OUT_NAMES) <= 1.0


def test_variants_vs_reference(ctx, golden):
    """Edge points (thresholds, asymptotic s_beta, pair cutoff, NES Taylor
    branches, n = 0 guards, pure thick / thin) under every flag set, both
    bremsstrahlung models, blocking, non-stimulated absorption."""
    g = golden("m1_variants.npz")
    tags = [k[4:] for k in g.files if k.startswith("out_") and k != "out_equilibrium_n10"]
    worst = 0.0
    for t in tags:
        say(f"variant {t}")
        args = (int(g[f"nquad_{t}"]), g[f"flags_{t}"], g[f"params_{t}"], int(g[f"stim_{t}"]))
        out = run(ctx, g["pts"], *args)
        _, m = O.bound_m1_opacities(g["pts"], *args)
        worst = max(worst, check_parity(t, out, g[f"out_{t}"], m, W.OUT_NAMES))
    say("variant equilibrium_n10 (distribution block given)")
    out = run(ctx, g["pts_eq"], 10, W.FLAGS_ALL, g["eq_params"], distr=g["distr_eq"])
    _, m = O.bound_m1_opacities(g["pts_eq"], 10, W.FLAGS_ALL, g["eq_params"], distr=g["distr_eq"])
    worst = max(worst, check_parity("equilibrium", out, g["out_equilibrium_n10"], m, W.OUT_NAMES))
    assert worst <= 1.0


@pytest.mark.parametrize("brem", [W.BREM_BRT06, W.BREM_HR98])
def test_n16_vs_oracle(ctx, brem):
    """The 128^3-grid distributions at n = 16 (beyond the reference's BS_N_MAX):
    against the n-generic oracle on a seeded sample."""
    pts = W.make_points(768, seed=1)
    prm = W.params_default_all(brem)
    say(f"n=16 brem={brem}: 768 points vs oracle")
    ref, m = O.bound_m1_opacities(pts, 16, W.FLAGS_ALL, prm)
    out = run(ctx, pts, 16, W.FLAGS_ALL, prm)
    assert check_parity(f"n16 brem{brem}", out, ref, m, W.OUT_NAMES) <= 1.0


def test_distinct_heavy_species_vs_oracle(ctx):
    pts = distinct_heavy_species(W.make_points(384, seed=5), seed=6)
    prm = W.params_default_all(W.BREM_BRT06)
    prm[7] = 0  # blocking on: the distribution enters every reaction
    say("anux moments distinct from nux: 384 points, n=16, FP64 and mixed")
    ref, m = O.bound_m1_opacities(pts, 16, W.FLAGS_ALL, prm)
    out = run(ctx, pts, 16, W.FLAGS_ALL, prm)
    assert check_parity("distinct anux", out, ref, m, W.OUT_NAMES) <= 1.0
    mixed = run(ctx, pts, 16, W.FLAGS_ALL, prm, precision=1)
    assert check_parity("distinct anux, mixed", mixed, ref, m, W.OUT_NAMES, rtol=1e-5) <= 1.0
    # the nux-shared path on the same points gives the nux / anux outputs the
    # oracle gives when anux copies nux
    same = np.array(pts, copy=True)
    for f in (W.P_N0, W.P_J0, W.P_CHI0):
        same[f + 3] = same[f + 2]
    ref2, m2 = O.bound_m1_opacities(same, 16, W.FLAGS_ALL, prm)
    assert check_parity("anux = nux", run(ctx, same, 16, W.FLAGS_ALL, prm), ref2, m2,
                        W.OUT_NAMES) <= 1.0


def test_mixed_orders_and_blocking_vs_oracle(ctx):
    pts = W.make_points(256, seed=21)
    for n1, n2, stim, blk in ((4, 12, 1, 0), (12, 5, 0, 1), (24, 24, 1, 0), (1, 2, 1, 1)):
        prm = W.params_default_all(W.BREM_HR98)
        prm[7] = blk
        say(f"n1={n1} n2={n2} stim={stim} neglect_blocking={blk}")
        ref, m = O.bound_m1_opacities(pts, n2, W.FLAGS_ALL, prm, stim=stim, nquad_1d=n1)
        out = run(ctx, pts, n2, W.FLAGS_ALL, prm, stim=stim, n1=n1)
        assert check_parity(f"n1={n1} n2={n2}", out, ref, m, W.OUT_NAMES) <= 1.0


@pytest.mark.parametrize("brem,stim,blk,prec", [(W.BREM_BRT06, 1, 1, 0), (W.BREM_HR98, 0, 0, 0),
                                                (W.BREM_BRT06, 1, 0, 1)])
def test_max_order_vs_oracle(ctx, brem, stim, blk, prec):
    """The largest quadrature the API accepts (n = 32: 64 nodes, one per
    lane of a wavefront) against the oracle; FP64 to the FP64 bar, the mixed
    form (precision 1) to 1e-5."""
    pts = W.make_points(96, seed=33)
    prm = W.params_default_all(brem)
    prm[7] = blk
    say(f"n=32 brem={brem} stim={stim} neglect_blocking={blk} precision={prec}")
    ref, m = O.bound_m1_opacities(pts, 32, W.FLAGS_ALL, prm, stim=stim)
    out = run(ctx, pts, 32, W.FLAGS_ALL, prm, stim=stim, precision=prec)
    rtol = TOL64 if prec == 0 else 1e-5
    assert check_parity(f"n=32 precision={prec}", out, ref, m, W.OUT_NAMES, rtol=rtol) <= 1.0


def test_batch_invariance(ctx):
    """A point's result does not depend on the batch it is computed in."""
    pts = W.make_points(3000, seed=4)
    prm = W.params_default_all(W.BREM_BRT06)
    full = run(ctx, pts, 10, W.FLAGS_ALL, prm)
    idx = np.array([0, 7, 999, 1500, 2999])
    part = run(ctx, pts[:, idx], 10, W.FLAGS_ALL, prm)
    assert np.array_equal(full[:, idx], part)


def test_host_entry_point_matches_device(ctx):
    pts = W.make_points(300, seed=8)
    prm = W.params_default_all(W.BREM_BRT06)
    ctx.set_quadrature(8)
    a = ctx.compute_m1_opacities(pts, flags=W.FLAGS_ALL, params=prm)
    b = run(ctx, pts, 8, W.FLAGS_ALL, prm)
    assert np.array_equal(a, b)


def test_grid128_full_properties(ctx):
    """The 128^3-point grid at full size (n = 16): every output
    finite where the inputs are regular, kappa zero exactly where the
    thresholds say so, and a seeded sample equal to the oracle."""
    cfg = W.WORKLOADS["grid128_n16_brt06"]
    say("grid128: generating 2,097,152 points")
    pts = W.make_points(cfg["npts"], seed=cfg["seed"])
    prm = W.params_default_all(W.BREM_BRT06)
    say("grid128: running the FP64 kernel")
    out = run(ctx, pts, 16, W.FLAGS_ALL, prm)
    assert np.isfinite(out).all()
    say("grid128: threshold zeros")
    n = pts[W.P_N0:W.P_N0 + 4]
    J = pts[W.P_J0:W.P_J0 + 4]
    assert (out[4:8][n <= 1e-21] == 0).all()
    assert (out[12:20][np.concatenate([J, J]) <= 1e-25] == 0).all()
    rng = np.random.default_rng(3)
    sel = np.sort(rng.choice(cfg["npts"], 384, replace=False))
    say("grid128: oracle on a 384-point sample")
    ref, m = O.bound_m1_opacities(np.ascontiguousarray(pts[:, sel]), 16, W.FLAGS_ALL, prm)
    assert check_parity("grid128 sample", out[:, sel], ref, m, W.OUT_NAMES) <= 1.0
