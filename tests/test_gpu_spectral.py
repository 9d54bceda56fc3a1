"""Parity of the HIP spectral opacity path (ComputeSpectralOpacities*,
m1_opacities.hpp:1316-1660) with the reference / the CPU oracle.

Every test calls the product through its C ABI (ctypes over libbnsamd.so)
and compares with the reference's own outputs stored in tests/golden (the six
points of the reference's spectral test; seeded and edge points under every
reaction / parameter variant, n = 4..20) or with the CPU oracle, which is
bit-exact with the reference, beyond them (n = 16 and 32 for every
variant).  FP64 bar per output: |gpu - ref| <= 1e-12 |ref| + 2 u m, m the
running error magnitude of that output (oracle/error_bound.cpp, u = 2^-53):
the plain 1e-12 relative bar where the output is well conditioned, the bound
every double evaluation obeys where it is formed by cancellation.
"""
import numpy as np
import pytest

import oracle as O
from bns_nurates_amd import workloads as W
from tests.gpu_util import check_parity, distinct_heavy_species, say

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

SP_NAMES = [f"{q}[{s}]" for q in ("j", "kappa", "j_s", "kappa_s") for s in range(4)]


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


def run(ctx, pts, nu, nquad, flags, params, stim=0, distr=None):
    ctx.set_quadrature(nquad)
    out = ctx.compute_spectral_opacities(dev(pts), dev(nu), flags=flags, params=params,
                                         stim_abs=stim,
                                         distr=None if distr is None else dev(distr))
    torch.cuda.synchronize()
    return out.cpu().numpy()[:, :, :pts.shape[1]]


def parity(tag, got, ref, mag):
    """check_parity over [16, nnu * npts] so that the worst output is named"""
    f = lambda a: np.asarray(a).reshape(16, -1)  # noqa: E731
    return check_parity(tag, f(got), f(ref), f(mag), SP_NAMES)


def test_reference_spectral_test_points(ctx, golden):
    """tests/tests.hpp:861-1261: six points x 100 energies, equilibrium
    distributions, all reactions, HR98, no corrections, n = 10."""
    g = golden("spectral_test6_n10_hr98.npz")
    say("reference spectral test: 6 points x 100 energies, n=10, HR98")
    args = (int(g["nquad"]), g["flags"], g["params"])
    out = run(ctx, g["pts"], g["nu"], *args, stim=int(g["stim"]), distr=g["distr"])
    _, m = O.bound_spectral_opacities(g["pts"], g["nu"], *args, stim=int(g["stim"]),
                                      distr=g["distr"])
    assert np.all(np.isfinite(out))
    assert parity("spectral test6", out, g["out"], m) <= 1.0


def test_variants_vs_reference(ctx, golden):
    """Seeded and edge points (cold / hot, degenerate, pair cut-off, NES Taylor
    branches, tiny and huge energies) under every single reaction, both
    bremsstrahlung models and GP19, medium correction, blocking, stimulated
    absorption, n = 4, 6, 10, 16, 20."""
    g = golden("spectral_variants.npz")
    tags = [k[4:] for k in g.files if k.startswith("out_") and k != "out_perpoint_n10_stim"]
    worst = 0.0
    for t in tags:
        say(f"variant {t}")
        args = (int(g[f"nquad_{t}"]), g[f"flags_{t}"], g[f"params_{t}"])
        stim = int(g[f"stim_{t}"])
        out = run(ctx, g["pts"], g["nu"], *args, stim=stim)
        _, m = O.bound_spectral_opacities(g["pts"], g["nu"], *args, stim=stim)
        worst = max(worst, parity(t, out, g[f"out_{t}"], m))
    say("variant perpoint_n10_stim (energies per point)")
    prm = W.params_default_all(W.BREM_BRT06)
    out = run(ctx, g["pts"], g["nu_perpoint"], 10, W.FLAGS_ALL, prm, stim=1)
    _, m = O.bound_spectral_opacities(g["pts"], g["nu_perpoint"], 10, W.FLAGS_ALL, prm, stim=1)
    worst = max(worst, parity("perpoint", out, g["out_perpoint_n10_stim"], m))
    assert worst <= 1.0


@pytest.mark.parametrize("nquad", [16, 32])
def test_large_orders_vs_oracle(ctx, nquad):
    """Orders past the reference's BS_N_MAX (n <= 20 for the spectral
    integrals): oracle only, both bremsstrahlung models, blocking on."""
    pts = W.make_points(96, seed=7 + nquad)
    nu = np.exp(np.linspace(np.log(0.1), np.log(800.0), 13))
    worst = 0.0
    for brem in (W.BREM_BRT06, W.BREM_HR98):
        prm = W.params_default_all(brem)
        prm[7] = 0
        say(f"n={nquad} brem={brem}: 96 points x 13 energies")
        out = run(ctx, pts, nu, nquad, W.FLAGS_ALL, prm, stim=1)
        ref, m = O.bound_spectral_opacities(pts, nu, nquad, W.FLAGS_ALL, prm, stim=1)
        worst = max(worst, parity(f"n{nquad} brem{brem}", out, ref, m))
    assert worst <= 1.0


def test_distinct_heavy_species_vs_oracle(ctx):
    """anux with its own moments (the kernel copies the nux distribution
    values when the two agree)."""
    pts = distinct_heavy_species(W.make_points(96, seed=15), seed=16)
    nu = np.exp(np.linspace(np.log(0.3), np.log(300.0), 9))
    prm = W.params_default_all(W.BREM_HR98)
    prm[7] = 0
    say("spectral, anux moments distinct from nux: 96 points x 9 energies, n=10")
    out = run(ctx, pts, nu, 10, W.FLAGS_ALL, prm, stim=1)
    ref, m = O.bound_spectral_opacities(pts, nu, 10, W.FLAGS_ALL, prm, stim=1)
    assert parity("distinct anux", out, ref, m) <= 1.0


def test_host_entry_matches_device(ctx):
    pts = W.make_points(40, seed=3)
    nu = np.exp(np.linspace(np.log(0.3), np.log(300.0), 11))
    prm = W.params_default_all(W.BREM_HR98)
    ctx.set_quadrature(10)
    h = ctx.compute_spectral_opacities(pts, nu, W.FLAGS_ALL, prm, stim_abs=1)
    d = run(ctx, pts, nu, 10, W.FLAGS_ALL, prm, stim=1)
    assert np.array_equal(h, d)
    nu2 = np.ascontiguousarray(nu[:, None] * (1.0 + 0.01 * np.arange(40))[None, :])
    h2 = ctx.compute_spectral_opacities(pts, nu2, W.FLAGS_ALL, prm, stim_abs=1)
    d2 = run(ctx, pts, nu2, 10, W.FLAGS_ALL, prm, stim=1)
    assert np.array_equal(h2, d2)


def test_batch_and_chunk_invariance(ctx):
    """A point's spectrum does not depend on the other points of the call or
    on how its energies fall into workgroup chunks (bit for bit)."""
    pts = W.make_points(50, seed=11)
    nu = np.exp(np.linspace(np.log(0.2), np.log(500.0), 37))
    prm = W.params_default_all(W.BREM_BRT06)
    full = run(ctx, pts, nu, 10, W.FLAGS_ALL, prm, stim=1)
    part = run(ctx, pts[:, 17:29], nu[5:31], 10, W.FLAGS_ALL, prm, stim=1)
    assert np.array_equal(full[:, 5:31, 17:29], part)
    one = run(ctx, pts[:, 3:4], nu[36:37], 10, W.FLAGS_ALL, prm, stim=1)
    assert np.array_equal(full[:, 36:37, 3:4], one)


def test_grid_scale_properties(ctx):
    """32^3 points of the grid distributions x 16 energies, n = 10: finite
    everywhere, non-negative emissivities and opacities where the oracle's
    are, the stimulated form equal to kappa + j / c of the plain form, and an
    oracle sample."""
    npts = 32 ** 3
    pts = W.make_points(npts, seed=2024)
    nu = np.exp(np.linspace(np.log(1.0), np.log(300.0), 16))
    prm = W.params_default_all(W.BREM_BRT06)
    say(f"grid: {npts} points x 16 energies")
    o0 = run(ctx, pts, nu, 10, W.FLAGS_ALL, prm, stim=0)
    o1 = run(ctx, pts, nu, 10, W.FLAGS_ALL, prm, stim=1)
    assert np.all(np.isfinite(o0)) and np.all(np.isfinite(o1))
    c = 2.99792458e+17
    assert np.array_equal(o0[:4], o1[:4]) and np.array_equal(o0[8:12], o1[8:12])
    np.testing.assert_allclose(o1[4:8], o0[4:8] + o0[:4] / c, rtol=1e-15, atol=0)
    idx = np.arange(0, npts, npts // 256)
    sub = np.ascontiguousarray(pts[:, idx])
    ref, m = O.bound_spectral_opacities(sub, nu, 10, W.FLAGS_ALL, prm, stim=0)
    assert parity("grid sample", o0[:, :, idx], ref, m) <= 1.0


@pytest.mark.parametrize("nquad", [10, 7])
def test_struct_layout_matches_field_layout(ctx, nquad):
    """bnsamd_spectral_opacities_aos writes the values of
    bnsamd_spectral_opacities as SpectralOpacities records [npts][nnu][16],
    bit for bit: a ragged last chunk (nnu = 37), shared and per-point
    energies, both stimulated forms, the fixed n = 10 instance and another
    order."""
    pts = W.make_points(70, seed=5)
    nu = np.exp(np.linspace(np.log(0.2), np.log(400.0), 37))
    prm = W.params_default_all(W.BREM_BRT06)
    ctx.set_quadrature(nquad)
    dpts = dev(pts)
    for stim in (0, 1):
        for nuv in (nu, nu[:, None] * (1.0 + 0.003 * np.arange(70))[None, :]):
            say(f"struct layout: n={nquad} stim={stim} nu{nuv.shape}")
            dnu = dev(nuv)
            f = ctx.compute_spectral_opacities(dpts, dnu, W.FLAGS_ALL, prm, stim_abs=stim)
            s = ctx.compute_spectral_opacities(dpts, dnu, W.FLAGS_ALL, prm, stim_abs=stim,
                                               layout="struct")
            torch.cuda.synchronize()
            assert tuple(s.shape) == (70, 37, 16)
            assert np.array_equal(f.cpu().numpy(), s.cpu().numpy().transpose(2, 1, 0))
