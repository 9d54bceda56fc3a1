"""GP19 bremsstrahlung on the GPU (bns_nurates_amd/csrc/gp19.h, its own
kernel instances) against the oracle's restatement of
kernel_brem_GP19.hpp:35-480, both over the same synthetic table
(tests/gp19_synth.py): the reference's table blob is absent, so the
physics values are parity unpinned and the oracle is pinned by mathematics
(tests/test_oracle_gp19.py).  The points reach every regime: in range,
below the table's density and temperature, above its temperature, Ye and
density above its edges.  Bar: the FP64 one, through the C ABI."""
import numpy as np
import pytest

import oracle as O
from bns_nurates_amd import workloads as W
from tests import gp19_synth as G
from tests.gpu_util import check_parity, say

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from bns_nurates_amd import api
    c = api.Context(0)
    c.set_gp19_table(G.table("smooth"))
    O.set_gp19_table(G.table("smooth"))
    yield c
    O.set_gp19_table(None)
    c.close()


def _points(n, seed):
    pts = W.make_points(n, seed=seed)
    pts[W.P_T, : n // 8] = np.linspace(51.0, 80.0, n // 8)  # above the table's T
    return np.ascontiguousarray(pts)


@pytest.mark.parametrize("nq,stim,blocking", [(10, 1, 0), (10, 0, 1), (16, 1, 1)])
def test_gray_gp19_matches_oracle(ctx, nq, stim, blocking):
    pts = _points(512, seed=20 + nq)
    nbfm = pts[W.P_NB] * 1e-18
    say(f"GP19 gray n={nq}: {np.mean(nbfm < 1e-4):.0%} below the table's density, "
        f"{np.mean(pts[W.P_T] < 2):.0%} below 2 MeV, {np.mean(pts[W.P_T] > 50):.0%} above 50 MeV")
    prm = W.params_default_all(W.BREM_GP19)
    prm[7] = 1 - blocking
    ctx.set_quadrature(nq)
    out = ctx.compute_m1_opacities(torch.from_numpy(pts).cuda(), flags=W.FLAGS_ALL, params=prm,
                                   stim_abs=stim).cpu().numpy()
    ref, m = O.bound_m1_opacities(pts, nq, W.FLAGS_ALL, prm, stim=stim)
    assert np.isfinite(out).all()
    assert check_parity(f"gp19 n={nq} stim={stim} block={blocking}", out, ref, m,
                        W.OUT_NAMES) <= 1.0
    # GP19 alone contributes
    only = np.array([0, 0, 1, 0, 0], np.int32)
    b = ctx.compute_m1_opacities(torch.from_numpy(pts).cuda(), flags=only, params=prm,
                                 stim_abs=stim).cpu().numpy()
    rb, mb = O.bound_m1_opacities(pts, nq, only, prm, stim=stim)
    assert np.abs(b).max() > 0
    assert check_parity(f"gp19 alone n={nq}", b, rb, mb, W.OUT_NAMES) <= 1.0


def test_gray_gp19_mixed_precision(ctx):
    pts = _points(256, seed=31)
    prm = W.params_default_all(W.BREM_GP19)
    ctx.set_quadrature(16)
    from bns_nurates_amd import api
    out = ctx.compute_m1_opacities(torch.from_numpy(pts).cuda(), flags=W.FLAGS_ALL, params=prm,
                                   precision=api.MIXED).cpu().numpy()
    ref, m = O.bound_m1_opacities(pts, 16, W.FLAGS_ALL, prm)
    assert check_parity("gp19 mixed", out, ref, m, W.OUT_NAMES, rtol=1e-5) <= 1.0


def test_spectral_gp19_matches_oracle(ctx):
    pts = _points(64, seed=33)
    nu = np.logspace(-1, 2.5, 20)
    prm = W.params_default_all(W.BREM_GP19)
    ctx.set_quadrature(10)
    out = ctx.compute_spectral_opacities(torch.from_numpy(pts).cuda(), torch.from_numpy(nu).cuda(),
                                         W.FLAGS_ALL, prm, stim_abs=0).cpu().numpy()
    ref, m = O.bound_spectral_opacities(pts, nu, 10, W.FLAGS_ALL, prm, stim=0)
    assert np.isfinite(out).all()
    assert check_parity("gp19 spectral", out, ref, m) <= 1.0


def test_cleared_table_adds_zero(ctx):
    """Without a table GP19 adds zero (the reference's table-free
    configuration): the same bits as bremsstrahlung switched off."""
    pts = _points(128, seed=34)
    prm = W.params_default_all(W.BREM_GP19)
    no_brem = np.array(W.FLAGS_ALL)
    no_brem[2] = 0
    ctx.set_quadrature(10)
    d = torch.from_numpy(pts).cuda()
    ctx.set_gp19_table(None)
    try:
        a = ctx.compute_m1_opacities(d, flags=W.FLAGS_ALL, params=prm).cpu().numpy()
        b = ctx.compute_m1_opacities(d, flags=no_brem, params=prm).cpu().numpy()
    finally:
        ctx.set_gp19_table(G.table("smooth"))
    assert np.array_equal(a, b)
    c = ctx.compute_m1_opacities(d, flags=W.FLAGS_ALL, params=prm).cpu().numpy()
    assert not np.array_equal(a, c)
