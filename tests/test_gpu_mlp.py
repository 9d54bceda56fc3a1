"""The surrogate MLP's own GPU kernels (bns_nurates_amd/csrc/mlp.hip) against
the numpy restatement oracle/mlp_oracle.py: forward pass and gradients on
the matrix cores, and training that lowers the loss on data the FP64 path
generates."""
import numpy as np
import pytest

import oracle.mlp_oracle as M
from bns_nurates_amd import surrogate as S
from bns_nurates_amd import workloads as W
from tests.gpu_util import say

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ens():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    e = S.Ensemble(0, members=2, max_batch=4096)
    yield e
    e.close()


def _params(seed):
    p = S.init_params(seed)
    rng = np.random.default_rng(seed + 100)
    wo, bo, _ = M.offsets()
    for l in range(M.LAYERS):
        p[bo[l]:bo[l] + M.LAYER_N[l]] = rng.normal(0, 0.05, M.LAYER_N[l])
    return p


def test_forward_matches_oracle(ens):
    rng = np.random.default_rng(1)
    p = _params(2)
    ens.set_params(0, p)
    x = rng.normal(size=(1000, M.FEAT)).astype(np.float32)
    y = torch.empty((1000, 15), dtype=torch.float32, device="cuda")
    ens.predict(0, torch.from_numpy(x).cuda(), y)
    zs, _ = M.forward(p, x)
    ref = zs[-1][:, :15]
    err = np.abs(y.cpu().numpy() - ref).max() / np.abs(ref).max()
    say(f"MLP forward vs oracle: max abs err / max |y| = {err:.2e}")
    assert err < 2e-2


@pytest.mark.parametrize("missing", [0.0, 0.15])
def test_gradients_match_oracle(ens, missing):
    """Loss and gradients on the matrix cores against the oracle; with
    `missing` of the targets NaN (outputs the FP64 path sets to zero), which
    both leave out of the loss."""
    rng = np.random.default_rng(4)
    p = _params(5)
    ens.set_params(1, p)
    x = rng.normal(size=(512, M.FEAT)).astype(np.float32)
    y = rng.normal(size=(512, M.OUT)).astype(np.float32)
    y[rng.random(y.shape) < missing] = np.nan
    loss, g = ens.gradients(1, torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda())
    lo, go = M.gradients(p, x, y)
    wo, bo, _ = M.offsets()
    say(f"MLP loss gpu {loss:.6f} oracle {lo:.6f}")
    assert abs(loss - lo) <= 1e-3 * lo
    for l in range(M.LAYERS):
        for name, a, b in (("W", wo[l], wo[l] + M.LAYER_K[l] * M.LAYER_N[l]),
                           ("b", bo[l], bo[l] + M.LAYER_N[l])):
            gg, oo = g[a:b].astype(np.float64), go[a:b]
            rel = np.linalg.norm(gg - oo) / max(np.linalg.norm(oo), 1e-30)
            cos = gg @ oo / max(np.linalg.norm(gg) * np.linalg.norm(oo), 1e-30)
            say(f"  layer {l} {name}: |dg|/|g| = {rel:.2e}, cos = {cos:.8f}")
            assert rel < 2e-2 and cos > 0.9998, (l, name)


def test_training_lowers_the_loss(ens):
    """One member, data from the FP64 path (n = 8) on 32768 seeded points."""
    from bns_nurates_amd import api
    pts = W.make_points(32768, seed=31)
    ctx = api.Context(0)
    ctx.set_quadrature(8)
    out = ctx.compute_m1_opacities(pts, flags=W.FLAGS_ALL,
                                   params=W.params_default_all(W.BREM_BRT06))
    ctx.close()
    f = S.features(pts)
    t, ok = S.targets(out)
    ok &= np.all(np.isfinite(f), axis=1)
    f, t = f[ok], t[ok]
    x = ((f - f.mean(0)) / f.std(0)).astype(np.float32)
    y = ((t - np.nanmean(t, 0)) / np.nanstd(t, 0)).astype(np.float32)
    ens.set_params(0, S.init_params(7))
    perm = torch.from_numpy(np.random.default_rng(0).permutation(len(x)).astype(np.int32)).cuda()
    hist = ens.train(0, torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda(), perm, epochs=6,
                     batch=1024, lr0=2e-3, lr1=2e-4)
    say(f"training loss per epoch: {np.array2string(hist, precision=4)}")
    assert np.isfinite(hist).all() and hist[-1] < 0.2 * hist[0] and hist[-1] < 0.05


def test_ensemble_training_matches_single_member_training(ens):
    """bnsamd_mlp_train_ensemble trains every member at once, each on its own
    stream and workspace; member e must end where bnsamd_mlp_train with the
    same initialisation and row order ends, bit for bit (every reduction of
    the training step adds its partial sums in a fixed order), and a second
    ensemble run from the same start must repeat the first exactly."""
    rng = np.random.default_rng(5)
    n = 8192
    x = torch.from_numpy(rng.normal(size=(n, M.FEAT)).astype(np.float32)).cuda()
    y = torch.from_numpy(rng.normal(size=(n, M.OUT)).astype(np.float32)).cuda()
    inits = [_params(40 + e) for e in range(ens.members)]
    perms = np.stack([rng.permutation(n).astype(np.int32) for _ in range(ens.members)])
    for e in range(ens.members):
        ens.set_params(e, inits[e])
    hist = ens.train_all(x, y, torch.from_numpy(perms).cuda(), epochs=2, batch=1024, lr0=1e-3,
                         lr1=1e-4)
    together = [ens.get_params(e) for e in range(ens.members)]
    for e in range(ens.members):
        ens.set_params(e, inits[e])
        h1 = ens.train(e, x, y, torch.from_numpy(perms[e]).cuda(), epochs=2, batch=1024,
                       lr0=1e-3, lr1=1e-4)
        alone = ens.get_params(e)
        moved = np.abs(alone - inits[e]).max()
        err = np.abs(together[e] - alone).max()
        say(f"member {e}: losses together {hist[e]} alone {h1}; max |d w| {err:.2e} "
            f"of a step of {moved:.2e}")
        assert moved > 0
        assert np.array_equal(hist[e], h1)
        assert np.array_equal(together[e], alone)
    for e in range(ens.members):
        ens.set_params(e, inits[e])
    hist2 = ens.train_all(x, y, torch.from_numpy(perms).cuda(), epochs=2, batch=1024, lr0=1e-3,
                          lr1=1e-4)
    assert np.array_equal(hist, hist2)
    for e in range(ens.members):
        assert np.array_equal(ens.get_params(e), together[e])
