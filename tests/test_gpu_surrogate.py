"""The learned path end to end: an ensemble trained on the GPU by the
library's own kernels on FP64-path data, the fused BF16-MFMA inference
kernel, and the per-point fallback to the FP64 path judged against it.

* points the ensemble does not trust (spread above the threshold, or
  features outside the training box) carry the FP64 path's bits exactly;
* accepted points stay within the stated tolerance of the FP64 path, and
  keep the exact zeros the FP64 path gives below the n / J thresholds;
* the fused kernel agrees with the training-path forward pass of each
  member (ensemble mean);
* out-of-box inputs always fall back.
"""
import numpy as np
import pytest

from bns_nurates_amd import surrogate as S
from bns_nurates_amd import workloads as W
from tests.gpu_util import say

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

NQ = 8
THR = 0.1


@pytest.fixture(scope="module")
def trained():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from bns_nurates_amd import api
    ctx = api.Context(0)
    prm = W.params_default_all(W.BREM_BRT06)
    say("training a 4-member ensemble on 1048576 points (n=8), 8 epochs")
    ens = S.train_ensemble(ctx, 1 << 20, NQ, prm, members=4, epochs=8, batch=4096, seed=500,
                           log=say)
    yield ctx, ens, prm
    ens.close()
    ctx.close()


def test_fallback_and_accepted_points(trained):
    ctx, ens, prm = trained
    pts = W.make_points(65536, seed=77)
    d = torch.from_numpy(pts).cuda()
    ctx.set_quadrature(NQ)
    ref = ctx.compute_m1_opacities(d, flags=W.FLAGS_ALL, params=prm)
    out = torch.full((20, pts.shape[1]), float("nan"), dtype=torch.float64, device="cuda")
    nfb = ens.m1_opacities(ctx, d, out, THR, W.FLAGS_ALL, prm)
    torch.cuda.synchronize()
    o, r = out.cpu().numpy(), ref.cpu().numpy()
    rows = S.OUT_ROWS + [3, 7, 11, 15, 19]
    with np.errstate(divide="ignore", invalid="ignore"):
        rel = np.where(r[rows] != 0, np.abs(o[rows] - r[rows]) / np.abs(r[rows]),
                       np.where(o[rows] == 0, 0.0, np.inf))
    same = np.all(o == r, axis=0)
    acc = ~same
    say(f"surrogate: {nfb} of {pts.shape[1]} points fell back ({nfb / pts.shape[1]:.1%}); "
        f"accepted: max rel err {rel[:, acc].max() if acc.any() else 0:.3e}, "
        f"median {np.median(rel[:, acc]) if acc.any() else 0:.3e}")
    assert np.isfinite(o).all()
    assert same.sum() >= nfb           # every fallback point carries the FP64 bits
    # accepted points exist and are, typically, within the threshold (the
    # spread is a proxy: single accepted outputs can be off by more)
    assert acc.sum() > 0
    assert np.median(rel[:, acc]) < THR
    # the threshold zeros of the opacities (n <= 1e-21, J <= 1e-25)
    zero = r == 0
    say(f"surrogate: {int(zero[:, acc].sum())} exact-zero outputs on accepted points")
    assert np.array_equal(o[zero], r[zero])


def test_threshold_zeros_on_accepted_points(trained):
    """Accepted points below the reference's n / J thresholds
    (m1_opacities.hpp:1199-1292) get exact zeros for kappa_0_a / kappa_a /
    kappa_s, as the FP64 path gives.  The trained box excludes such points
    (their targets have no logarithm), so the box and the residual range are
    opened here: every point is then accepted and the zeros must come from
    the kernel's own threshold rule."""
    ctx, ens, prm = trained
    pts = W.make_points(4096, seed=80)
    rng = np.random.default_rng(3)
    for s in range(3):
        for fld, thr in ((W.P_N0 + s, 1e-21), (W.P_J0 + s, 1e-25)):
            idx = rng.choice(pts.shape[1], 64, replace=False)
            pts[fld, idx] = thr * rng.choice([0.0, 0.5, 1.0], 64)
            if s == 2:
                pts[fld + 1, idx] = pts[fld, idx]  # anux copies nux
    nm = ens.norm
    keep = {k: np.array(getattr(nm, k)[:]) for k in
            ("feat_mean", "feat_std", "out_mean", "out_std", "box_lo", "box_hi", "res_lo", "res_hi")}
    lin = np.array([[nm.lin[i][o] for o in range(S.OUT)] for i in range(S.FEAT + 1)])
    big = np.full(S.FEAT, 1e30)
    wide = np.full(S.OUT, 1e30)
    d = torch.from_numpy(pts).cuda()
    ctx.set_quadrature(NQ)
    ref = ctx.compute_m1_opacities(d, flags=W.FLAGS_ALL, params=prm).cpu().numpy()
    out = torch.full((20, pts.shape[1]), float("nan"), dtype=torch.float64, device="cuda")
    try:
        ens.set_norm(keep["feat_mean"], keep["feat_std"], keep["out_mean"], keep["out_std"],
                     -big, big, lin, -wide, wide)
        nfb = ens.m1_opacities(ctx, d, out, 1e30, W.FLAGS_ALL, prm)
        torch.cuda.synchronize()
    finally:
        ens.set_norm(keep["feat_mean"], keep["feat_std"], keep["out_mean"], keep["out_std"],
                     keep["box_lo"], keep["box_hi"], lin, keep["res_lo"], keep["res_hi"])
    o = out.cpu().numpy()
    zero = ref == 0
    say(f"threshold zeros: {int(zero.sum())} zero outputs, {nfb} fallback points")
    # log(0) features fall outside any box; the points with n or J exactly 0
    # fall back, the rest are the network's with the kernel's zeros
    assert nfb < pts.shape[1] // 2
    assert zero.sum() > 0
    assert np.array_equal(o[zero], ref[zero])


def test_fused_kernel_matches_training_path(trained):
    ctx, ens, prm = trained
    pts = W.make_points(4096, seed=78)
    with np.errstate(divide="ignore"):
        f = S.features(pts)
    nm = ens.norm
    fm, fs = np.array(nm.feat_mean), np.array(nm.feat_std)
    tm, ts = np.array(nm.out_mean), np.array(nm.out_std)
    x = torch.from_numpy(((f - fm) / fs).astype(np.float32)).cuda()
    ys = []
    for e in range(ens.members):
        y = torch.empty((pts.shape[1], 15), dtype=torch.float32, device="cuda")
        ens.predict(e, x, y)
        ys.append(y.cpu().numpy())
    lin = np.c_[np.ones(len(f)), (f - fm) / fs] @ ens.lin
    mean = np.mean(ys, axis=0) * ts + tm + lin
    d = torch.from_numpy(pts).cuda()
    out = torch.zeros((20, pts.shape[1]), dtype=torch.float64, device="cuda")
    ctx.set_quadrature(NQ)
    ens.m1_opacities(ctx, d, out, 1e30, W.FLAGS_ALL, prm)  # no spread fallback
    torch.cuda.synchronize()
    o = out.cpu().numpy()[S.OUT_ROWS].T
    ref = ctx.compute_m1_opacities(d, flags=W.FLAGS_ALL, params=prm).cpu().numpy()
    fell = np.all(out.cpu().numpy() == ref, axis=0)  # recomputed by the FP64 path
    inbox = np.all((f >= np.array(nm.box_lo)) & (f <= np.array(nm.box_hi)), axis=1)
    use = inbox & ~fell
    # in the network's own (standardised) units: BF16 rounding noise only
    # (outputs set to zero by the n / J thresholds are not the network's)
    pos = o[use] > 0
    with np.errstate(divide="ignore"):
        err = np.where(pos, np.abs(np.log(o[use]) - mean[use]) / ts, 0.0).max()
    say(f"fused inference vs training-path forward: max |d y| = {err:.3e} on {use.sum()} points")
    assert use.sum() > 0 and err < 5e-2


def test_out_of_box_points_fall_back(trained):
    ctx, ens, prm = trained
    pts = W.make_points(256, seed=79)
    pts[W.P_T, :64] = 500.0  # far above the training range of T
    d = torch.from_numpy(pts).cuda()
    ctx.set_quadrature(NQ)
    ref = ctx.compute_m1_opacities(d, flags=W.FLAGS_ALL, params=prm).cpu().numpy()
    out = torch.zeros((20, 256), dtype=torch.float64, device="cuda")
    nfb = ens.m1_opacities(ctx, d, out, 1e30, W.FLAGS_ALL, prm)
    torch.cuda.synchronize()
    assert nfb >= 64
    assert np.array_equal(out.cpu().numpy()[:, :64], ref[:, :64])
