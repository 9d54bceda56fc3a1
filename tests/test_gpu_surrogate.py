"""The learned path end to end: an ensemble trained on the GPU by the
library's own kernels on FP64-path data, the fused BF16-MFMA inference
kernel, and the per-point fallback to the FP64 path judged against it.

* points the ensemble does not trust (spread above the threshold, or
  features outside the training box) carry the FP64 path's bits exactly;
* accepted points stay within the stated tolerance of the FP64 path;
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
    rel = np.abs(o[rows] - r[rows]) / np.abs(r[rows])
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
    err = (np.abs(np.log(o[use]) - mean[use]) / ts).max()
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
