"""The learned form's calibration (bns_nurates_amd/surrogate.py calibrate)
on synthetic predictions, CPU only: a stand-in ensemble records the
acceptance coefficients.  Errors grow with the spread, and a minority of
points are confidently wrong (small spread, large error) as the network's
are near the reference's discontinuities."""
import numpy as np
import pytest

from bns_nurates_amd import surrogate as S


class _Ens:
    def set_acceptance(self, a, k):
        self.a, self.k = np.asarray(a, np.float64), np.asarray(k, np.float64)

    def set_judge(self, judge, threshold=0.0):
        self.judge = judge


def _data(n, seed, wrong=2e-4):
    rng = np.random.default_rng(seed)
    logref = rng.normal(0, 3, (S.OUT, n))
    spread = np.exp(rng.normal(np.log(0.001), 0.7, (S.OUT, n)))
    err = spread * rng.normal(0, 1.0, (S.OUT, n))
    bad = rng.random(n) < wrong
    err[:, bad] += 0.5 * spread[:, bad] / spread[:, bad].mean() * rng.choice([-1, 1], (S.OUT, bad.sum()))
    mean = logref + err
    ref = np.zeros((20, n))
    ref[S.OUT_ROWS] = np.exp(logref)
    noise = np.zeros((20, n))
    outside = (rng.random(n) < 0.05).astype(np.int32)
    return mean, spread.astype(np.float32), outside, ref, noise


@pytest.mark.parametrize("tol", [1e-2, 3e-2])
def test_rule_holds_the_bar_on_held_out_data(tol):
    ens = _Ens()
    rep = S.calibrate(ens, None, tol, None, 16, data=_data(1 << 17, 1), coverage=0.999)
    assert rep["split_A"]["accepted"] >= 3000
    mean, spread, outside, ref, noise = _data(1 << 17, 2)
    b = np.log1p(tol)
    pe = ens.a[:, None] + ens.k[:, None] * spread
    acc = (outside == 0) & np.all(pe <= b, axis=0)
    pred = np.zeros_like(ref)
    pred[S.OUT_ROWS] = np.exp(mean)
    viol = acc & np.any(S.violations(pred, ref, noise, tol), axis=0)
    assert acc.sum() > 0.05 * acc.size
    assert viol.sum() <= 2e-3 * acc.sum(), (int(viol.sum()), int(acc.sum()))
    # threshold zeros (spread -1) never block acceptance
    sp2 = spread.copy()
    sp2[3] = -1.0
    pe2 = np.where(sp2 < 0, 0.0, ens.a[:, None] + ens.k[:, None] * sp2)
    assert np.all(pe2 <= b, axis=0).sum() >= acc.sum()


def test_nothing_accepts_when_every_point_is_wrong():
    ens = _Ens()
    mean, spread, outside, ref, noise = _data(1 << 14, 3)
    mean = mean + 1.0  # every output off by e - 1
    rep = S.calibrate(ens, None, 1e-2, None, 16, data=(mean, spread, outside, ref, noise))
    assert rep["split_B"]["accepted"] == 0
    assert np.all(ens.k >= 1e38)


def test_accept_threshold_ties_and_minimum():
    score = np.array([1.0, 1.0, 2.0, 3.0, np.inf])
    bad = np.array([False, True, False, False, False])
    # 1 violation in the first two (tied) points: rate 1/2, then 1/3, 1/4
    assert S._accept_threshold(score, bad, 0.3, 1) == 3.0
    assert S._accept_threshold(score, bad, 0.3, 5) == -np.inf
    assert S._accept_threshold(score, bad, 0.0, 1) == -np.inf


def test_judge_margins():
    """The judge's targets: log(|pred - ref| / (tol |ref| + NOISE_K D)),
    clamped; outputs the FP64 path sets to zero (spread < 0) and exact
    predictions at the lower clamp; a wrong value where the bar is zero at
    the upper one."""
    tol = 0.1
    n = 6
    ref = np.zeros((20, n))
    noise = np.zeros((20, n))
    rows = S.OUT_ROWS
    ref[rows] = 2.0
    mean = np.full((S.OUT, n), np.log(2.0))
    spread = np.zeros((S.OUT, n), np.float32)
    mean[0, 1] = np.log(2.0 * 1.05)       # 5% off at tol 10%: margin log(0.1 / 0.2)
    mean[0, 2] = np.log(2.0 * 3.0)        # 200% off: log(4 / 0.2) = log 20, below the clamp
    noise[rows[0], 3] = 0.5               # the bar grows with the FP64 noise
    mean[0, 3] = np.log(2.0 * 1.5)        # |d| = 1, bar 0.2 + 16 * 0.5 = 8.2
    spread[1, 4] = -1.0                   # a threshold zero: exact
    ref[rows[1], 4] = 0.0
    ref[rows[2], 5] = 0.0                 # bar 0 and a positive prediction
    m = S.judge_margins(mean, spread, ref, noise, tol)
    assert m.shape == (S.OUT, n)
    assert m[0, 0] == S.MARGIN_LO          # exact
    assert np.isclose(m[0, 1], np.log(0.1 / 0.2))
    assert np.isclose(m[0, 2], min(np.log(4.0 / 0.2), S.MARGIN_HI))
    assert np.isclose(m[0, 3], np.log(1.0 / 8.2))
    assert m[1, 4] == S.MARGIN_LO
    assert m[2, 5] == S.MARGIN_HI
