"""The learned form's bar excuses an output only as far as the FP64 path's
own rounding noise goes: |pred - ref| <= tol |ref| + NOISE_K D, D the change
of the FP64 output under 2^-46 perturbations of T and mu_e
(bns_nurates_amd/surrogate.py fp64_noise).  Pinned here on the CPU against
the oracle's running-error magnitude m (oracle/error_bound.cpp), with the
oracle's FP64 build standing in for the FP64 path:

* every output that m shows to be determined to tol (64 u m <= tol |ref|)
  is judged by the D bar as well (NOISE_K D <= tol |ref|): the bar is never
  looser than the oracle's condition-aware one;
* every output the FP64 path returns negative (an emissivity or opacity:
  rounding noise only) is excused (NOISE_K D >= |ref|);
* well-determined outputs (64 u m <= 1e-8 |ref|) get no excuse above 1e-8.
"""
import numpy as np
import pytest

import oracle as O
from bns_nurates_amd import surrogate as S
from bns_nurates_amd import workloads as W


@pytest.fixture(scope="module")
def sample():
    if not (O.available("bound") and O.available("orc")):
        O.build()
    pts = W.make_points(2048, seed=901)
    prm = W.params_default_all(W.BREM_BRT06)
    ref, mag = O.bound_m1_opacities(pts, 16, W.FLAGS_ALL, prm)
    r0 = O.m1_opacities(pts, 16, W.FLAGS_ALL, prm)
    D = np.zeros_like(r0)
    for f, fac in ((W.P_T, 1 + S.PERTURB), (W.P_MUE, 1 - S.PERTURB)):
        p = pts.copy()
        p[f] *= fac
        D = np.maximum(D, np.abs(O.m1_opacities(p, 16, W.FLAGS_ALL, prm) - r0))
    rows = S.OUT_ROWS
    return r0[rows], mag[rows], D[rows]


@pytest.mark.parametrize("tol", [1e-2, 3e-2, 1e-1])
def test_bar_is_never_looser_than_the_running_error_bound(sample, tol):
    R, M, D = sample
    det_m = 64 * O.U * np.abs(M) <= tol * np.abs(R)
    det_d = S.NOISE_K * D <= tol * np.abs(R)
    assert det_m.sum() > 0.9 * det_m.size
    assert not np.any(det_m & ~det_d), int(np.sum(det_m & ~det_d))


def test_negative_fp64_outputs_are_excused(sample):
    R, _, D = sample
    neg = R < 0
    assert neg.sum() > 0  # the sample reaches the degenerate regime
    assert np.all(S.NOISE_K * D[neg] >= np.abs(R[neg]))


def test_well_determined_outputs_get_no_excuse(sample):
    R, M, D = sample
    well = (64 * O.U * np.abs(M) <= 1e-8 * np.abs(R)) & (R != 0)
    assert well.sum() > 0.5 * well.size
    assert np.all(S.NOISE_K * D[well] <= 1e-8 * np.abs(R[well]))


def test_violations_rule():
    """violations(): beyond tol |ref| + NOISE_K D, exact zeros pass,
    non-finite predictions fail."""
    ref = np.ones((20, 4))
    noise = np.zeros((20, 4))
    pred = ref.copy()
    pred[S.OUT_ROWS[0], 1] = 1.0 + 2e-2      # beyond tol 1e-2
    pred[S.OUT_ROWS[1], 2] = 1.0 + 0.5e-2    # within
    pred[S.OUT_ROWS[2], 3] = np.nan
    noise[S.OUT_ROWS[3], 2] = 1.0            # a noisy output excuses
    pred[S.OUT_ROWS[3], 2] = 1.5
    ref[S.OUT_ROWS[4], 0] = pred[S.OUT_ROWS[4], 0] = 0.0
    bad = np.any(S.violations(pred, ref, noise, 1e-2), axis=0)
    assert list(bad) == [False, True, False, True]
