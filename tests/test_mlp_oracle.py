"""Pins of the MLP restatement (oracle/mlp_oracle.py) on the CPU: its
back-propagation is the derivative of its forward pass (central finite
differences of the float64 network), its BF16 rounding is round-to-nearest-
even, and its layout matches the product's parameter blob."""
import numpy as np

import oracle.mlp_oracle as M
from bns_nurates_amd import surrogate as S


def test_layout_matches_product():
    assert M.offsets() == S.offsets()
    assert M.offsets()[2] == 209936


def test_bf16_rounding():
    # BF16 spacing at 1 is 2^-7: 1 + 2^-8 is a tie (to even: 1), 1 + 3 2^-8 a
    # tie (to even: 1 + 2^-6), 1 + 3 2^-9 rounds up, 1 + 2^-9 down
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -9, -2.5,
                  3.0e38], np.float32)
    r = M.bf16(x)
    assert list(r[:6]) == [1.0, 1.0, 1.0 + 2 ** -6, 1.0 + 2 ** -7, 1.0, -2.5]
    assert (M.bf16(r) == r).all()


def test_gradients_are_the_derivative_of_the_loss():
    rng = np.random.default_rng(3)
    p = S.init_params(1).astype(np.float64)
    wo, bo, _ = M.offsets()
    for l in range(M.LAYERS):  # nonzero biases so that their gradients are exercised
        p[bo[l]:bo[l] + M.LAYER_N[l]] = rng.normal(0, 0.1, M.LAYER_N[l])
    p[wo[0] + M.FEAT * M.LAYER_N[0]:bo[0]] = 0.0
    x = rng.normal(size=(16, M.FEAT))
    y = rng.normal(size=(16, M.OUT))
    _, g = M.gradients(p, x, y, exact=True)
    idx = [wo[0] + 3, wo[1] + 1000, wo[2] + 5000, wo[3] + 77, wo[4] + 15 * 16 + 2, bo[0] + 9,
           bo[2] + 100, bo[4] + 14, bo[4] + 0]
    h = 1e-5
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.gradients(pp, x, y, exact=True)[0] - M.gradients(pm, x, y, exact=True)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])


def test_missing_targets_carry_no_error():
    """A NaN target (an output the FP64 path sets to zero) is left out of the
    loss: the gradients equal those with that target replaced by the
    network's own output, and the loss keeps its rows * 15 normalisation."""
    rng = np.random.default_rng(8)
    p = S.init_params(2).astype(np.float64)
    x = rng.normal(size=(32, M.FEAT))
    y = rng.normal(size=(32, M.OUT))
    miss = rng.random(y.shape) < 0.2
    zs, _ = M.forward(p, x, exact=True)
    y_nan, y_own = y.copy(), y.copy()
    y_nan[miss] = np.nan
    y_own[miss] = zs[-1][:, :M.OUT][miss]
    l1, g1 = M.gradients(p, x, y_nan, exact=True)
    l2, g2 = M.gradients(p, x, y_own, exact=True)
    # the targets are held in float32: a target equal to the output carries
    # the output's float32 rounding (relative 6e-8) as its error
    assert np.isfinite(g1).all() and np.isfinite(l1)
    assert abs(l1 - l2) <= 1e-12 * l2
    assert np.abs(g1 - g2).max() <= 1e-6 * np.abs(g2).max()
    # and a missing target is not a zero target
    y_zero = y.copy()
    y_zero[miss] = 0.0
    _, g3 = M.gradients(p, x, y_zero, exact=True)
    assert np.abs(g1 - g3).max() > 1e-3 * np.abs(g3).max()
