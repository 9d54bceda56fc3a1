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
