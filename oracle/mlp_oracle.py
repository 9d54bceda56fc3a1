"""Numpy restatement of the surrogate MLP (bns_nurates_amd/csrc/mlp.hip).

TEST INFRASTRUCTURE ONLY: the GPU tests use it as the checker of the MLP's
forward pass and gradients.  It rounds to BF16 at the same places as the
kernels (inputs as high and low parts, weights, hidden activations,
back-propagated gradients)
and accumulates in float64, so that it differs from the GPU only by the
order of the FP32 accumulations and by BF16 roundings that such differences
flip.  The network has no counterpart in the reference; its parity is to
this restatement.
"""
import numpy as np

FEAT, OUT, K0, NL, HID, LAYERS = 15, 15, 32, 16, 256, 5
LAYER_K = [K0, HID, HID, HID, HID]
LAYER_N = [HID, HID, HID, HID, NL]


def offsets():
    """(weight offsets, bias offsets, floats per member) of the parameter blob."""
    wo, bo, o = [], [], 0
    for k, n in zip(LAYER_K, LAYER_N):
        wo.append(o)
        o += k * n
        bo.append(o)
        o += n
    return wo, bo, o


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).reshape(a.shape)


def unpack(params):
    wo, bo, _ = offsets()
    W = [params[wo[l]:wo[l] + LAYER_K[l] * LAYER_N[l]].reshape(LAYER_K[l], LAYER_N[l])
         for l in range(LAYERS)]
    b = [params[bo[l]:bo[l] + LAYER_N[l]] for l in range(LAYERS)]
    return W, b


def _sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def first_layer_inputs(x, exact=False):
    """The first layer's input rows [rows, 32]: feature i as x_hi = bf16(x)
    in column i and x_lo = bf16(x - x_hi) in column 15 + i (exact: x and 0)."""
    if exact:
        h = np.zeros((x.shape[0], K0), np.float64)
        h[:, :FEAT] = x
        return h
    h = np.zeros((x.shape[0], K0), np.float32)
    hi = bf16(np.asarray(x, np.float32))
    h[:, :FEAT] = hi
    h[:, FEAT:2 * FEAT] = bf16(np.asarray(x, np.float32) - hi)
    return h


def tied_w0(w0):
    """W_0 with rows 15..29 (the low parts' weights) mirroring rows 0..14."""
    w = np.array(w0, copy=True)
    w[FEAT:2 * FEAT] = w[:FEAT]
    return w


def forward(params, x, exact=False):
    """Pre-activations Z_l and activations H_l of one member on normalised
    features x [rows, 15]: rounded as the kernels round (BF16 operands, the
    features split into BF16 high and low parts sharing their weights, FP32
    results), or, with exact=True, in float64 throughout (the mathematical
    network, against which the gradient formulas are pinned)."""
    rnd = (lambda a: np.asarray(a, np.float64)) if exact else bf16
    dt = np.float64 if exact else np.float32
    W, b = unpack(params)
    W = [tied_w0(W[0])] + W[1:]
    hs = [first_layer_inputs(x, exact)]
    zs = []
    for l in range(LAYERS):
        z = (hs[-1].astype(np.float64) @ rnd(W[l]).astype(np.float64)).astype(dt)
        z = (z + b[l]).astype(dt)
        zs.append(z)
        if l + 1 < LAYERS:
            hs.append(rnd(z * _sig(z.astype(np.float64))))
    return zs, hs


def gradients(params, x, y, exact=False):
    """Loss and gradient blob of the mean-squared error over the 15 outputs
    (rounded as the kernels round, or exact as in forward).  A NaN target is
    a missing one (an output the FP64 path sets to zero): its error counts as
    0; the normalisation stays rows * 15."""
    rnd = (lambda a: np.asarray(a, np.float64)) if exact else bf16
    dt = np.float64 if exact else np.float32
    W, _ = unpack(params)
    W = [tied_w0(W[0])] + W[1:]
    zs, hs = forward(params, x, exact)
    rows = x.shape[0]
    wo, bo, n = offsets()
    g = np.zeros(n, np.float64)
    t = np.zeros((rows, NL), np.float32)
    t[:, :OUT] = y
    e = zs[-1].astype(np.float64) - t
    e[:, OUT:] = 0.0
    e[np.isnan(e)] = 0.0
    loss = float((e ** 2).sum() / (rows * OUT))
    d = (e * (2.0 / (rows * OUT))).astype(dt)
    for l in range(LAYERS - 1, -1, -1):
        dz = rnd(d)
        g[bo[l]:bo[l] + LAYER_N[l]] = d.astype(np.float64).sum(axis=0)
        g[wo[l]:wo[l] + LAYER_K[l] * LAYER_N[l]] = (
            hs[l].astype(np.float64).T @ dz.astype(np.float64)).ravel()
        if l == 0:
            # tied first layer: the low-part rows' gradient belongs to the
            # feature rows
            g0 = g[wo[0]:wo[0] + LAYER_K[0] * LAYER_N[0]].reshape(LAYER_K[0], LAYER_N[0])
            g0[:FEAT] += g0[FEAT:2 * FEAT]
            g0[FEAT:2 * FEAT] = 0.0
        if l:
            dh = dz.astype(np.float64) @ rnd(W[l]).astype(np.float64).T
            z = zs[l - 1].astype(np.float64)
            s = _sig(z)
            d = (dh * s * (1.0 + z * (1.0 - s))).astype(dt)
    return loss, g
