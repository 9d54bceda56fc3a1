"""The learned form of the gray M1 opacities: an ensemble of small MLPs
(bns_nurates_amd/csrc/mlp.hip) trained on data the FP64 path generates, with
per-point fallback to the FP64 path.

Features of a point (12): log nb, log T, Ye, and for nue, anue, nux: log n,
log J, chi (the heavy antineutrino copies nux).  Outputs (15): for nue,
anue, nux the logarithms of eta_0, eta, kappa_0_a, kappa_a, kappa_s.  Both
are standardised with statistics of the training set; the training box is
the per-feature range of the training set.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import api
from . import workloads as W

FEAT, OUT = 12, 15
# output block rows of the 15 outputs, species-major
OUT_ROWS = [r for s in range(3) for r in (0 + s, 8 + s, 4 + s, 12 + s, 16 + s)]


class MlpNorm(C.Structure):
    _fields_ = [("feat_mean", C.c_float * 12), ("feat_std", C.c_float * 12),
                ("out_mean", C.c_float * 15), ("out_std", C.c_float * 15),
                ("box_lo", C.c_float * 12), ("box_hi", C.c_float * 12)]


def features(pts: np.ndarray) -> np.ndarray:
    """Raw features [npts, 12] (float64) of a point block [21, npts]."""
    f = np.empty((pts.shape[1], FEAT))
    f[:, 0] = np.log(pts[W.P_NB])
    f[:, 1] = np.log(pts[W.P_T])
    f[:, 2] = pts[W.P_YP]
    for s in range(3):
        f[:, 3 + 3 * s] = np.log(pts[W.P_N0 + s])
        f[:, 4 + 3 * s] = np.log(pts[W.P_J0 + s])
        f[:, 5 + 3 * s] = pts[W.P_CHI0 + s]
    return f


def targets(out: np.ndarray):
    """Log outputs [npts, 15] of an output block [20, npts] and the mask of
    points whose 15 outputs are all positive and finite."""
    o = out[OUT_ROWS].T
    ok = np.all(np.isfinite(o) & (o > 0), axis=1)
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.log(np.where(o > 0, o, 1.0)), ok


# layer shapes (padded) and the parameter blob layout of bns_nurates_amd/csrc/mlp.h
LAYER_K = [32, 256, 256, 256, 256]
LAYER_N = [256, 256, 256, 256, 16]


def offsets():
    """(weight offsets, bias offsets, floats per member)."""
    wo, bo, o = [], [], 0
    for k, n in zip(LAYER_K, LAYER_N):
        wo.append(o)
        o += k * n
        bo.append(o)
        o += n
    return wo, bo, o


def init_params(seed: int) -> np.ndarray:
    """Glorot-uniform weights (zero in the padded rows and columns), zero biases."""
    rng = np.random.default_rng(seed)
    wo, bo, n = offsets()
    p = np.zeros(n, np.float32)
    fan_in = [FEAT, 256, 256, 256, 256]
    fan_out = [256, 256, 256, 256, OUT]
    for l in range(len(LAYER_K)):
        k, nn = LAYER_K[l], LAYER_N[l]
        a = np.sqrt(6.0 / (fan_in[l] + fan_out[l]))
        w = np.zeros((k, nn), np.float32)
        w[:fan_in[l], :fan_out[l]] = rng.uniform(-a, a, (fan_in[l], fan_out[l]))
        p[wo[l]:wo[l] + k * nn] = w.ravel()
    return p


def _fptr(t):
    return C.c_void_p(t.data_ptr())


class Ensemble:
    """An ensemble of `members` MLPs on one device (bnsamd_mlp_*).  Device
    arrays are torch tensors (float32, contiguous)."""

    def __init__(self, device: int = 0, members: int = 4, max_batch: int = 8192):
        self.lib = api.load_library()
        h = C.c_void_p()
        rc = self.lib.bnsamd_mlp_create(int(device), int(members), int(max_batch), C.byref(h))
        if rc:
            raise api.BnsAmdError(f"bnsamd_mlp_create: {api._STATUS.get(rc, rc)}")
        self.h, self.members, self.max_batch = h, members, max_batch
        self.nparams = self.lib.bnsamd_mlp_param_count()
        self.norm = None

    def close(self):
        if getattr(self, "h", None):
            self.lib.bnsamd_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc:
            msg = self.lib.bnsamd_mlp_last_error(self.h).decode(errors="replace")
            raise api.BnsAmdError(f"{what}: {api._STATUS.get(rc, rc)}: {msg}")

    def set_params(self, member: int, params: np.ndarray):
        p = np.ascontiguousarray(params, np.float32)
        assert p.size == self.nparams
        self._check(self.lib.bnsamd_mlp_set_params(self.h, member, p.ctypes.data),
                    "bnsamd_mlp_set_params")

    def get_params(self, member: int) -> np.ndarray:
        p = np.empty(self.nparams, np.float32)
        self._check(self.lib.bnsamd_mlp_get_params(self.h, member, p.ctypes.data),
                    "bnsamd_mlp_get_params")
        return p

    def set_norm(self, feat_mean, feat_std, out_mean, out_std, box_lo, box_hi):
        n = MlpNorm()
        for name, v in (("feat_mean", feat_mean), ("feat_std", feat_std), ("out_mean", out_mean),
                        ("out_std", out_std), ("box_lo", box_lo), ("box_hi", box_hi)):
            arr = getattr(n, name)
            for i, x in enumerate(np.asarray(v, np.float64)):
                arr[i] = float(x)
        self._check(self.lib.bnsamd_mlp_set_norm(self.h, C.byref(n)), "bnsamd_mlp_set_norm")
        self.norm = n

    def gradients(self, member: int, x, y):
        """Loss and gradient blob on rows of device tensors x [rows, 12], y [rows, 15]."""
        g = np.empty(self.nparams, np.float32)
        loss = C.c_float()
        self._check(self.lib.bnsamd_mlp_gradients(self.h, member, x.shape[0], _fptr(x), _fptr(y),
                                                  g.ctypes.data, C.byref(loss)),
                    "bnsamd_mlp_gradients")
        return loss.value, g

    def train(self, member: int, x, y, perm, epochs: int, batch: int, lr0: float, lr1: float):
        hist = np.zeros(epochs, np.float32)
        self._check(self.lib.bnsamd_mlp_train(self.h, member, x.shape[0], _fptr(x), _fptr(y),
                                              _fptr(perm), int(epochs), int(batch), float(lr0),
                                              float(lr1), hist.ctypes.data), "bnsamd_mlp_train")
        return hist

    def predict(self, member: int, x, y):
        self._check(self.lib.bnsamd_mlp_predict(self.h, member, x.shape[0], _fptr(x), _fptr(y)),
                    "bnsamd_mlp_predict")
