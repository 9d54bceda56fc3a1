"""The learned form of the gray M1 opacities: an ensemble of small MLPs
(bns_nurates_amd/csrc/mlp.hip) trained on data the FP64 path generates, with
per-point fallback to the FP64 path.

Features of a point (15): log nb, log T, Ye, and for nue, anue, nux: log n,
log J, chi (the heavy antineutrino copies nux); then mu_e / T,
(mu_n - mu_p) / T and 1 / T.  The last three are inputs of the opacity
function itself (the chemical potentials of the point block): the rates
depend on them through exponentials such as e^{-eta_e}, which a linear
model in these features carries exactly.  Outputs (15): for nue,
anue, nux the logarithms of eta_0, eta, kappa_0_a, kappa_a, kappa_s.  Both
are standardised with statistics of the training set; the training box is
the per-feature range of the training set.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import api
from . import workloads as W

FEAT, OUT = 15, 15
# output block rows of the 15 outputs, species-major
OUT_ROWS = [r for s in range(3) for r in (0 + s, 8 + s, 4 + s, 12 + s, 16 + s)]


class MlpNorm(C.Structure):
    _fields_ = [("feat_mean", C.c_float * FEAT), ("feat_std", C.c_float * FEAT),
                ("out_mean", C.c_float * OUT), ("out_std", C.c_float * OUT),
                ("box_lo", C.c_float * FEAT), ("box_hi", C.c_float * FEAT),
                ("lin", (C.c_float * OUT) * (FEAT + 1)),
                ("res_lo", C.c_float * 15), ("res_hi", C.c_float * 15),
                ("err_a", C.c_float * 15), ("err_k", C.c_float * 15)]


class MlpDomain(C.Structure):
    _fields_ = [("flags", api._Flags), ("params", api._Params), ("stim_abs", C.c_int),
                ("nquad", C.c_int), ("dU", C.c_double), ("dm_eff", C.c_double),
                ("beta_guard", C.c_double)]


def features(pts: np.ndarray) -> np.ndarray:
    """Raw features [npts, 15] (float64) of a point block [21, npts]; a zero
    density gives -inf (outside any training box)."""
    f = np.empty((pts.shape[1], FEAT))
    f[:, 0] = np.log(pts[W.P_NB])
    f[:, 1] = np.log(pts[W.P_T])
    f[:, 2] = pts[W.P_YP]
    for s in range(3):
        f[:, 3 + 3 * s] = np.log(pts[W.P_N0 + s])
        f[:, 4 + 3 * s] = np.log(pts[W.P_J0 + s])
        f[:, 5 + 3 * s] = pts[W.P_CHI0 + s]
    rT = 1.0 / pts[W.P_T]
    f[:, 12] = pts[W.P_MUE] * rT
    f[:, 13] = (pts[W.P_MUN] - pts[W.P_MUP]) * rT
    f[:, 14] = rT
    return f


def targets(out: np.ndarray):
    """Log outputs [npts, 15] of an output block [20, npts], NaN where an
    output is not positive and finite (the FP64 path's threshold zeros: such
    a target is missing, not trained on), and the mask of points with at
    least one target."""
    o = out[OUT_ROWS].T
    pos = np.isfinite(o) & (o > 0)
    with np.errstate(divide="ignore", invalid="ignore"):
        t = np.where(pos, np.log(np.where(pos, o, 1.0)), np.nan)
    return t, np.any(pos, axis=1)


def masked_linear_fit(X: np.ndarray, t: np.ndarray) -> np.ndarray:
    """Least-squares coefficients [cols, outputs] of t ~ X per output, each
    on the rows where its target is present (outputs sharing a mask of
    present rows are fitted together)."""
    pres = ~np.isnan(t)
    A = np.zeros((X.shape[1], t.shape[1]))
    keys = [pres[:, o].tobytes() for o in range(t.shape[1])]
    done = set()
    for o in range(t.shape[1]):
        if o in done:
            continue
        group = [j for j in range(t.shape[1]) if keys[j] == keys[o]]
        rows = pres[:, o]
        A[:, group] = np.linalg.lstsq(X[rows], t[rows][:, group], rcond=None)[0]
        done.update(group)
    return A


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
        if getattr(self, "judge", None) is not None:
            self.judge.close()
            self.judge = None
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

    def set_norm(self, feat_mean, feat_std, out_mean, out_std, box_lo, box_hi, lin, res_lo,
                 res_hi):
        n = MlpNorm()
        for name, v in (("feat_mean", feat_mean), ("feat_std", feat_std), ("out_mean", out_mean),
                        ("out_std", out_std), ("box_lo", box_lo), ("box_hi", box_hi),
                        ("res_lo", res_lo), ("res_hi", res_hi),
                        ("err_a", np.zeros(OUT)), ("err_k", np.ones(OUT))):
            arr = getattr(n, name)
            for i, x in enumerate(np.asarray(v, np.float64)):
                arr[i] = float(x)
        lin = np.asarray(lin, np.float64)
        for i in range(FEAT + 1):
            for o in range(OUT):
                n.lin[i][o] = float(lin[i, o])
        self._check(self.lib.bnsamd_mlp_set_norm(self.h, C.byref(n)), "bnsamd_mlp_set_norm")
        self.norm = n

    def gradients(self, member: int, x, y):
        """Loss and gradient blob on rows of device tensors x [rows, 15], y [rows, 15]."""
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

    def train_all(self, x, y, perms, epochs: int, batch: int, lr0: float, lr1: float):
        """Every member at once (bnsamd_mlp_train_ensemble): member e reads
        rows in the order perms[e] (device int32 [members, n]); returns the
        epoch losses [members, epochs]."""
        assert perms.shape == (self.members, x.shape[0]) and perms.is_contiguous()
        hist = np.zeros((self.members, epochs), np.float32)
        self._check(self.lib.bnsamd_mlp_train_ensemble(
            self.h, x.shape[0], _fptr(x), _fptr(y), _fptr(perms), int(epochs), int(batch),
            float(lr0), float(lr1), hist.ctypes.data), "bnsamd_mlp_train_ensemble")
        return hist

    def predict(self, member: int, x, y):
        self._check(self.lib.bnsamd_mlp_predict(self.h, member, x.shape[0], _fptr(x), _fptr(y)),
                    "bnsamd_mlp_predict")

    def set_domain(self, flags, params, stim_abs: int, nquad: int, dU: float, dm_eff: float,
                   beta_guard: float = 0.0):
        """The configuration of the training targets (bnsamd_mlp_set_domain)."""
        d = MlpDomain(api.as_flags(flags).c(), api.as_params(params).c(), int(stim_abs),
                      int(nquad), float(dU), float(dm_eff), float(beta_guard))
        self._check(self.lib.bnsamd_mlp_set_domain(self.h, C.byref(d)), "bnsamd_mlp_set_domain")
        self.domain = d

    def set_acceptance(self, err_a, err_k):
        """Predicted log error err_a + err_k * spread per output
        (bnsamd_mlp_set_acceptance)."""
        a = np.ascontiguousarray(err_a, np.float32)
        k = np.ascontiguousarray(err_k, np.float32)
        assert a.shape == (OUT,) and k.shape == (OUT,)
        self._check(self.lib.bnsamd_mlp_set_acceptance(self.h, a.ctypes.data, k.ctypes.data),
                    "bnsamd_mlp_set_acceptance")
        self.err_a, self.err_k = a.copy(), k.copy()

    def set_judge(self, judge: "Ensemble | None", threshold: float = 0.0):
        """Acceptance judge (bnsamd_mlp_set_judge): member 0 of `judge`
        (copied), rejecting a point whose largest predicted log margin
        exceeds threshold; None removes it."""
        self._check(self.lib.bnsamd_mlp_set_judge(self.h, judge.h if judge is not None else None,
                                                  float(threshold)), "bnsamd_mlp_set_judge")
        self.judge_threshold = None if judge is None else float(threshold)

    def infer(self, ctx, pts):
        """Raw predictions on a device point block pts [21, >= n]: mean log
        outputs [15, n] (float64), their spreads [15, n] (float32) and the
        outside-the-box-or-domain flags [n] (int32), device tensors."""
        import torch
        n = pts.shape[1]
        pp, ld = api._tensor_ptr(pts, 21, n, "pts")
        ctx._follow_torch_stream()
        mean = torch.empty((OUT, ld), dtype=torch.float64, device=pts.device)
        spread = torch.empty((OUT, ld), dtype=torch.float32, device=pts.device)
        outside = torch.empty(ld, dtype=torch.int32, device=pts.device)
        self._check(self.lib.bnsamd_mlp_infer(ctx.h, self.h, n, ld, C.c_void_p(pp),
                                              C.c_void_p(mean.data_ptr()),
                                              C.c_void_p(spread.data_ptr()),
                                              C.c_void_p(outside.data_ptr())), "bnsamd_mlp_infer")
        return mean[:, :n], spread[:, :n], outside[:n]

    def m1_opacities(self, ctx, pts, out, error_bound: float, flags=None, params=None,
                     stim_abs: int = 1) -> int:
        """The learned path on a device point block pts [21, >= n] into out
        [20, >= n] (torch float64 tensors, ctx's quadrature for the fallback):
        points whose predicted log error exceeds error_bound on any output
        (log1p(tol) after calibrate(..., tol)) are recomputed by the FP64
        path; returns their number.  Synchronous."""
        n = pts.shape[1]
        pp, ld = api._tensor_ptr(pts, 21, n, "pts")
        op, ldo = api._tensor_ptr(out, 20, n, "out")
        if ldo != ld:
            raise api.BnsAmdError("pts and out must share the leading dimension")
        ctx._follow_torch_stream()
        fl, pr = api.as_flags(flags).c(), api.as_params(params).c()
        nfb = C.c_int64()
        self._check(self.lib.bnsamd_surrogate_m1_opacities(
            ctx.h, self.h, n, ld, C.c_void_p(pp), C.c_float(error_bound), C.byref(fl),
            C.byref(pr), int(stim_abs), C.c_void_p(op), C.byref(nfb)),
            "bnsamd_surrogate_m1_opacities")
        return int(nfb.value)


def check_domain_rules(pts: np.ndarray):
    """The point-block rules of the training domain (bnsamd_mlp_domain):
    anux moments equal to nux's, yn = 1 - yp, one dU and one dm_eff value.
    Returns (dU, dm_eff)."""
    for f in (W.P_N0, W.P_J0, W.P_CHI0):
        assert np.array_equal(pts[f + 3], pts[f + 2]), "anux moments must equal nux's"
    assert np.array_equal(pts[W.P_YN], 1.0 - pts[W.P_YP]), "yn must be 1 - yp"
    du, dm = np.unique(pts[W.P_DU]), np.unique(pts[W.P_DMEFF])
    assert du.size == 1 and dm.size == 1, "dU and dm_eff must be constant"
    return float(du[0]), float(dm[0])


def train_ensemble(ctx, npts: int, nquad: int, params, members: int = 4, epochs: int = 8,
                   batch: int = 4096, lr0: float = 4e-3, lr1: float = 2e-5, seed: int = 1000,
                   chunk: int = 1 << 21, partial_targets: bool = False, ln_loss: bool = False,
                   beta_guard: float = 0.1, log=None) -> "Ensemble":
    """An ensemble trained on npts seeded points (workload distributions,
    seed `seed`) with targets from the FP64 path at quadrature order nquad
    (all reactions, stimulated absorption); member e starts from its own
    initialisation and data order.  ln_loss: the residuals share one scale,
    so that the loss is the mean squared error in natural-log units (the
    unit of the acceptance bar) instead of per-output standardised units.
    beta_guard: points with |mu_np| < beta_guard T (mu_np = mu_n - mu_p - dQ -
    dU of the beta rates) are neither trained on nor accepted: the
    reference's beta rates jump across the branch surfaces of EtaNNAbs
    there (formulas_beta_iso.hpp:79-96), which no smooth network fits."""
    import torch
    pts = W.make_points(npts, seed=seed)
    dU, dm_eff = check_domain_rules(pts)
    prm = api.as_params(params)
    shift = (dm_eff if prm.use_dm_eff else W.Q_NP) + (dU if prm.use_dU else 0.0)
    ctx.set_quadrature(nquad)
    outs = []
    for a in range(0, npts, chunk):
        d = torch.from_numpy(np.ascontiguousarray(pts[:, a:a + chunk])).cuda()
        outs.append(ctx.compute_m1_opacities(d, flags=W.FLAGS_ALL, params=params).cpu().numpy())
    out = np.concatenate(outs, axis=1)
    with np.errstate(divide="ignore"):
        f = features(pts)
    # points with finite features and all 15 outputs positive; with
    # partial_targets also those where the FP64 path sets some outputs to
    # zero (n or J below its thresholds): those outputs are missing targets,
    # left out of the fits and the loss.  (Measured on the 128^3 grid: the
    # training box then holds 99.99% of the points instead of 93.2%, but the
    # extreme-degeneracy points it adds cost accuracy elsewhere -- all 15
    # outputs within 3e-2 on 12.8% of the points instead of 31.8% -- for the
    # same speed, so they are left out by default.)
    t, ok = targets(out)
    if not partial_targets:
        ok &= ~np.any(np.isnan(t), axis=1)
    ok &= np.all(np.isfinite(f), axis=1)
    ok &= ~(np.abs(pts[W.P_MUN] - pts[W.P_MUP] - shift) < beta_guard * pts[W.P_T])
    f, t = f[ok], t[ok]
    fm, fs = f.mean(0), f.std(0)
    xn = (f - fm) / fs
    # linear model of the log outputs in the standardised features; the
    # network learns the standardised residual
    X1 = np.c_[np.ones(len(xn)), xn]
    A = masked_linear_fit(X1, t)
    res = t - X1 @ A
    del X1
    tm, ts = np.nanmean(res, 0), np.nanstd(res, 0)
    if ln_loss:
        ts = np.full_like(ts, np.sqrt(np.mean(ts ** 2)))
    ys = (res - tm) / ts
    x = torch.from_numpy(xn.astype(np.float32)).cuda()
    y = torch.from_numpy(ys.astype(np.float32)).cuda()
    ens = Ensemble(ctx.device, members=members, max_batch=batch)
    ens.set_norm(fm, fs, tm, ts, f.min(0), f.max(0), A, np.nanmin(ys, 0), np.nanmax(ys, 0))
    ens.set_domain(W.FLAGS_ALL, params, 1, nquad, dU, dm_eff, beta_guard)
    ens.lin = A
    ens.n_train, ens.history = int(ok.sum()), []
    # each member from its own initialisation and data order, all trained at
    # once on their own streams
    perms = np.empty((members, len(f)), np.int32)
    for e in range(members):
        ens.set_params(e, init_params(seed + 17 * e))
        perms[e] = np.random.default_rng(seed + 31 * e).permutation(len(f))
    hist = ens.train_all(x, y, torch.from_numpy(perms).cuda(), epochs, batch, lr0, lr1)
    del perms
    for e in range(members):
        ens.history.append(hist[e])
        if log:
            log(f"surrogate member {e}: epoch losses {np.array2string(hist[e], precision=5)}")
    return ens


# ---- acceptance rule ------------------------------------------------------------
# The FP64 path's own rounding noise: outputs formed by cancellation (pair and
# NES/NPS rates at large degeneracy, some stimulated-absorption opacities)
# carry rounding errors far above 1e-2 of their value, and can even come out
# negative.  A relative tolerance against such a value judges nothing, so an
# accepted output violates the bar when |pred - ref| > tol |ref| + NOISE_K D,
# D the FP64 path's change under input perturbations of PERTURB (relative)
# in T and mu_e: far above an ulp, so that the roundings change, and far
# below any tolerance in its effect on the true value.
PERTURB = 2.0 ** -46
NOISE_K = 16.0


def fp64_noise(ctx, d, params, ref, flags=None):
    """Noise estimate D [20, n] of the FP64 outputs ref of the device point
    block d (ctx's quadrature): the largest change of each output when T or
    mu_e is scaled by 1 +- PERTURB."""
    import torch
    flags = W.FLAGS_ALL if flags is None else flags
    n = d.shape[1]
    D = torch.zeros_like(ref[:, :n])
    for field, f in ((W.P_T, 1.0 + PERTURB), (W.P_MUE, 1.0 - PERTURB)):
        dp = d.clone()
        dp[field] *= f
        r2 = ctx.compute_m1_opacities(dp, flags=flags, params=params)
        D = torch.maximum(D, (r2[:, :n] - ref[:, :n]).abs())
        del dp, r2
    return D


def violations(pred: np.ndarray, ref: np.ndarray, noise: np.ndarray, tol: float) -> np.ndarray:
    """[15, n] True where an output of the learned form misses the bar
    |pred - ref| > tol |ref| + NOISE_K D (rows OUT_ROWS of the output
    blocks; exact zeros on both sides pass)."""
    p, r, D = pred[OUT_ROWS], ref[OUT_ROWS], noise[OUT_ROWS]
    with np.errstate(invalid="ignore"):
        bad = np.abs(p - r) > tol * np.abs(r) + NOISE_K * D
    return bad | ~np.isfinite(p)


def prediction_errors(mean: np.ndarray, spread: np.ndarray, outside: np.ndarray,
                      ref: np.ndarray, noise: np.ndarray, tol: float):
    """Per output and point: the log error |mean - log ref| where the FP64
    value is positive and determined to tol (NOISE_K D <= tol |ref|; NaN
    elsewhere: such outputs do not shape the error model), the bar's verdict
    on exp(mean) (violations()), the spread, and the mask of points inside
    the training box and domain.  Threshold zeros (spread < 0) are exact."""
    t = _log_targets(ref)
    r, D = ref[OUT_ROWS], noise[OUT_ROWS]
    judged = (r > 0) & (NOISE_K * D <= tol * np.abs(r))
    with np.errstate(invalid="ignore", over="ignore"):
        e = np.where(judged, np.abs(mean - t), np.nan)
        pred = np.zeros_like(ref)
        pred[OUT_ROWS] = np.exp(mean)
    zero_pred = spread < 0.0
    pred[OUT_ROWS] = np.where(zero_pred, 0.0, pred[OUT_ROWS])
    bad = violations(pred, ref, noise, tol)
    e = np.where(zero_pred, 0.0, e)
    return e, bad, spread.astype(np.float64), outside == 0


def _log_targets(ref: np.ndarray) -> np.ndarray:
    """Natural logs [15, n] of the FP64 outputs, NaN where an output is not
    positive."""
    o = ref[OUT_ROWS]
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(o > 0, np.log(np.where(o > 0, o, 1.0)), np.nan)


def clopper_pearson_upper(k: int, n: int, conf: float = 0.95) -> float:
    """One-sided upper confidence bound of a binomial proportion."""
    if n == 0 or k >= n:
        return 1.0
    from scipy.stats import beta
    return float(beta.ppf(conf, k + 1, n - k))


def _accept_threshold(score: np.ndarray, bad: np.ndarray, target: float, minacc: int):
    """The largest score threshold t for which the points with score <= t
    number at least minacc and have a violation fraction <= target (-inf
    when there is none)."""
    o = np.argsort(score, kind="stable")
    sc, bb = score[o], bad[o]
    fin = np.isfinite(sc)
    k = np.arange(1, len(sc) + 1)
    ok = (np.cumsum(bb) <= target * k) & (k >= minacc) & fin
    # ties: a threshold accepts every point at its score
    last_of_tie = np.r_[sc[1:] != sc[:-1], True]
    ok &= last_of_tie
    return float(sc[np.nonzero(ok)[0].max()]) if ok.any() else -np.inf


def calibrate(ens: "Ensemble", ctx, tol: float, params, nquad: int, ncal: int = 1 << 20,
              seed: int = 2000, coverage: float = 0.999, q: float = 0.99, minacc: int = 3000,
              log=None, data=None) -> dict:
    """Calibrates the acceptance rule for a relative tolerance tol on every
    output and sets it on the ensemble; the learned form is then called with
    error_bound = log1p(tol).

    Two disjoint seeded splits of ncal points from the workload
    distributions, with FP64 targets (and their noise estimate, fp64_noise)
    from the same path and quadrature as the training targets.  A rule
    accepts a point when max_o (a_o + k_o s_o) <= t over its outputs (s_o
    the spread, threshold zeros excepted).  Three shapes (a, k) are tried:
    the spread itself, the spread over its median per output, and per output
    the q-quantile ratio of log error to (spread + delta_o) with the delta_o
    (a multiple of log1p(tol)) that accepts most points for that output
    alone.  On split A each shape gets the threshold t that accepts most
    points while at most (1 - coverage) / 2 of them have an output beyond the
    bar (violations()), and at least minacc are accepted; the shape that
    accepts most wins, and split B reports it (accepted fraction, violations
    among the accepted, 95% Clopper-Pearson upper bound of their rate).
    data: the splits' (mean, spread, outside, ref, noise) from
    calibration_data(), to share them between tolerances."""
    b = float(np.log1p(tol))
    if data is None:
        data = calibration_data(ens, ctx, params, nquad, 2 * ncal, seed)
    mean, spread, outside, ref, noise = data
    n = ref.shape[1]
    e, bad, sd, inside = prediction_errors(mean, spread, outside, ref, noise, tol)
    A, B = slice(0, n // 2), slice(n // 2, n)
    sdp = np.where(sd < 0, 0.0, sd)
    live_all = inside[None, :] & (sd >= 0) & np.isfinite(sd)
    med = np.array([np.median(sdp[o, A][live_all[o, A]]) if live_all[o, A].any() else 1.0
                    for o in range(OUT)])
    shapes = {"spread": (np.zeros(OUT), np.ones(OUT)),
              "spread_over_median": (np.zeros(OUT), 1.0 / np.maximum(med, 1e-30))}
    qa, qk = np.zeros(OUT), np.ones(OUT)
    for o in range(OUT):
        live = live_all[o, A] & ~np.isnan(e[o, A])
        if not live.any():
            continue
        eo, so = e[o, A][live], sd[o, A][live]
        best = None
        for frac in (0.0, 0.03, 0.1, 0.3, 1.0):
            delta = frac * b
            with np.errstate(divide="ignore", invalid="ignore"):
                r = np.where(eo == 0, 0.0, eo / (so + delta))
            r = np.where(np.isnan(r), np.inf, r)
            k = float(np.quantile(r, q))
            if not np.isfinite(k) or k <= 0:
                continue
            acc = float(np.mean(k * (so + delta) <= b))
            if best is None or acc > best[0]:
                best = (acc, k, delta)
        if best is not None:
            qa[o], qk[o] = best[1] * best[2], best[1]
    shapes["quantile_ratio"] = (qa, qk)

    def score(a_, k_, cols):
        with np.errstate(invalid="ignore", over="ignore"):
            pe = np.where(sd[:, cols] < 0, 0.0, a_[:, None] + k_[:, None] * sdp[:, cols])
        return np.where(inside[cols], pe.max(axis=0), np.inf)

    anybad = np.any(bad, axis=0)
    target = (1.0 - coverage) / 2
    fits = {}
    for name, (a_, k_) in shapes.items():
        sA = score(a_, k_, A)
        t = _accept_threshold(sA, anybad[A], target, minacc)
        fits[name] = (t, int((sA <= t).sum()) if np.isfinite(t) else 0)
    name = max(fits, key=lambda k: fits[k][1])
    t = fits[name][0]
    a_, k_ = shapes[name]
    if np.isfinite(t) and t > 0:
        a_set, k_set = a_ * b / t, k_ * b / t
    else:  # nothing accepts: every output's predicted error above any bound
        a_set, k_set = np.full(OUT, 3e38), np.full(OUT, 3e38)
    a_set = np.where(np.isfinite(a_set) & (a_set < 3e38), a_set, 3e38).astype(np.float32)
    k_set = np.where(np.isfinite(k_set) & (k_set < 3e38), k_set, 3e38).astype(np.float32)
    ens.set_acceptance(a_set, k_set)
    ens.set_judge(None)  # this rule alone decides
    rep = {"tol": tol, "error_bound_log": b, "rule": name,
           "candidates_split_A_accepted": {k: v[1] for k, v in fits.items()},
           "coverage_target": coverage, "calibration_points": n,
           "bar": f"|pred - ref| <= tol |ref| + {NOISE_K:g} D (D: FP64 change under "
                  f"2^-46 perturbations of T and mu_e)",
           "inside_frac": round(float(inside.mean()), 4),
           "err_a": [float(f"{v:.3g}") for v in a_set], "err_k": [float(f"{v:.3g}") for v in k_set]}
    for tag, cols in (("split_A", A), ("split_B", B)):
        with np.errstate(invalid="ignore", over="ignore"):
            pe = np.where(sd[:, cols] < 0, 0.0,
                          a_set[:, None].astype(np.float64) + k_set[:, None] * sdp[:, cols])
        acc = inside[cols] & np.all(pe <= b, axis=0)
        na, nv = int(acc.sum()), int((acc & anybad[cols]).sum())
        rep[tag] = {"accepted_frac": round(na / int(np.sum(np.ones(n, bool)[cols])), 5),
                    "accepted": na, "violations": nv,
                    "violation_rate_ucb95": float(f"{clopper_pearson_upper(nv, na):.3g}")}
    if log:
        log(f"calibrate tol={tol}: rule {name}, {rep['split_B']}")
    return rep


# ---- acceptance judge -------------------------------------------------------------
# The ensemble's spread separates good from bad predictions poorly: at tol 0.1
# the network has every output within the bar on ~97% of the points inside its
# domain, but no spread threshold isolates the other 3%.  A judge network of
# the members' shape (15 features -> 15 outputs), trained by the library's own
# kernels on held-out points, predicts each output's log margin to the bar;
# a point is accepted when every predicted margin lies below a threshold
# calibrated on further held-out points.  The margins are clamped to
# [MARGIN_LO, MARGIN_HI] so that the network's capacity goes to the range
# around the threshold (about -1.7 at tol 0.1): on the 128^3 grid's
# distributions this clamp accepted 63% at tol 0.1 and 6.1% at 3e-2, against
# 61% and 5.0% with [-8, 4] (same coverage).
MARGIN_LO, MARGIN_HI = -3.0, 0.5


def judge_margins(mean: np.ndarray, spread: np.ndarray, ref: np.ndarray, noise: np.ndarray,
                  tol: float) -> np.ndarray:
    """Log margins [15, n] of the ensemble's predictions exp(mean) to the bar
    |pred - ref| <= tol |ref| + NOISE_K D, clamped to [MARGIN_LO, MARGIN_HI]:
    log(|pred - ref| / (tol |ref| + NOISE_K D)); exact outputs (threshold
    zeros, spread < 0) at MARGIN_LO."""
    r, D = ref[OUT_ROWS], noise[OUT_ROWS]
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        p = np.where(spread < 0, 0.0, np.exp(mean))
        m = np.log(np.abs(p - r) / (tol * np.abs(r) + NOISE_K * D))
    m = np.where(np.isnan(m), MARGIN_LO, m)  # 0 / 0: exact
    return np.clip(m, MARGIN_LO, MARGIN_HI)


def _norm_arrays(ens: "Ensemble"):
    n = ens.norm
    return (np.array(n.feat_mean[:], np.float64), np.array(n.feat_std[:], np.float64),
            np.array(n.box_lo[:], np.float64), np.array(n.box_hi[:], np.float64))


def train_judge(ens: "Ensemble", ctx, tol: float, pts: np.ndarray, data, rows, epochs: int = 16,
                batch: int = 4096, lr0: float = 4e-3, lr1: float = 2e-5, seed: int = 3000,
                log=None) -> "Ensemble":
    """A one-member network that predicts judge_margins() of ens at tol from
    the 15 features, trained (library kernels, MSE on the standardised
    residual of a linear model, as the members) on the points `rows` of the
    calibration points pts / data (calibration_data()) that lie inside the
    training box and domain.  Its feature standardisation and box are the
    ensemble's, as bnsamd_mlp_set_judge requires."""
    import torch
    mean, spread, outside, ref, noise = data
    rows = np.asarray(rows)
    ok = rows[outside[rows] == 0]
    t = judge_margins(mean[:, ok], spread[:, ok], ref[:, ok], noise[:, ok], tol).T
    with np.errstate(divide="ignore"):
        f = features(pts[:, ok])
    fm, fs, lo, hi = _norm_arrays(ens)
    xn = (f - fm.astype(np.float32)) / fs.astype(np.float32)
    X1 = np.c_[np.ones(len(xn)), xn]
    A = masked_linear_fit(X1, t)
    res = t - X1 @ A
    del X1
    tm, ts = res.mean(0), res.std(0)
    ts = np.where(ts > 0, ts, 1.0)
    ys = (res - tm) / ts
    j = Ensemble(ctx.device, members=1, max_batch=batch)
    j.set_norm(fm, fs, tm, ts, lo, hi, A, ys.min(0), ys.max(0))
    d = ens.domain
    j.domain = d
    j._check(j.lib.bnsamd_mlp_set_domain(j.h, C.byref(d)), "bnsamd_mlp_set_domain")
    j.set_params(0, init_params(seed))
    perm = np.random.default_rng(seed).permutation(len(xn)).astype(np.int32)[None, :]
    hist = j.train_all(torch.from_numpy(xn.astype(np.float32)).cuda(),
                       torch.from_numpy(ys.astype(np.float32)).cuda(),
                       torch.from_numpy(np.ascontiguousarray(perm)).cuda(), epochs, batch, lr0, lr1)
    j.n_train, j.history = int(len(xn)), [hist[0]]
    if log:
        log(f"judge tol={tol}: {len(xn)} points, epoch losses {np.array2string(hist[0], precision=4)}")
    return j


def judge_scores(j: "Ensemble", ctx, pts: np.ndarray, outside: np.ndarray) -> np.ndarray:
    """Largest predicted margin of each point (inf outside the training box
    or domain, or where a predicted residual leaves the judge's training
    range; exact outputs excepted), as the learned form's kernel forms it."""
    import torch
    d = torch.from_numpy(np.ascontiguousarray(pts)).cuda()
    m, sp, out = (x.cpu().numpy() for x in j.infer(ctx, d))
    del d
    sc = np.where(sp < 0, -np.inf, np.where(np.isfinite(sp), m, np.inf)).max(axis=0)
    return np.where((outside == 0) & (out == 0), sc, np.inf)


def calibrate_judge(ens: "Ensemble", ctx, tol: float, params, nquad: int, data=None,
                    pts=None, ncal: int = 1 << 22, seed: int = 2000, coverage: float = 0.999,
                    minacc: int = 3000, epochs: int = 16, log=None,
                    cal_frac: float = 0.5) -> dict:
    """Trains a judge for tol on the first half of ncal held-out points,
    calibrates its threshold on the third quarter (the largest that accepts
    at most cal_frac (1 - coverage) points with an output beyond the bar, and
    at least minacc points), reports on the last quarter, and sets it on the
    ensemble (the spread rule off: err_a = err_k = 0).  data / pts:
    calibration_data() and the points it was made on (seed `seed`)."""
    if data is None:
        data = calibration_data(ens, ctx, params, nquad, ncal, seed)
        pts = W.make_points(ncal, seed=seed)
    mean, spread, outside, ref, noise = data
    n = ref.shape[1]
    A1, A2, B = np.arange(0, n // 2), np.arange(n // 2, 3 * n // 4), np.arange(3 * n // 4, n)
    j = train_judge(ens, ctx, tol, pts, data, A1, epochs=epochs, log=log)
    ctx.set_quadrature(nquad)
    _, bad, _, _ = prediction_errors(mean, spread, outside, ref, noise, tol)
    anybad = np.any(bad, axis=0)
    target = cal_frac * (1.0 - coverage)
    sA2 = judge_scores(j, ctx, pts[:, A2], outside[A2])
    t = _accept_threshold(sA2, anybad[A2], target, minacc)
    rep = {"tol": tol, "rule": "judge", "coverage_target": coverage, "calibration_points": n,
           "judge_train_points": j.n_train, "judge_epochs": epochs,
           "judge_final_loss": float(f"{j.history[0][-1]:.4g}"),
           "bar": f"|pred - ref| <= tol |ref| + {NOISE_K:g} D (D: FP64 change under "
                  f"2^-46 perturbations of T and mu_e)",
           "threshold": float(t) if np.isfinite(t) else None,
           "inside_frac": round(float((outside == 0).mean()), 4)}
    for tag, cols, sc in (("split_A2", A2, sA2),
                          ("split_B", B, judge_scores(j, ctx, pts[:, B], outside[B]))):
        acc = sc <= t
        na, nv = int(acc.sum()), int((acc & anybad[cols]).sum())
        rep[tag] = {"accepted_frac": round(na / len(cols), 5), "accepted": na, "violations": nv,
                    "violation_rate_ucb95": float(f"{clopper_pearson_upper(nv, na):.3g}")}
    # the spread rule off; a threshold of -inf rejects every point.  The
    # library keeps its own copy of the judge; ens.judge keeps this one for
    # diagnostics (judge_scores) until the next calibration or close()
    ens.set_acceptance(np.zeros(OUT), np.zeros(OUT))
    ens.set_judge(j, t if np.isfinite(t) else -np.inf)
    if getattr(ens, "judge", None) is not None:
        ens.judge.close()
    ens.judge = j
    if log:
        log(f"calibrate_judge tol={tol}: threshold {rep['threshold']}, {rep['split_B']}")
    return rep


def calibration_data(ens: "Ensemble", ctx, params, nquad: int, npts: int, seed: int):
    """Raw predictions and FP64 targets on npts seeded workload points:
    (mean [15, n], spread [15, n], outside [n], ref [20, n], noise [20, n]),
    host arrays."""
    import torch
    pts = W.make_points(npts, seed=seed)
    ctx.set_quadrature(nquad)
    d = torch.from_numpy(pts).cuda()
    ref = ctx.compute_m1_opacities(d, flags=W.FLAGS_ALL, params=params)
    noise = fp64_noise(ctx, d, params, ref)
    mean, spread, outside = (t.cpu().numpy() for t in ens.infer(ctx, d))
    out = (mean, spread, outside, ref.cpu().numpy(), noise.cpu().numpy())
    del d, ref, noise
    return out
