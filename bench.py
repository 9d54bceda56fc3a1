"""Throughput of the MI355X gray M1 opacity path on the 128^3-point grid workload.

One step = one pass of ComputeM1Opacities (all reactions, BRT06
bremsstrahlung, Gauss-Legendre n = 16 -> 32 energy nodes) over the 128^3 =
2,097,152 seeded points of the workload, inputs already resident in HBM.
Prints one JSON line (the driver's contract):

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Beside the points/s it reports the roofline of the fused kernel (algorithmic
FP64 FLOPs per launch / kernel time against the FP64 vector peak), the HR98
variant, the mixed-precision form, the spectral path
(ComputeSpectralOpacitiesNotStimulatedAbs over 32^3 points x 100 energies,
with its own roofline and CPU baseline), the serving sweep, the learned form,
and the CPU baseline: the CPU oracle timed on the host cores on a bounded
sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "M1-opacity points/s, FP64 plain vs mixed/surrogate at stated rel err; p99 latency"

# Algorithmic work per point of the gray path at n = 16, all reactions, from
# the FLOP-counting build of the CPU restatement (oracle/flop_count.cpp) on the
# 128^3-grid workload (+, -, *, /, sqrt = 1 FLOP each; exp/log/pow/atan/cbrt
# tallied apart).  tests/test_bench_constants.py re-derives them.
FLOPS_PER_POINT = {"brt06": 1114336.0, "hr98": 1361608.0}
TRANSCENDENTALS_PER_POINT = {"brt06": {"exp": 19309.3, "pow": 5528.0, "log": 197.7},
                             "hr98": {"exp": 23201.3, "pow": 38264.0, "log": 1.7,
                                      "atan": 3168.0, "cbrt": 528.0}}
PEAK_FP64_VECTOR_TFLOPS = 78.6  # MI355X datasheet, FP64 vector (FMA = 2)

# Measured with rocprofv3 on k_gray_m1_f64 at n = 16, BRT06
# (profiles/k_gray_m1_f64_grid128_n16_brt06.md):
# HBM bytes per point per launch, FETCH_SIZE (x2, calibrated on a 1 GiB copy)
# + WRITE_SIZE (x1) over the 2,097,152-point launch ...
TRAFFIC_BYTES_PER_POINT = (346834.0 + 327680) * 1024 / 2097152
# ... against the algorithmic 21 input + 20 output doubles per point
ALGORITHMIC_BYTES_PER_POINT = 8 * (21 + 20)
# hardware FP64 FLOPs per point, (ADD + MUL + TRANS + 2 FMA) x 64 lanes from
# the SQ_INSTS_VALU_*_F64 counters over the 2,097,152-point launch
HW_FLOPS_PER_POINT = 1399451.8


# The spectral path (ComputeSpectralOpacitiesNotStimulatedAbs): 32^3 points
# of the grid workload x the reference spectral test's 100 energies
# 10^(-1 + 4k/100) MeV (tests/tests.hpp:882-894), Gauss-Legendre n = 10, all
# reactions, BRT06, default_all corrections.  Algorithmic work per (point,
# energy) evaluation from the counting build of the reuse-aware restatement
# on this workload (tests/test_bench_constants.py re-derives it).
SPECTRAL_NPTS = 32 ** 3
SPECTRAL_NQUAD = 10
SPECTRAL_NU = 10.0 ** (-1.0 + 4.0 * np.arange(100) / 100.0)
SPECTRAL_FLOPS_PER_EVAL = 25387.1
SPECTRAL_TRANSCENDENTALS_PER_EVAL = {"exp": 440.5, "pow": 109.9, "log": 7.6}


# outputs whose accuracy the reduced-precision form reports: number and
# energy emissivities, absorption and scattering coefficients of nue, anue, nux
MIXED_ROWS = [0, 1, 2, 8, 9, 10, 4, 5, 6, 12, 13, 14, 16, 17, 18]
MIXED_RTOL = 1e-5


def mixed_error_stats(got, ref, mag):
    """Accuracy of a reduced-precision result against the FP64 oracle.

    Per output: [max, median] relative error over the points whose FP64
    value is itself determined to 1e-7 (64 u m <= 1e-7 |ref|, m the running
    error magnitude) and the fraction of such points; over every point, the
    worst excess over the bar |got - ref| <= 1e-5 |ref| + 64 u m (<= 1
    passes)."""
    import oracle as O
    from bns_nurates_amd import workloads as W
    ex = O.parity_excess(got, ref, mag, rtol=MIXED_RTOL)
    rel, det_frac = {}, {}
    for r in MIXED_ROWS:
        g, f, m = got[r], ref[r], mag[r]
        det = (64 * O.U * np.abs(m) <= 1e-7 * np.abs(f)) & (f != 0)
        e = np.abs(g[det] - f[det]) / np.abs(f[det]) if det.any() else np.zeros(1)
        rel[W.OUT_NAMES[r]] = [float(f"{e.max():.3g}"), float(f"{np.median(e):.3g}")]
        det_frac[W.OUT_NAMES[r]] = float(f"{det.mean():.3g}")
    return {"bar_excess_max": float(f"{ex[MIXED_ROWS].max():.3g}"), "rel_err_max_median": rel,
            "fp64_determined_frac": det_frac}


def reference_float_comparison(ctx, sp, params, nquad=10):
    """The reference's own all-float build (REAL_TYPE=float, oracle/_ref)
    and the mixed form on the same points at n = 10 (the reference's largest
    order), both against the FP64 oracle: the upper bound of all-FP32 error
    the mixed form is set against."""
    import torch
    import oracle as O
    from bns_nurates_amd import api
    from bns_nurates_amd import workloads as W
    ref, mag = O.bound_m1_opacities(sp, nquad, W.FLAGS_ALL, params)
    f32 = O.m1_opacities(sp, nquad, W.FLAGS_ALL, params, kind="ref_f32")
    ctx.set_quadrature(nquad)
    mix = ctx.compute_m1_opacities(torch.from_numpy(sp).cuda(), flags=W.FLAGS_ALL, params=params,
                                   precision=api.MIXED)
    torch.cuda.synchronize()
    mix = mix.cpu().numpy()

    def summary(got):
        g, f = got[MIXED_ROWS], ref[MIXED_ROWS]
        det = (64 * O.U * np.abs(mag[MIXED_ROWS]) <= 1e-7 * np.abs(f)) & (f != 0)
        with np.errstate(divide="ignore", invalid="ignore"):
            e = np.abs(g - f) / np.abs(f)
        e = e[det]
        fin = e[np.isfinite(e)]
        ex = O.parity_excess(got, ref, mag, rtol=MIXED_RTOL)[MIXED_ROWS]
        # the maximum over finite errors; non-finite outputs are counted apart
        return {"rel_err_max_finite": float(f"{fin.max():.3g}"),
                "rel_err_median": float(f"{np.median(fin):.3g}"),
                "nonfinite_frac": float(f"{(~np.isfinite(g)).mean():.3g}"),
                "bar_1e-5_fail_frac": float(f"{(ex > 1).mean():.3g}")}
    return {"sample": f"first {sp.shape[1]} points of the grid workload, n={nquad}, BRT06",
            "reference_all_float": summary(f32), "mixed": summary(mix)}


# request mix of the serving workload: meshblock sizes and weights, share of
# single-point calls, quadrature orders, precisions
SERVE_SIZES = [8 ** 3, 16 ** 3, 32 ** 3, 64 ** 3]
SERVE_WEIGHTS = [0.4, 0.3, 0.2, 0.1]
SERVE_SINGLE = 0.05
SERVE_ORDERS = [8, 10, 16]
SERVE_CHUNK = 32768  # a 64^3 request runs as 8 pieces interleaved with the others


def serving_sweep(templates=96, fractions=(0.3, 0.6, 0.9), seconds=2.0, seed=3):
    """Poisson arrivals of opacity requests at several rates (fractions of
    the closed-loop capacity) through the continuous-batching server; p50/p99
    latency and served points/s per rate."""
    import threading
    from collections import deque
    from bns_nurates_amd import api
    from bns_nurates_amd import workloads as W
    rng = np.random.default_rng(seed)
    # request templates: own inputs (slices of one seeded pool), pinned
    sizes = [1 if rng.random() < SERVE_SINGLE else int(rng.choice(SERVE_SIZES, p=SERVE_WEIGHTS))
             for _ in range(templates)]
    pool = W.make_points(max(sizes) * 2 + sum(sizes) // 4, seed=seed)
    tpl = []
    for i, n in enumerate(sizes):
        a = api.PinnedBuffer((21, n))
        o = api.PinnedBuffer((20, n))
        off = int(rng.integers(0, pool.shape[1] - n + 1))
        a.array[...] = pool[:, off:off + n]
        tpl.append(dict(n=n, pts=a, out=o, nquad=int(rng.choice(SERVE_ORDERS)),
                        precision=int(rng.integers(0, 2))))
    prm = W.params_default_all(W.BREM_BRT06)
    srv = api.Server(0, hbm_budget_bytes=4 << 30, streams=4, max_batch_points=1 << 20,
                     chunk_points=SERVE_CHUNK)

    def submit(t):
        return srv.submit(t["pts"].array, t["out"].array, t["nquad"], W.FLAGS_ALL, prm,
                          precision=t["precision"])

    # warm-up (table builds, first launches) and closed-loop capacity
    for t in tpl:
        srv.wait(submit(t))
    t0 = time.perf_counter()
    for _ in range(2):
        ids = [submit(t) for t in tpl]
        for i in ids:
            srv.wait(i)
    cap_pts = 2 * sum(sizes) / (time.perf_counter() - t0)
    mean_pts = sum(sizes) / len(sizes)
    cap_rps = cap_pts / mean_pts
    log(f"serving: closed-loop capacity {cap_pts:.3g} points/s = {cap_rps:.0f} requests/s")
    rows = []
    for frac in fractions:
        rate = frac * cap_rps
        free = deque(range(len(tpl)))
        inflight = deque()
        lock = threading.Condition()
        lat, served = [], [0]
        stop = [False]
        b0 = srv.stats()["batches"]

        def waiter():
            while True:
                with lock:
                    while not inflight and not stop[0]:
                        lock.wait()
                    if not inflight and stop[0]:
                        return
                    rid, k = inflight.popleft()
                lat.append((srv.wait(rid), tpl[k]["n"]))
                with lock:
                    served[0] += tpl[k]["n"]
                    free.append(k)
                    lock.notify_all()

        th = threading.Thread(target=waiter)
        th.start()
        t_start = time.perf_counter()
        t_next = t_start
        n_sub = 0
        while t_next - t_start < seconds:
            now = time.perf_counter()
            if now < t_next:
                time.sleep(min(t_next - now, 1e-3))
                continue
            with lock:
                while not free:
                    lock.wait()
                k = free.popleft()
                inflight.append((submit(tpl[k]), k))
                lock.notify_all()
            n_sub += 1
            t_next += rng.exponential(1.0 / rate)
        with lock:
            stop[0] = True
            lock.notify_all()
        th.join()
        wall = time.perf_counter() - t_start
        lat_ms = np.array([x for x, _ in lat]) * 1e3
        size = np.array([n for _, n in lat])
        nb = srv.stats()["batches"] - b0
        per_class = {}
        for name, lo, hi in (("single", 1, 1), ("<=16^3", 2, 16 ** 3), ("32^3", 16 ** 3 + 1, 32 ** 3),
                             ("64^3", 32 ** 3 + 1, 64 ** 3)):
            sel = (size >= lo) & (size <= hi)
            if sel.any():
                per_class[name] = {"requests": int(sel.sum()),
                                   "p50_ms": float(f"{np.percentile(lat_ms[sel], 50):.3g}"),
                                   "p99_ms": float(f"{np.percentile(lat_ms[sel], 99):.3g}")}
        rows.append({"arrival_rate_rps": round(rate, 1), "requests": n_sub,
                     "p50_ms": float(f"{np.percentile(lat_ms, 50):.3g}"),
                     "p99_ms": float(f"{np.percentile(lat_ms, 99):.3g}"),
                     "by_size": per_class,
                     "points_per_s": float(f"{served[0] / wall:.4g}"),
                     "requests_per_batch": float(f"{n_sub / max(nb, 1):.3g}")})
        log(f"serving at {rate:.0f} req/s: {rows[-1]}")
    st = srv.stats()
    srv.close()
    for t in tpl:
        t["pts"].close()
        t["out"].close()
    return {"workload": "Poisson arrivals; meshblocks 8^3/16^3/32^3/64^3 (0.4/0.3/0.2/0.1) "
                        "+ 5% single points; n in {8,10,16}; FP64 or mixed per request; "
                        "host inputs in pinned memory (PCIe copies included in latency); "
                        f"requests above {SERVE_CHUNK} points as interleaved pieces, smallest "
                        "request first",
            "capacity_points_per_s": float(f"{cap_pts:.4g}"), "sweep": rows,
            "server": {"streams": 4, "hbm_budget_bytes": st["arena_bytes"],
                       "max_inflight_bytes": st["max_inflight_bytes"],
                       "max_concurrent_batches": st["max_concurrent_batches"]}}


def surrogate_eval(ctx, dpts, ref, fp64_ms, params, nquad, train_points, epochs, tols,
                   ln_loss=False, steps=3, coverage=0.999, judge_epochs=16, cal_frac=0.5):
    """The learned form: an ensemble of 4 MLPs trained on the GPU (own
    kernels) on train_points seeded points (seed 1000, disjoint from the
    evaluation grid) with FP64-path targets at order nquad; for each relative
    tolerance its acceptance rule is calibrated on 2 x 2^20 further seeded
    points (seed 2000) and the learned form is run on the workload points
    with per-point fallback to the FP64 path; accuracy judged against the
    FP64 outputs ref (the 128^3 grid, held out from training and
    calibration)."""
    import torch
    from bns_nurates_amd import surrogate as S
    from bns_nurates_amd import workloads as W
    t0 = time.perf_counter()
    ens = S.train_ensemble(ctx, train_points, nquad, params, members=4, epochs=epochs,
                           batch=4096, seed=1000, ln_loss=ln_loss, log=log)
    train_s = time.perf_counter() - t0
    npts = dpts.shape[1]
    r = ref.cpu().numpy()
    rows = S.OUT_ROWS
    res = {"members": 4,
           "network": "15 -> 4 x 256 SiLU -> 15, BF16 MFMA, FP32 accumulation; features "
                      "log nb, log T, Ye, (log n, log J, chi) x 3 species, mu_e/T, "
                      "(mu_n-mu_p)/T, 1/T (15 where the workload names 10: the three "
                      "degeneracy features are inputs of the opacity function), each as "
                      "BF16 high + low parts",
           "train_points": int(ens.n_train), "epochs": epochs, "train_wall_s": round(train_s, 1),
           "loss_units": "natural log" if ln_loss else "standardised per output",
           "final_epoch_loss": [float(f"{h[-1]:.4g}") for h in ens.history],
           "fp64_kernel_ms": fp64_ms, "coverage_target": coverage, "tolerances": []}
    out = torch.empty_like(ref)
    # the network alone, judged by the FP64 path: every point inside the
    # training box and domain accepted (predicted error 0); the fraction of
    # all points whose 15 predicted outputs are all within a relative error
    ctx.set_quadrature(nquad)
    ens.set_acceptance(np.zeros(S.OUT), np.zeros(S.OUT))
    ens.m1_opacities(ctx, dpts, out, 1.0, W.FLAGS_ALL, params)
    o = out.cpu().numpy()
    pred = ~np.all(o == r, axis=0)
    with np.errstate(divide="ignore", invalid="ignore"):
        worst = np.max(np.where(o[rows] == r[rows], 0.0,
                                np.abs(o[rows] - r[rows]) / np.abs(r[rows])), axis=0)
    res["network_alone"] = {
        "predicted_fraction": round(float(pred.mean()), 4),
        "all_outputs_within": {str(tol): round(float((pred & (worst <= tol)).mean()), 4)
                               for tol in (1e-2, 3e-2, 1e-1)}}
    log(f"surrogate network alone: {res['network_alone']}")
    # FP64 noise estimate of the grid outputs (the bar excuses an output only
    # as far as the FP64 path's own rounding noise goes)
    ctx.set_quadrature(nquad)
    noise = S.fp64_noise(ctx, dpts, params, ref).cpu().numpy()
    ncal = 1 << 22
    caldata = S.calibration_data(ens, ctx, params, nquad, ncal, seed=2000)
    calpts = W.make_points(ncal, seed=2000)
    import oracle as O
    for tol in tols:
        # the ensemble-spread rule on the first 2^21 calibration points, for
        # comparison; then the judge (trained and calibrated on all 2^22),
        # which the learned form runs with
        half = tuple(x[..., :ncal // 2] for x in caldata)
        spread_rule = S.calibrate(ens, ctx, tol, params, nquad, coverage=coverage, log=log,
                                  data=half)
        cal = S.calibrate_judge(ens, ctx, tol, params, nquad, data=caldata, pts=calpts,
                                coverage=coverage, epochs=judge_epochs, log=log,
                                cal_frac=cal_frac)
        cal["spread_rule_split_B"] = spread_rule["split_B"]
        bound = float(np.log1p(tol))
        ctx.set_quadrature(nquad)
        nfb = ens.m1_opacities(ctx, dpts, out, bound, W.FLAGS_ALL, params)  # warm-up
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            nfb = ens.m1_opacities(ctx, dpts, out, bound, W.FLAGS_ALL, params)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t1) * 1e3 / steps
        o = out.cpu().numpy()
        acc = ~np.all(o == r, axis=0)
        with np.errstate(divide="ignore", invalid="ignore"):
            rel = np.where(r[rows] != 0, np.abs(o[rows] - r[rows]) / np.abs(r[rows]),
                           np.where(o[rows] == 0, 0.0, np.inf))
        nacc = int(acc.sum())
        nviol = int((acc & np.any(S.violations(o, r, noise, tol), axis=0)).sum())
        nplain = int((acc & np.any(rel > tol, axis=0)).sum())
        per = {W.OUT_NAMES[q]: [float(f"{rel[i, acc].max():.3g}"),
                                float(f"{np.median(rel[i, acc]):.3g}")]
               for i, q in enumerate(rows)} if nacc else {}
        # the oracle's condition-aware bar on up to 2048 accepted points
        orc = None
        if nacc:
            idx = np.nonzero(acc)[0]
            idx = idx[np.linspace(0, idx.size - 1, min(2048, idx.size)).astype(np.int64)]
            pts_h = dpts[:, idx].cpu().numpy()
            oref, mag = O.bound_m1_opacities(np.ascontiguousarray(pts_h), nquad, W.FLAGS_ALL,
                                             params, nthreads=cpu_threads())
            bad_m = (np.abs(o[rows][:, idx] - oref[rows]) >
                     tol * np.abs(oref[rows]) + O.K_FP64 * O.U * np.abs(mag[rows]))
            nbm = int(np.any(bad_m, axis=0).sum())
            orc = {"points": int(idx.size), "beyond": nbm,
                   "bar": "|pred - ref| <= tol |ref| + 2 u m (oracle running error m)"}
        res["tolerances"].append({
            "tol": tol, "error_bound_log": bound, "calibration": cal,
            "fallback_fraction": round(nfb / npts, 5), "accepted_points": nacc,
            "bar": cal["bar"],
            "accepted_beyond_bar": nviol,
            "violation_rate": float(f"{nviol / max(nacc, 1):.3g}"),
            "violation_rate_ucb95": float(f"{S.clopper_pearson_upper(nviol, nacc):.3g}"),
            "accepted_beyond_tol_ignoring_fp64_noise": nplain,
            "oracle_sample": orc,
            "ms_per_pass": round(ms, 2), "points_per_s": float(f"{npts / (ms * 1e-3):.4g}"),
            "speedup_vs_fp64": round(fp64_ms / ms, 3),
            "accepted_rel_err_max_median": per})
        log(f"surrogate tol={tol}: fallback {nfb / npts:.2%}, {nviol} of {nacc} accepted beyond "
            f"the bar ({nplain} beyond tol alone), oracle sample {orc}, {ms:.1f} ms/pass "
            f"({fp64_ms / ms:.2f}x FP64)")
    ens.close()
    return res


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def config0_eval(ctx, steps=20, repeats=5):
    """The reference's own CPU-runnable case (4096 seeded points, n = 10,
    BRT06, all reactions): GPU kernel time, and the genuine reference
    (oracle/_ref, compiled from the reference's headers, OpenMP over points)
    on the host cores over the same points, with the parity of the two."""
    import torch
    import oracle as O
    from bns_nurates_amd import workloads as W
    cfg = W.WORKLOADS["pts4096_n10_brt06"]
    pts = W.make_points(cfg["npts"], seed=cfg["seed"])
    prm = W.params_default_all(cfg["brem"])
    npts = pts.shape[1]
    ctx.set_quadrature(cfg["nquad"])
    dp = torch.from_numpy(pts).cuda()
    out = torch.empty((20, npts), dtype=torch.float64, device="cuda")
    for _ in range(3):
        ctx.compute_m1_opacities(dp, flags=W.FLAGS_ALL, params=prm, out=out)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    t0 = time.perf_counter()
    for a, b in ev:
        a.record()
        ctx.compute_m1_opacities(dp, flags=W.FLAGS_ALL, params=prm, out=out)
        b.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    got = out.cpu().numpy()
    res = {"workload": "pts4096_n10_brt06: 4096 seeded points (seed 0), n=10, all reactions, "
                       "BRT06, default_all, distributions reconstructed from M1",
           "value": npts * steps / wall, "unit": "points/s", "kernel_ms": kern_ms,
           "note": "4096 workgroups do not fill the 256 CUs; the grid workload is the "
                   "throughput figure"}
    if O.available("ref"):
        threads = cpu_threads()
        O.m1_opacities(pts[:, :256], cfg["nquad"], W.FLAGS_ALL, prm, kind="ref", nthreads=threads)
        ts = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            ref = O.m1_opacities(pts, cfg["nquad"], W.FLAGS_ALL, prm, kind="ref",
                                 nthreads=threads)
            ts.append(time.perf_counter() - t0)
        _, mag = O.bound_m1_opacities(pts, cfg["nquad"], W.FLAGS_ALL, prm)
        res["cpu_reference"] = {
            "value": npts / float(np.median(ts)), "unit": "points/s", "cores": threads,
            "kind": "reference",
            "sample": f"all {npts} points, median of {repeats} passes; the reference's own "
                      "ComputeM1Opacities compiled from its headers (oracle/Makefile, -O2 "
                      "-ffp-contract=off, OpenMP over points)"}
        res["gpu_kernel_over_reference"] = (npts / (kern_ms * 1e-3)) / res["cpu_reference"]["value"]
        res["parity_vs_reference"] = {"bar_excess_max": float(O.parity_excess(got, ref, mag).max())}
        log(f"config 0: GPU {npts / (kern_ms * 1e-3):.4g} points/s (kernel), reference "
            f"{res['cpu_reference']['value']:.4g} points/s on {threads} threads")
    else:
        res["cpu_reference"] = None
        log("config 0: oracle/_ref not built here; reference not timed")
    return res


def spectral_eval(ctx, steps, warmup, cpu=True, target_s=8.0, layouts=("struct", "fields"),
                  warmup_s=0.3):
    """The spectral leg: events-timed passes of the spectral kernel over the
    spectral workload (inputs resident in HBM) in each output layout -- one
    SpectralOpacities record per (point, energy) (bnsamd_spectral_opacities_aos)
    and the field-major [16][nnu][npts] block (bnsamd_spectral_opacities) --,
    an oracle parity sample of each, and the CPU restatement timed on a
    bounded prefix of the points.  The first layout is the line's figure."""
    import torch
    import oracle as O
    from bns_nurates_amd import workloads as W
    sp = W.make_points(SPECTRAL_NPTS, seed=1)
    npts, nnu = sp.shape[1], len(SPECTRAL_NU)
    dsp = torch.from_numpy(sp).cuda()
    dnu = torch.from_numpy(SPECTRAL_NU).cuda()
    prm = W.params_default_all(W.BREM_BRT06)
    ctx.set_quadrature(SPECTRAL_NQUAD)
    idx = np.arange(0, npts, npts // 32)[:32]
    ref, mag = O.bound_spectral_opacities(np.ascontiguousarray(sp[:, idx]), SPECTRAL_NU,
                                          SPECTRAL_NQUAD, W.FLAGS_ALL, prm, stim=0)
    nev = npts * nnu
    legs = {}
    for layout in layouts:
        if layout == "struct":
            out = torch.empty((npts, nnu, 16), dtype=torch.float64, device="cuda")
        else:
            out = torch.empty((16 * nnu, npts), dtype=torch.float64, device="cuda")

        def run():
            ctx.compute_spectral_opacities(dsp, dnu, W.FLAGS_ALL, prm, stim_abs=0, out=out,
                                           layout=layout)

        # a launch is ~5 ms: the warm-up runs for warmup_s of wall time (at
        # least `warmup` launches) so that the clocks have left their idle
        # state, as in a loop that keeps the device busy
        t0 = time.perf_counter()
        nw = 0
        while nw < warmup or time.perf_counter() - t0 < warmup_s:
            run()
            nw += 1
            if nw % 8 == 0:
                torch.cuda.synchronize()
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(steps)]
        log(f"spectral ({layout} layout): {nw} warm-up launches, timing {steps} steps "
            f"({npts} points x {nnu} energies)")
        t0 = time.perf_counter()
        for a, b in ev:
            a.record()
            run()
            b.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        kern_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
        o = out.cpu().numpy()
        o = o.transpose(2, 1, 0) if layout == "struct" else o.reshape(16, nnu, npts)
        assert np.isfinite(o).all()
        excess = float(O.parity_excess(o[:, :, idx], ref, mag).max())
        achieved = SPECTRAL_FLOPS_PER_EVAL * nev / (kern_ms * 1e-3) / 1e12
        legs[layout] = {"value": nev * steps / wall, "kernel_ms": kern_ms, "achieved": achieved,
                        "bar_excess_max": excess, "warmup_launches": nw}
        log(f"spectral ({layout}): {legs[layout]['value']:.4g} evals/s, kernel {kern_ms:.2f} ms, "
            f"{achieved:.2f} TFLOP/s, parity excess {excess:.3g}")
        del out
    main = legs[layouts[0]]
    res = {"workload": f"spectral_grid32_e100_n10_brt06: {npts} points of the grid workload x "
                       f"{nnu} energies 10^(-1+4k/100) MeV (the reference spectral test's grid), "
                       "n=10, all reactions, BRT06, default_all, not stimulated, distributions "
                       "reconstructed from M1",
           "layout": {"struct": "SpectralOpacities[npts][nnu] (bnsamd_spectral_opacities_aos)",
                      "fields": "[16][nnu][npts] (bnsamd_spectral_opacities)"}[layouts[0]],
           "value": main["value"], "unit": "(point, energy) evaluations/s",
           "steps": steps, "warmup_launches": main["warmup_launches"],
           "kernel_ms": main["kernel_ms"],
           "roofline": {"bound": "fp64 vector ALU", "achieved": main["achieved"],
                        "peak": PEAK_FP64_VECTOR_TFLOPS, "unit": "TFLOP/s",
                        "frac": main["achieved"] / PEAK_FP64_VECTOR_TFLOPS,
                        "kernel": "k_spectral_f64",
                        "flops_per_eval": SPECTRAL_FLOPS_PER_EVAL,
                        "transcendentals_per_eval": SPECTRAL_TRANSCENDENTALS_PER_EVAL,
                        "algorithmic_bytes_per_eval": 16 * 8 + 8 * 21 / nnu},
           "parity_sample": {"points": len(idx), "energies": nnu,
                             "bar_excess_max": main["bar_excess_max"]},
           "layouts": {k: {"kernel_ms": v["kernel_ms"], "value": v["value"],
                           "bar_excess_max": v["bar_excess_max"]} for k, v in legs.items()}}
    if cpu:
        kind = "fast" if O.available("fast") else "orc"
        threads = cpu_threads()
        probe = np.ascontiguousarray(sp[:, :2 * threads])
        t0 = time.perf_counter()
        O.spectral_opacities(probe, SPECTRAL_NU, SPECTRAL_NQUAD, W.FLAGS_ALL, prm, stim=0,
                             kind=kind, nthreads=threads)
        rate = probe.shape[1] / (time.perf_counter() - t0)
        n = int(min(max(rate * target_s, threads), npts))
        sample = np.ascontiguousarray(sp[:, :n])
        t0 = time.perf_counter()
        O.spectral_opacities(sample, SPECTRAL_NU, SPECTRAL_NQUAD, W.FLAGS_ALL, prm, stim=0,
                             kind=kind, nthreads=threads)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": n * nnu / dt, "unit": "(point, energy) evaluations/s",
                               "cores": threads, "kind": "port",
                               "sample": f"first {n} points x {nnu} energies, {dt:.1f} s wall on "
                                         f"{threads} threads, reuse-aware restatement "
                                         "oracle/m1_oracle.hpp built -O3 -march=x86-64-v3"}
        res["gpu_over_cpu"] = res["value"] / res["cpu_baseline"]["value"]
        log(f"spectral CPU baseline: {res['cpu_baseline']['value']:.4g} evals/s")
    return res


def cpu_threads():
    """Threads for the CPU legs: the CPUs this process may run on, capped by
    OMP_NUM_THREADS when it is set (the host's share for one GPU)."""
    n = len(os.sched_getaffinity(0))
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return min(n, omp) if omp > 0 else n


def host_info():
    """CPU model, SMT state and CPU counts of the host the baseline ran on."""
    model = None
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    smt = None
    try:
        smt = {"1": "on", "0": "off"}.get(open("/sys/devices/system/cpu/smt/active").read().strip())
    except OSError:
        pass
    return {"cpu_model": model, "smt": smt, "affinity_cpus": len(os.sched_getaffinity(0)),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(pts, nquad, params, pass_s=2.0, repeats=5, single_s=4.0):
    """The CPU restatement of the path (FP64, OpenMP over points,
    schedule(dynamic), built with -O3 for AVX2/FMA) on a bounded prefix of
    the workload: the median of `repeats` passes on every thread the process
    may use, after one warm-up pass, and one pass on a single thread.  The
    parallel loop over points is the reference drivers' Kokkos parallel_for
    (tests/tests.hpp:272-274)."""
    import oracle as O
    from bns_nurates_amd import workloads as W
    kind = "fast" if O.available("fast") else "orc"
    threads = cpu_threads()
    probe = np.ascontiguousarray(pts[:, : 64 * threads])
    t0 = time.perf_counter()
    O.m1_opacities(probe, nquad, W.FLAGS_ALL, params, nthreads=threads, kind=kind)
    rate = probe.shape[1] / (time.perf_counter() - t0)
    n = int(min(max(rate * pass_s, 1024), pts.shape[1]))
    sample = np.ascontiguousarray(pts[:, :n])
    ts = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        O.m1_opacities(sample, nquad, W.FLAGS_ALL, params, nthreads=threads, kind=kind)
        ts.append(time.perf_counter() - t0)
    dt = float(np.median(ts))
    n1 = int(max(64, min(rate / threads * single_s, n)))
    one = np.ascontiguousarray(pts[:, :n1])
    t0 = time.perf_counter()
    O.m1_opacities(one, nquad, W.FLAGS_ALL, params, nthreads=1, kind=kind)
    dt1 = time.perf_counter() - t0
    build = ("-O3 -march=x86-64-v3 (FMA)" if kind == "fast"
             else "-O2 -ffp-contract=off (the bit-exact build)")
    return {"value": n / dt, "unit": "points/s", "cores": threads, "kind": "port",
            "single_thread_points_per_s": n1 / dt1,
            "passes_s": [round(t, 3) for t in ts], "host": host_info(),
            "sample": f"first {n} points of the workload, n={nquad}, FP64, median of {repeats} "
                      f"passes ({dt:.2f} s each) on {threads} threads after a warm-up; single "
                      f"thread: first {n1} points; oracle/m1_oracle.hpp built {build}; the "
                      f"same code built -O2 -ffp-contract=off is bit-exact with the reference "
                      f"at n<=10 (the reference itself stops at 2n <= 20)"}


def grid_parity(ctx, pts, out, nquad, params, nsample=2048):
    """Parity of the headline pass's outputs against the FP64 oracle on a
    strided sample of the grid: the worst excess over the bar |gpu - ref| <=
    1e-12 |ref| + 2 u m (m the oracle's running error magnitude), the plain
    relative error over all outputs, and over the outputs whose FP64 value
    the rounding-error bound determines to 1e-11, 1e-10, 1e-9."""
    import oracle as O
    from bns_nurates_amd import workloads as W
    idx = np.linspace(0, pts.shape[1] - 1, nsample).astype(np.int64)
    sp = np.ascontiguousarray(pts[:, idx])
    got = out.cpu().numpy()[:, idx]
    ref, mag = O.bound_m1_opacities(sp, nquad, W.FLAGS_ALL, params, nthreads=cpu_threads())
    ex = O.parity_excess(got, ref, mag)
    with np.errstate(divide="ignore", invalid="ignore"):
        rel = np.where(got == ref, 0.0, np.abs(got - ref) / np.abs(ref))
    tiers = {}
    for cut in (1e-11, 1e-10, 1e-9):
        det = (O.K_FP64 * O.U * np.abs(mag) <= cut * np.abs(ref)) & (ref != 0)
        tiers[f"{cut:g}"] = {"frac": float(f"{det.mean():.3g}"),
                             "rel_err_max": float(f"{rel[det].max():.3g}") if det.any() else None}
    return {"sample": f"{nsample} points strided over the grid, all 20 outputs",
            "bar": "|gpu - ref| <= 1e-12 |ref| + 2 u m",
            "bar_excess_max": float(f"{ex.max():.3g}"),
            "rel_err_max_all_outputs": float(f"{rel.max():.3g}"),
            "frac_outputs_within_1e-12": float(f"{np.mean(rel <= 1e-12):.4g}"),
            "fp64_determined": tiers,
            "note": "fp64_determined[c]: outputs whose FP64 rounding-error bound 2 u m is at most "
                    "c |ref|, with the GPU's largest relative error on them; the other outputs "
                    "are cancellation-dominated, where the bar's 2 u m term applies"}


def host_roundtrip(ctx, pts, params, passes=3):
    """Points/s of the pass from and to host memory (pinned buffers):
    upload of the 21-field block, the kernel, download of the 20 outputs,
    synchronous, as the reference's drivers pay the copies around the loop
    (tests/tests.hpp:139-153)."""
    from bns_nurates_amd import api
    from bns_nurates_amd import workloads as W
    npts = pts.shape[1]
    hp = api.PinnedBuffer((21, npts))
    ho = api.PinnedBuffer((20, npts))
    hp.array[...] = pts
    ctx.compute_m1_opacities(hp.array, flags=W.FLAGS_ALL, params=params, out=ho.array)
    ts = []
    for _ in range(passes):
        t0 = time.perf_counter()
        ctx.compute_m1_opacities(hp.array, flags=W.FLAGS_ALL, params=params, out=ho.array)
        ts.append(time.perf_counter() - t0)
    hp.close()
    ho.close()
    dt = float(np.median(ts))
    return {"value": npts / dt, "unit": "points/s", "ms_per_pass": round(dt * 1e3, 2),
            "passes": passes,
            "what": "bnsamd_m1_opacities_host from pinned host buffers: H2D of 21 x 8 B/point, "
                    "kernel, D2H of 20 x 8 B/point, synchronous (median of passes)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--npts", type=int, default=128 ** 3)
    ap.add_argument("--nquad", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-hr98", action="store_true", help="skip the HR98 passes")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision passes")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving sweep")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the learned form")
    ap.add_argument("--no-spectral", action="store_true", help="skip the spectral leg")
    ap.add_argument("--surrogate-train-points", type=int, default=10_000_000)
    ap.add_argument("--surrogate-epochs", type=int, default=96)
    ap.add_argument("--surrogate-judge-epochs", type=int, default=16,
                    help="training epochs of each tolerance's acceptance judge")
    ap.add_argument("--surrogate-cal-frac", type=float, default=0.75,
                    help="violators the judge's threshold allows on its calibration split, "
                         "as a fraction of 1 - coverage")
    ap.add_argument("--surrogate-std-loss", action="store_true",
                    help="train the learned form on the squared error in per-output "
                         "standardised units instead of natural-log units")
    ap.add_argument("--spectral-layout", choices=("both", "struct", "fields"), default="both",
                    help="output layouts of the spectral leg (struct: SpectralOpacities records; "
                         "fields: [16][nnu][npts]); the first timed is the line's figure")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="before timing, run one device-to-device copy of this many MiB "
                         "(the known-size kernel that scales rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(points shard across GPUs with no exchange; only the one-GPU path is built)", file=sys.stderr)
        sys.exit(2)

    import torch
    from bns_nurates_amd import api
    from bns_nurates_amd import workloads as W

    if not torch.cuda.is_available():
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)
    log(f"generating {args.npts} points (128^3-grid workload distributions, seed 1)")
    pts = W.make_points(args.npts, seed=1)
    dpts = torch.from_numpy(pts).cuda()
    out = torch.empty((W.O_NFIELDS, args.npts), dtype=torch.float64, device="cuda")
    ctx = api.Context(0)
    ctx.set_quadrature(args.nquad)

    if args.copy_mib > 0:
        nel = args.copy_mib * (1 << 20) // 8
        src = torch.ones(nel, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: {args.copy_mib} MiB read + {args.copy_mib} MiB written")
        del src, dst

    def run_pass(params, precision=api.FP64):
        ctx.compute_m1_opacities(dpts, flags=W.FLAGS_ALL, params=params, out=out,
                                 precision=precision)

    # the kernel runs on torch's current stream, where the events are recorded
    res = {}
    passes = [("brt06", api.FP64)]
    if not args.no_hr98:
        passes.append(("hr98", api.FP64))
    if not args.no_mixed:
        passes.append(("brt06_mixed", api.MIXED))
        if not args.no_hr98:
            passes.append(("hr98_mixed", api.MIXED))
    for brem, prec in passes:
        params = W.params_default_all(W.BREM_BRT06 if brem.startswith("brt06") else W.BREM_HR98)
        steps = args.steps if brem == "brt06" else max(1, args.steps // 2)
        for _ in range(args.warmup):
            run_pass(params, prec)
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(steps)]
        log(f"{brem}: timing {steps} steps")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for a, b in ev:
            a.record()
            run_pass(params, prec)
            b.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        kern_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
        res[brem] = dict(wall=wall, steps=steps, kernel_ms=kern_ms,
                         pps=args.npts * steps / wall)
        log(f"{brem}: {res[brem]['pps']:.4g} points/s, kernel {kern_ms:.2f} ms/launch")
        assert np.isfinite(out.cpu().numpy()).all()
        if brem == "brt06":
            out_brt06 = out.clone()  # the headline pass's outputs, for the parity sample

    b = res["brt06"]
    flops = FLOPS_PER_POINT["brt06"] * args.npts if args.nquad == 16 else None
    achieved = flops / (b["kernel_ms"] * 1e-3) / 1e12 if flops else None
    line = {
        "metric": METRIC,
        "value": b["pps"],
        "unit": "points/s",
        "n_gpus": 1,
        "steps": b["steps"],
        "warmup": args.warmup,
        "ms_per_step": 1e3 * b["wall"] / b["steps"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "grid128_n16_brt06", "npts": args.npts,
                   "nquad": args.nquad, "energy_nodes": 2 * args.nquad,
                   "reactions": "beta+pair+brem+nes+iso", "brem": "BRT06",
                   "opacity_params": "default_all", "distribution": "reconstructed from M1"},
        "roofline": {"bound": "fp64 vector ALU", "achieved": achieved, "peak": PEAK_FP64_VECTOR_TFLOPS,
                     "unit": "TFLOP/s",
                     "frac": (achieved / PEAK_FP64_VECTOR_TFLOPS) if achieved else None,
                     "traffic": TRAFFIC_BYTES_PER_POINT * args.npts if args.nquad == 16 else None,
                     "bound_detail": "FP64 compute on the vector ALU (MI355X FP64 vector and "
                                     "matrix peaks are both 78.6 TFLOP/s); HBM traffic "
                                     f"{TRAFFIC_BYTES_PER_POINT:.1f} B/point vs algorithmic "
                                     f"{ALGORITHMIC_BYTES_PER_POINT} B/point",
                     "kernel": "k_gray_m1_f64", "kernel_ms": b["kernel_ms"],
                     "flops_per_point": FLOPS_PER_POINT["brt06"],
                     "hw_flops_per_point": HW_FLOPS_PER_POINT,
                     "hw_achieved": HW_FLOPS_PER_POINT * args.npts / (b["kernel_ms"] * 1e-3) / 1e12,
                     "transcendentals_per_point": TRANSCENDENTALS_PER_POINT["brt06"]},
    }
    log("parity sample of the headline pass against the oracle")
    line["parity"] = grid_parity(ctx, pts, out_brt06, args.nquad, W.params_default_all(W.BREM_BRT06))
    del out_brt06
    log(f"parity: {line['parity']}")
    line["host_roundtrip"] = host_roundtrip(ctx, pts, W.params_default_all(W.BREM_BRT06))
    log(f"host round trip: {line['host_roundtrip']}")
    if "hr98" in res:
        h = res["hr98"]
        line["hr98"] = {"value": h["pps"], "unit": "points/s", "kernel_ms": h["kernel_ms"],
                        "achieved_tflops": FLOPS_PER_POINT["hr98"] * args.npts /
                        (h["kernel_ms"] * 1e-3) / 1e12 if args.nquad == 16 else None}
    for brem in ("brt06", "hr98"):
        mk = brem + "_mixed"
        if mk not in res:
            continue
        h = res[mk]
        params = W.params_default_all(W.BREM_BRT06 if brem == "brt06" else W.BREM_HR98)
        # accuracy against the FP64 oracle: the first 2048 points of the
        # workload and 512 cold degenerate tail points (T < 1 MeV, eta_e > 50)
        import oracle as O
        if not O.available("bound"):
            O.build()
        acc = {}
        for tag, sp in (("sample", np.ascontiguousarray(pts[:, :2048])),
                        ("tail_T<1_eta_e>50", W.make_tail_points(512, seed=5))):
            ctx.set_quadrature(args.nquad)
            got = ctx.compute_m1_opacities(torch.from_numpy(sp).cuda(), flags=W.FLAGS_ALL,
                                           params=params, precision=api.MIXED)
            torch.cuda.synchronize()
            ref, mag = O.bound_m1_opacities(sp, args.nquad, W.FLAGS_ALL, params)
            acc[tag] = mixed_error_stats(got.cpu().numpy(), ref, mag)
        line.setdefault("mixed", {})[brem] = {
            "value": h["pps"], "unit": "points/s", "kernel_ms": h["kernel_ms"],
            "speedup_vs_fp64": res[brem]["kernel_ms"] / h["kernel_ms"] if brem in res else None,
            "fp32_parts": "distribution values (exp/log), HR98 channel functions",
            "rel_err_vs_fp64_oracle": acc}
        if brem == "brt06" and O.available("ref_f32"):
            line["mixed"]["vs_reference_all_float"] = reference_float_comparison(
                ctx, np.ascontiguousarray(pts[:, :2048]), params)
    if not args.no_cpu:
        line["config0"] = config0_eval(ctx)
    if not args.no_spectral:
        lay = {"both": ("struct", "fields"), "struct": ("struct",),
               "fields": ("fields",)}[args.spectral_layout]
        line["spectral"] = spectral_eval(ctx, max(10, 2 * args.steps), max(1, args.warmup),
                                         cpu=not args.no_cpu, layouts=lay)
    if not args.no_serving:
        line["serving"] = serving_sweep()
    if not args.no_surrogate:
        params = W.params_default_all(W.BREM_BRT06)
        ctx.set_quadrature(args.nquad)
        ctx.compute_m1_opacities(dpts, flags=W.FLAGS_ALL, params=params, out=out)
        ref = out.clone()
        line["surrogate"] = surrogate_eval(ctx, dpts, ref, b["kernel_ms"], params, args.nquad,
                                           args.surrogate_train_points, args.surrogate_epochs,
                                           tols=(1e-2, 3e-2, 1e-1),
                                           ln_loss=not args.surrogate_std_loss,
                                           judge_epochs=args.surrogate_judge_epochs,
                                           cal_frac=args.surrogate_cal_frac)
    if not args.no_cpu:
        log("CPU baseline (oracle) on a bounded sample")
        line["cpu_baseline"] = cpu_baseline(pts, args.nquad, W.params_default_all(W.BREM_BRT06))
        line["gpu_over_cpu"] = b["pps"] / line["cpu_baseline"]["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
