"""Helpers shared by the GPU tests."""
import sys

import numpy as np


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_err(got, ref):
    """Elementwise relative error; 0 where both are equal (incl. both zero,
    both NaN, or both below 1e-290), inf where exactly one is NaN or one is
    zero and the other is not tiny."""
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    both_nan = np.isnan(got) & np.isnan(ref)
    # values that both underflowed below 1e-290 carry no digits to compare
    tiny = (np.abs(got) < 1e-290) & (np.abs(ref) < 1e-290)
    eq = (got == ref) | both_nan | tiny
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.abs(got - ref) / np.abs(ref)
    r = np.where(eq, 0.0, r)
    r = np.where(np.isnan(r), np.inf, r)
    return r


def report(tag, got, ref, names=None):
    r = rel_err(got, ref)
    per = r.max(axis=1) if r.ndim == 2 else r
    say(f"[{tag}] max rel err {r.max():.3e}  median {np.median(r):.3e}")
    if r.max() > 0:
        loc = np.unravel_index(np.argmax(r), r.shape)
        say(f"   worst at {loc}: got {np.asarray(got)[loc]!r} ref {np.asarray(ref)[loc]!r}")
    if names is not None and r.ndim == 2:
        worst = np.argsort(per)[-3:][::-1]
        say("   worst outputs: " + ", ".join(f"{names[i]}={per[i]:.2e}" for i in worst))
    return r


def check_parity(tag, got, ref, mag, names=None, rtol=1e-12, k=None):
    """Condition-aware parity: |got - ref| <= rtol |ref| + k U m, m the
    running error magnitude of the oracle (oracle/error_bound.cpp).  Prints
    the excess and the plain relative error over the well-conditioned
    outputs (m <= 1e3 |ref|); returns the largest excess (pass: <= 1)."""
    import oracle as O
    ex = O.parity_excess(got, ref, mag, rtol=rtol, k=k)
    r = rel_err(got, ref)
    well = np.abs(mag) <= 1e3 * np.abs(ref)
    rw = r[well] if well.any() else np.zeros(1)
    say(f"[{tag}] excess max {ex.max():.3e} | rel err (well-conditioned, {well.mean():.1%} of "
        f"outputs) max {rw.max():.3e} median {np.median(rw):.3e} | rel err all max {r.max():.3e}")
    if ex.max() > 0:
        loc = np.unravel_index(np.argmax(ex), ex.shape)
        nm = names[loc[0]] if names is not None and len(loc) == 2 else loc
        say(f"   worst excess at {nm} point {loc[-1]}: got {np.asarray(got)[loc]!r} "
            f"ref {np.asarray(ref)[loc]!r} m {np.asarray(mag)[loc]!r}")
    return float(ex.max())


def distinct_heavy_species(pts, seed):
    """A copy of a point block whose anux has its own moments (the workloads
    give anux those of nux; the kernels copy the nux distribution values when
    the two agree, so this exercises the other branch)."""
    from bns_nurates_amd import workloads as W
    p = np.array(pts, copy=True)
    rng = np.random.default_rng(seed)
    npts = p.shape[1]
    p[W.P_N0 + 3] *= np.exp(rng.uniform(-1.0, 1.0, npts))
    p[W.P_J0 + 3] *= np.exp(rng.uniform(-1.0, 1.0, npts))
    p[W.P_CHI0 + 3] = rng.uniform(1.0 / 3.0, 1.0, npts)
    return p
