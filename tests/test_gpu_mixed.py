"""The mixed-precision form (BNSAMD_MIXED) against the FP64 oracle.

Bar per output: |mixed - ref| <= 1e-5 |ref| + 64 u m, i.e. a relative error
of at most 1e-5 on top of what the FP64 value itself leaves undetermined
(m: running error magnitude, oracle/error_bound.cpp), on the 128^3-grid
distributions, the cold degenerate tail (T < 1 MeV, eta_e > 50) and every
flag/parameter variant of the reference fixtures.
"""
import numpy as np
import pytest

import bench
import oracle as O
from bns_nurates_amd import workloads as W
from tests.gpu_util import say

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

MIXED = 1


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from bns_nurates_amd import api
    c = api.Context(0)
    yield c
    c.close()


def run(ctx, pts, nquad, flags, params, stim=1, precision=MIXED):
    ctx.set_quadrature(nquad)
    out = ctx.compute_m1_opacities(torch.from_numpy(np.ascontiguousarray(pts)).cuda(),
                                   flags=flags, params=params, stim_abs=stim,
                                   precision=precision)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def check(tag, got, ref, mag):
    ex = O.parity_excess(got, ref, mag, rtol=bench.MIXED_RTOL)
    st = bench.mixed_error_stats(got, ref, mag)
    worst = max(v[0] for v in st["rel_err_max_median"].values())
    say(f"[{tag}] bar excess {ex.max():.3e}; max rel err over FP64-determined outputs {worst:.2e}")
    if ex.max() > 1:
        loc = np.unravel_index(np.argmax(ex), ex.shape)
        say(f"   worst at {W.OUT_NAMES[loc[0]]} point {loc[1]}: got {got[loc]!r} "
            f"ref {ref[loc]!r} m {mag[loc]!r}")
    return float(ex.max())


@pytest.mark.parametrize("brem", [W.BREM_BRT06, W.BREM_HR98])
def test_mixed_grid_distributions_n16(ctx, brem):
    pts = W.make_points(1024, seed=11)
    prm = W.params_default_all(brem)
    say(f"mixed n=16 brem={brem}: 1024 points")
    got = run(ctx, pts, 16, W.FLAGS_ALL, prm)
    ref, m = O.bound_m1_opacities(pts, 16, W.FLAGS_ALL, prm)
    assert check(f"mixed n16 brem{brem}", got, ref, m) <= 1.0


def test_mixed_cold_degenerate_tail(ctx):
    pts = W.make_tail_points(512, seed=5)
    for brem in (W.BREM_BRT06, W.BREM_HR98):
        prm = W.params_default_all(brem)
        say(f"mixed tail brem={brem}: 512 points, T<1 MeV, eta_e>50")
        got = run(ctx, pts, 16, W.FLAGS_ALL, prm)
        ref, m = O.bound_m1_opacities(pts, 16, W.FLAGS_ALL, prm)
        assert check(f"mixed tail brem{brem}", got, ref, m) <= 1.0


def test_mixed_variants_vs_reference(ctx, golden):
    """Every flag set, blocking on/off, non-stimulated, both brem models,
    against the reference's own outputs."""
    g = golden("m1_variants.npz")
    tags = [k[4:] for k in g.files if k.startswith("out_") and k != "out_equilibrium_n10"]
    worst = 0.0
    for t in tags:
        say(f"mixed variant {t}")
        flags, params = g[f"flags_{t}"], g[f"params_{t}"]
        nq, stim = int(g[f"nquad_{t}"]), int(g[f"stim_{t}"])
        got = run(ctx, g["pts"], nq, flags, params, stim=stim)
        _, m = O.bound_m1_opacities(g["pts"], nq, flags, params, stim=stim)
        worst = max(worst, check(f"mixed {t}", got, g[f"out_{t}"], m))
    assert worst <= 1.0
