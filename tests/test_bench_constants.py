"""The algorithmic FLOP figures bench.py reports per point agree with the
counting build of the oracle on the bench workload (128^3-grid distributions,
n = 16)."""
import numpy as np
import pytest

import bench
import oracle as O
from bns_nurates_amd import workloads as W


@pytest.mark.skipif(not O.available("count"), reason="make -C oracle first")
@pytest.mark.parametrize("brem,name", [(W.BREM_BRT06, "brt06"), (W.BREM_HR98, "hr98")])
def test_flops_per_point(brem, name):
    pts = np.ascontiguousarray(W.make_points(1024, seed=1)[:, :256])
    c, _ = O.count_m1_opacities(pts, 16, W.FLAGS_ALL, W.params_default_all(brem))
    per = sum(c[k] for k in ("add", "mul", "div", "sqrt")) / pts.shape[1]
    assert abs(per / bench.FLOPS_PER_POINT[name] - 1.0) < 0.02, per
    assert abs(c["exp"] / pts.shape[1] / bench.TRANSCENDENTALS_PER_POINT[name]["exp"] - 1) < 0.05


@pytest.mark.skipif(not O.available("count"), reason="make -C oracle first")
def test_spectral_flops_per_eval():
    pts = np.ascontiguousarray(W.make_points(bench.SPECTRAL_NPTS, seed=1)[:, ::512])
    c, _ = O.count_spectral_opacities(pts, bench.SPECTRAL_NU, bench.SPECTRAL_NQUAD, W.FLAGS_ALL,
                                      W.params_default_all(W.BREM_BRT06), stim=0)
    nev = pts.shape[1] * len(bench.SPECTRAL_NU)
    per = sum(c[k] for k in ("add", "mul", "div", "sqrt")) / nev
    assert abs(per / bench.SPECTRAL_FLOPS_PER_EVAL - 1.0) < 0.02, per
    assert abs(c["exp"] / nev / bench.SPECTRAL_TRANSCENDENTALS_PER_EVAL["exp"] - 1) < 0.05
