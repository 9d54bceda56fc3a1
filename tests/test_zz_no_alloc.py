"""bnsamd_m1_opacities_host_static, the host entry point that allocates
nothing (it stages through the library's static device block), against
bnsamd_m1_opacities_host on the same inputs, through the C ABI: two calls
in a row, the second smaller and with a caller-supplied distribution block,
bit-identical outputs; and the size limit rejected before any work."""
import ctypes as C

import numpy as np
import pytest

from bns_nurates_amd import api
from bns_nurates_amd import workloads as W
from tests.gpu_util import say

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    c = api.Context(0)
    c.set_quadrature(10)
    yield c
    c.close()


def _call(lib, name, ctx, pts, distr, flags, params, stim, prec):
    n = pts.shape[1]
    out = np.full((api.OUT_NFIELDS, n), np.nan)
    rc = getattr(lib, name)(ctx.h, n, pts.ctypes.data, distr.ctypes.data if distr is not None
                            else None, C.byref(flags), C.byref(params), stim, prec,
                            out.ctypes.data)
    return rc, out


def test_static_entry_matches_the_host_entry(ctx):
    lib = api.load_library()
    fl = api.as_flags(W.FLAGS_ALL).c()
    cases = [
        # (npts, seed, with distribution block, params, stim, precision)
        (16384, 11, False, W.params_default_all(W.BREM_BRT06), 1, api.FP64),
        (3001, 12, True, W.params_default_all(W.BREM_HR98), 0, api.FP64),
        (777, 13, False, W.params_default_all(W.BREM_BRT06), 1, api.MIXED),
    ]
    for npts, seed, with_distr, prm, stim, prec in cases:
        say(f"static entry: {npts} points, distr={with_distr}, stim={stim}, precision={prec}")
        pts = np.ascontiguousarray(W.make_points(npts, seed=seed))
        distr = None
        if with_distr:
            distr = np.empty((api.DI_NFIELDS, npts))
            assert lib.bnsamd_distr_from_m1_host(ctx.h, npts, pts.ctypes.data,
                                                 distr.ctypes.data) == 0
            distr[api.DI_NFIELDS - 4:] *= 1.01  # not the reconstruction's own values
        pr = api.as_params(prm).c()
        rc1, got = _call(lib, "bnsamd_m1_opacities_host_static", ctx, pts, distr, fl, pr, stim,
                         prec)
        rc2, ref = _call(lib, "bnsamd_m1_opacities_host", ctx, pts, distr, fl, pr, stim, prec)
        assert rc1 == 0 and rc2 == 0, (rc1, rc2, lib.bnsamd_last_error(ctx.h))
        assert np.isfinite(ref).all()
        assert np.array_equal(got, ref)


def test_static_entry_rejects_too_many_points(ctx):
    lib = api.load_library()
    n = 16385  # BNSAMD_STATIC_MAX_POINTS + 1
    pts = np.zeros((api.PT_NFIELDS, n))
    out = np.zeros((api.OUT_NFIELDS, n))
    fl = api.as_flags(W.FLAGS_ALL).c()
    pr = api.as_params(W.params_default_all(W.BREM_BRT06)).c()
    rc = lib.bnsamd_m1_opacities_host_static(ctx.h, n, pts.ctypes.data, None, C.byref(fl),
                                             C.byref(pr), 1, 0, out.ctypes.data)
    assert rc == 1
    assert b"BNSAMD_STATIC_MAX_POINTS" in lib.bnsamd_last_error(ctx.h)
    assert not out.any()
