"""The C ABI contract (include/bns_nurates_amd.h) of the gray and spectral
entry points: status codes and messages for bad arguments (never an abort),
empty inputs as no-ops, inputs left unmodified, the quadrature limits."""
import ctypes as C

import numpy as np
import pytest

from bns_nurates_amd import api
from bns_nurates_amd import workloads as W

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

OK, INVALID, NO_QUAD = 0, 1, 4


@pytest.fixture(scope="module")
def lib():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return api.load_library()


@pytest.fixture()
def ctx(lib):
    h = C.c_void_p()
    assert lib.bnsamd_create(0, C.byref(h)) == OK
    yield h
    torch.cuda.synchronize()
    lib.bnsamd_destroy(h)


def quad(lib, h, n):
    x, w = api.gauss_legendre(n)
    dp = C.POINTER(C.c_double)
    return lib.bnsamd_set_quadrature(h, n, x.ctypes.data_as(dp), w.ctypes.data_as(dp), n,
                                     x.ctypes.data_as(dp), w.ctypes.data_as(dp))


def message(lib, h):
    return lib.bnsamd_last_error(h).decode()


def gray(lib, h, npts, ld, pts, out, stim=1, prec=api.FP64, flags=None, params=None):
    fl = (flags or api.OpacityFlags()).c()
    pr = (params or api.OpacityParams(brem_implementation=api.BREM_BRT06)).c()
    return lib.bnsamd_m1_opacities(h, npts, ld, pts, None, C.byref(fl), C.byref(pr), stim, prec,
                                   out)


def spectral(lib, h, npts, ld, pts, nnu, nu, nu_ld, out, stim=0, params=None):
    fl = api.OpacityFlags().c()
    pr = (params or api.OpacityParams(brem_implementation=api.BREM_BRT06)).c()
    return lib.bnsamd_spectral_opacities(h, npts, ld, pts, None, nnu, nu, nu_ld, C.byref(fl),
                                         C.byref(pr), stim, out)


def test_quadrature_is_required_and_bounded(lib, ctx):
    p = torch.from_numpy(W.make_points(8, seed=1)).cuda()
    o = torch.empty((20, 8), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()  # the context launches on its own stream
    assert gray(lib, ctx, 8, 8, p.data_ptr(), o.data_ptr()) == NO_QUAD
    assert "quadrature" in message(lib, ctx)
    nu = torch.ones(3, dtype=torch.float64, device="cuda")
    s = torch.empty((48, 8), dtype=torch.float64, device="cuda")
    assert spectral(lib, ctx, 8, 8, p.data_ptr(), 3, nu.data_ptr(), 0, s.data_ptr()) == NO_QUAD
    assert quad(lib, ctx, 33) == INVALID
    x = np.array([0.5, 1.0])  # a node outside (0, 1)
    dp = C.POINTER(C.c_double)
    assert lib.bnsamd_set_quadrature(ctx, 0, x.ctypes.data_as(dp), x.ctypes.data_as(dp), 0,
                                     x.ctypes.data_as(dp), x.ctypes.data_as(dp)) == INVALID
    assert lib.bnsamd_set_quadrature(ctx, 2, x.ctypes.data_as(dp), x.ctypes.data_as(dp), 2,
                                     x.ctypes.data_as(dp), x.ctypes.data_as(dp)) == INVALID
    assert quad(lib, ctx, 32) == OK and quad(lib, ctx, 1) == OK


def test_bad_arguments_return_invalid(lib, ctx):
    assert quad(lib, ctx, 6) == OK
    n = 16
    p = torch.from_numpy(W.make_points(n, seed=2)).cuda()
    o = torch.empty((20, n), dtype=torch.float64, device="cuda")
    pp, op = p.data_ptr(), o.data_ptr()
    assert gray(lib, ctx, n, n - 1, pp, op) == INVALID            # ld < npts
    assert gray(lib, ctx, -1, n, pp, op) == INVALID
    assert gray(lib, ctx, n, n, None, op) == INVALID
    assert gray(lib, ctx, n, n, pp, op, stim=2) == INVALID
    assert "stim_abs" in message(lib, ctx)
    assert gray(lib, ctx, n, n, pp, op, prec=7) == INVALID
    bad = api.OpacityParams(brem_implementation=3)
    assert gray(lib, ctx, n, n, pp, op, params=bad) == INVALID
    assert "brem" in message(lib, ctx)
    nu = torch.ones((2, n), dtype=torch.float64, device="cuda")
    s = torch.empty((32, n), dtype=torch.float64, device="cuda")
    sp = s.data_ptr()
    torch.cuda.synchronize()
    assert spectral(lib, ctx, n, n, pp, 2, nu.data_ptr(), n - 1, sp) == INVALID  # nu_ld < npts
    assert spectral(lib, ctx, n, n, pp, -2, nu.data_ptr(), 0, sp) == INVALID
    assert spectral(lib, ctx, n, n, pp, 2, None, 0, sp) == INVALID
    assert spectral(lib, ctx, n, n, pp, 2, nu.data_ptr(), 0, sp, stim=5) == INVALID
    assert spectral(lib, ctx, n, n, pp, 2, nu.data_ptr(), 0, sp, params=bad) == INVALID
    assert lib.bnsamd_synchronize(ctx) == OK  # nothing was launched, the context is usable


def test_empty_inputs_are_noops_and_inputs_are_not_modified(lib, ctx):
    assert quad(lib, ctx, 6) == OK
    n = 32
    host = W.make_points(n, seed=3)
    p = torch.from_numpy(host).cuda()
    o = torch.full((20, n), -7.0, dtype=torch.float64, device="cuda")
    assert gray(lib, ctx, 0, n, p.data_ptr(), o.data_ptr()) == OK
    nu = torch.tensor([1.0, 10.0, 100.0], dtype=torch.float64, device="cuda")
    s = torch.full((48, n), -7.0, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()  # the context launches on its own stream
    assert spectral(lib, ctx, 0, n, p.data_ptr(), 3, nu.data_ptr(), 0, s.data_ptr()) == OK
    assert spectral(lib, ctx, n, n, p.data_ptr(), 0, nu.data_ptr(), 0, s.data_ptr()) == OK
    assert lib.bnsamd_synchronize(ctx) == OK
    assert torch.all(o == -7.0) and torch.all(s == -7.0)
    # a real call writes its outputs and leaves the inputs as they were
    assert gray(lib, ctx, n, n, p.data_ptr(), o.data_ptr()) == OK
    assert spectral(lib, ctx, n, n, p.data_ptr(), 3, nu.data_ptr(), 0, s.data_ptr()) == OK
    assert lib.bnsamd_synchronize(ctx) == OK
    assert np.array_equal(p.cpu().numpy(), host)
    assert np.array_equal(nu.cpu().numpy(), [1.0, 10.0, 100.0])
    assert torch.all(torch.isfinite(o)) and torch.all(torch.isfinite(s))
