"""Argument checks of the Python host API that must reject a call before
anything is launched (CPU only: a stand-in library records the calls that
reach it, and stand-in device tensors carry shapes and strides)."""
import pytest

from bns_nurates_amd import api


class _Tensor:
    """Shape, strides and pointer of a [rows, cols] float64 device tensor."""
    dtype = "torch.float64"
    is_cuda = True

    def __init__(self, rows, cols, ld, ptr=0x10000):
        self.shape = (rows, cols)
        self._ld = ld
        self._ptr = ptr

    def dim(self):
        return 2

    def stride(self, i):
        return (self._ld, 1)[i]

    def data_ptr(self):
        return self._ptr


class _Lib:
    def __init__(self):
        self.calls = []

    def __getattr__(self, name):
        def f(*a):
            self.calls.append(name)
            return 0
        return f


def _ctx():
    c = object.__new__(api.Context)
    c.lib = _Lib()
    c.h = None
    c.device = 0
    c._quad = (16, 16)
    c._stream_explicit = True  # no stream query
    return c


@pytest.mark.parametrize("method", ["distr_from_m1", "equilibrium"])
def test_distribution_block_with_another_leading_dimension_is_rejected(method):
    """pts as a column slice of a wider block (ld 4096) with a compact
    distribution block (ld 1000): the kernel would address the distribution
    rows at the point block's stride, so the call must fail unlaunched."""
    torch = pytest.importorskip("torch")  # noqa: F841 (the methods import it)
    c = _ctx()
    pts = _Tensor(api.PT_NFIELDS, 1000, 4096)
    distr = _Tensor(api.DI_NFIELDS, 1000, 1000, ptr=0x20000)
    with pytest.raises(api.BnsAmdError, match="leading dimension"):
        getattr(c, method)(pts, distr)
    assert not [n for n in c.lib.calls if n != "bnsamd_set_stream"]


def test_matching_leading_dimension_reaches_the_library():
    pytest.importorskip("torch")
    c = _ctx()
    pts = _Tensor(api.PT_NFIELDS, 1000, 4096)
    distr = _Tensor(api.DI_NFIELDS, 1000, 4096, ptr=0x20000)
    c.distr_from_m1(pts, distr)
    c.equilibrium(pts, distr)
    assert c.lib.calls.count("bnsamd_distr_from_m1") == 1
    assert c.lib.calls.count("bnsamd_equilibrium") == 1


def test_host_output_array_is_checked():
    import numpy as np
    c = _ctx()
    pts = np.zeros((api.PT_NFIELDS, 8))
    with pytest.raises(api.BnsAmdError, match="out"):
        c.compute_m1_opacities(pts, out=np.zeros((20, 7)))
    with pytest.raises(api.BnsAmdError, match="out"):
        c.compute_m1_opacities(pts, out=np.zeros((20, 8), np.float32))
    assert "bnsamd_m1_opacities_host" not in c.lib.calls


def test_gp19_table_shape_is_checked():
    import numpy as np
    from tests import gp19_synth as G
    c = _ctx()
    t = G.table("smooth")
    t["data"] = t["data"][:, :, :, :-1]
    with pytest.raises(api.BnsAmdError, match="gp19"):
        c.set_gp19_table(t)
    assert "bnsamd_set_gp19_table" not in c.lib.calls


def test_spectral_layout_is_checked():
    import numpy as np
    c = _ctx()
    pts = np.zeros((api.PT_NFIELDS, 4))
    with pytest.raises(api.BnsAmdError, match="layout"):
        c.compute_spectral_opacities(pts, np.ones(3), layout="points")
    with pytest.raises(api.BnsAmdError, match="device tensors"):
        c.compute_spectral_opacities(pts, np.ones(3), layout="struct")
    assert not c.lib.calls
