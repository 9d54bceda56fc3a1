"""The C-ABI library loads and exports every entry point include/*.h
declares (CPU only: nothing is called that needs a device)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "bns_nurates_amd", "libbnsamd.so")


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"\b(bnsamd_[a-z0-9_]+)\s*\(", src):
            names.add(m.group(1))
    return names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libbnsamd.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    names = declared_symbols()
    assert len(names) >= 14
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libbnsamd.so not built")
def test_host_gauss_legendre_matches_reference_rule(golden):
    # host-only entry point: no device needed
    from bns_nurates_amd import api
    g = golden("gauss_legendre.npz")
    for n in range(1, 21):
        x, w = api.gauss_legendre(n)
        assert (x == g[f"x{n}"]).all() and (w == g[f"w{n}"]).all(), n


def test_header_layout_matches_workloads():
    from bns_nurates_amd import workloads as W
    src = open(os.path.join(ROOT, "include", "bns_nurates_amd.h")).read()
    val = lambda name: int(re.search(name + r"\s*=\s*(\d+)", src).group(1))
    assert val("BNSAMD_PT_NFIELDS") == W.P_NFIELDS
    assert val("BNSAMD_PT_N") == W.P_N0 and val("BNSAMD_PT_J") == W.P_J0
    assert val("BNSAMD_PT_CHI") == W.P_CHI0
    assert val("BNSAMD_DI_NFIELDS") == W.D_NFIELDS and val("BNSAMD_OUT_NFIELDS") == W.O_NFIELDS


def test_flag_and_param_coercion():
    import numpy as np
    from bns_nurates_amd import api
    from bns_nurates_amd import workloads as W
    f = api.as_flags(W.FLAGS_ALL)
    assert (f.use_abs_em, f.use_iso) == (1, 1)
    p = api.as_params(W.params_default_all(W.BREM_BRT06))
    assert p.brem_implementation == 1 and p.neglect_blocking == 1
    assert api.as_flags(None) == api.OpacityFlags()
    with pytest.raises(api.BnsAmdError):
        api.as_params(np.zeros(3))


@pytest.mark.skipif(not os.path.exists(LIB), reason="libbnsamd.so not built")
def test_python_mirror_declares_every_signature():
    """Every C entry point has ctypes argtypes/restype in api.load_library:
    an undeclared one would receive Python ints as 32-bit C ints and
    truncate pointers."""
    from bns_nurates_amd import api
    lib = api.load_library()
    missing = [n for n in sorted(declared_symbols()) if getattr(lib, n).argtypes is None]
    assert not missing, missing
