"""The header-only C++ host API (include/bns_nurates_amd.hpp).

CPU: the header compiles on its own with the reference's type and function
names a caller of bns_nurates uses (bns_nurates.hpp:76-474,
m1_opacities.hpp:1297-1313, distribution.hpp:254-473, integration.hpp:28).
GPU: tests/cpp/m1_host_api_check drives it the way the reference's test
driver drives ComputeM1Opacities (tests/tests.hpp:700-760) over the 4096-point
fixture, per point and batched, and the outputs are held to the parity bar
against the reference's own outputs.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from bns_nurates_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "m1_host_api_check")
SPEC_DRIVER = os.path.join(ROOT, "tests", "cpp", "spectral_host_api_check")

NAMES_SRC = r"""
#include "bns_nurates_amd.hpp"
#include <type_traits>
int main() {
  MyQuadrature q = quadrature_default;
  q.nx = 10;
  GreyOpacityParams g{};
  g.opacity_flags = opacity_flags_default_all;
  g.opacity_pars = opacity_params_default_all;
  g.opacity_pars.brem_implementation = BREM_BRT06;
  g.eos_pars.nb = 1e-3; g.eos_pars.temp = 10.; g.eos_pars.ye = 0.3; g.eos_pars.yp = 0.3;
  g.eos_pars.yn = 0.7; g.eos_pars.mu_p = 900.; g.eos_pars.mu_n = 950.; g.eos_pars.mu_e = 20.;
  g.eos_pars.mu_mu = 0.; g.eos_pars.dU = 0.; g.eos_pars.dm_eff = 0.;
  g.m1_pars.n[id_nue] = 1.; g.m1_pars.J[id_anue] = 1.; g.m1_pars.chi[id_nux] = 0.5;
  g.m1_pars.H[id_anux][3] = 0.;
  g.distr_pars.w_t[0] = g.distr_pars.temp_t[0] = g.distr_pars.eta_t[0] = 0.;
  g.distr_pars.w_f[0] = g.distr_pars.temp_f[0] = g.distr_pars.c_f[0] = g.distr_pars.beta_f[0] = 0.;
  static_assert(std::is_same<decltype(ComputeM1Opacities(&q, &q, &g)), M1Opacities>::value, "");
  static_assert(std::is_same<decltype(CalculateDistrParamsFromM1(&g.m1_pars, &g.eos_pars)),
                             NuDistributionParams>::value, "");
  static_assert(std::is_same<decltype(NuEquilibriumParams(&g.eos_pars)),
                             NuDistributionParams>::value, "");
  M1Opacities (*fp)(MyQuadrature*, MyQuadrature*, GreyOpacityParams*) =
      &ComputeM1OpacitiesNotStimulated;
  void (*gl)(MyQuadrature*) = &GaussLegendre;
  void (*eq)(const MyEOSParams*, const NuDistributionParams*, M1Quantities*) =
      &ComputeM1DensitiesEq;
  (void)fp; (void)gl; (void)eq;
  static_assert(std::is_same<decltype(ComputeSpectralOpacitiesNotStimulatedAbs(10., &q, &g)),
                             SpectralOpacities>::value, "");
  SpectralOpacities (*sp)(const BS_REAL, MyQuadrature*, GreyOpacityParams*) =
      &ComputeSpectralOpacitiesStimulatedAbs;
  (void)sp;
  SpectralOpacities so{};
  (void)(so.j[0] + so.kappa[0] + so.j_s[0] + so.kappa_s[0]);
  M1Opacities o{};
  return int(o.eta_0[0] + o.kappa_0_a[0] + o.eta[0] + o.kappa_a[0] + o.kappa_s[total_num_species - 1]);
}
"""


def test_header_compiles_with_reference_names(tmp_path):
    src = tmp_path / "names.cpp"
    src.write_text(NAMES_SRC)
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                        "-I", os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def _write_input(path, pts, nquad, stim, flags, params):
    npts = pts.shape[1]
    with open(path, "wb") as f:
        f.write(np.int64(npts).tobytes())
        f.write(np.array([nquad, stim], np.int32).tobytes())
        f.write(np.asarray(flags, np.int32).tobytes())
        f.write(np.asarray(params, np.int32).tobytes())
        f.write(np.ascontiguousarray(pts, np.float64).tobytes())


@pytest.mark.gpu
def test_cpp_driver_matches_reference_outputs(golden, tmp_path):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    assert os.path.exists(DRIVER), "tests/cpp/m1_host_api_check not built (run build())"
    from tests.gpu_util import check_parity, say
    g = golden("m1_pts4096_n10_brt06.npz")
    pts, ref = g["pts"], g["out"]
    npts, nsingle = pts.shape[1], 64
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    _write_input(inp, pts, 10, int(g["stim"]), g["flags"], g["params"])
    say(f"C++ host API driver: {npts} points, {nsingle} per-point calls")
    r = subprocess.run([DRIVER, str(inp), str(outp), str(nsingle)], capture_output=True,
                       text=True, timeout=120)
    say(r.stdout.strip(), r.stderr.strip())
    assert r.returncode == 0
    raw = np.fromfile(outp, dtype=np.float64)
    single = raw[:20 * nsingle].reshape(20, nsingle)
    batch = raw[20 * nsingle:20 * (nsingle + npts)].reshape(20, npts)
    distr = raw[20 * (nsingle + npts):].reshape(28, npts)
    # per-point and batched calls give the same bits
    assert np.array_equal(single, batch[:, :nsingle], equal_nan=True)
    # the reconstruction the driver made, against the oracle's
    d_ref, d_mag = O.bound_distr_from_m1(pts)
    assert O.parity_excess(distr, d_ref, d_mag).max() <= 1.0
    _, m = O.bound_m1_opacities(pts, 10, g["flags"], g["params"], int(g["stim"]))
    assert check_parity("C++ host API, 4096 pts brt06", batch, ref, m, W.OUT_NAMES) <= 1.0


@pytest.mark.gpu
def test_cpp_spectral_driver_matches_reference_outputs(golden, tmp_path):
    """tests/cpp/spectral_host_api_check makes the reference spectral test's
    calls (six points, equilibrium distributions from NuEquilibriumParams,
    100 energies, one ComputeSpectralOpacitiesNotStimulatedAbs per (energy,
    point)) through the C++ API, then the same pairs batched."""
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    assert os.path.exists(SPEC_DRIVER), "tests/cpp/spectral_host_api_check not built (run build())"
    from tests.gpu_util import check_parity, say
    g = golden("spectral_test6_n10_hr98.npz")
    pts, nu, ref = g["pts"], g["nu"], g["out"]
    npts, nnu = pts.shape[1], nu.shape[0]
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        f.write(np.array([npts, nnu], np.int64).tobytes())
        f.write(np.array([int(g["nquad"]), int(g["stim"])], np.int32).tobytes())
        f.write(np.asarray(g["flags"], np.int32).tobytes())
        f.write(np.asarray(g["params"], np.int32).tobytes())
        f.write(np.ascontiguousarray(pts, np.float64).tobytes())
        f.write(np.ascontiguousarray(nu, np.float64).tobytes())
    say(f"C++ spectral driver: {npts} points x {nnu} energies")
    r = subprocess.run([SPEC_DRIVER, str(inp), str(outp)], capture_output=True, text=True,
                       timeout=120)
    say(r.stdout.strip(), r.stderr.strip())
    assert r.returncode == 0
    raw = np.fromfile(outp, dtype=np.float64)
    k = 16 * nnu * npts
    single = raw[:k].reshape(16, nnu, npts)
    batch = raw[k:2 * k].reshape(16, nnu, npts)
    distr = raw[2 * k:].reshape(28, npts)
    assert np.array_equal(single, batch, equal_nan=True)
    # the equilibrium distributions the driver made, against the reference's
    np.testing.assert_allclose(distr, g["distr"], rtol=1e-14, atol=0)
    _, m = O.bound_spectral_opacities(pts, nu, int(g["nquad"]), g["flags"], g["params"],
                                      stim=int(g["stim"]), distr=g["distr"])
    names = [f"{q}[{s}]" for q in ("j", "kappa", "j_s", "kappa_s") for s in range(4)]
    assert check_parity("C++ host API, spectral test6", batch.reshape(16, -1),
                        ref.reshape(16, -1), m.reshape(16, -1), names) <= 1.0
