"""The deep-tail form of the Fermi integrals (bns_nurates_amd/csrc/gray_math.h,
fdi_bundle with DEEP): below x = -40 the exponential form of every order
F_0..F_5 (functions.hpp:1551-2837) rounds to e^x a0, so a wavefront whose
lanes all lie there may skip the rationals without changing a bit.  The host
build of the same header checks the identity on 4.8 M samples, and checks
that it fails at x = -36 (inside the bound's margin), so that a wrong bound
would be seen.  The device path itself is covered bit for bit by the GPU
parity tests (positron NES bundles at eta > 40)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "tests", "cpp", "fdi_deep_check")


def test_deep_tail_identity():
    assert os.path.exists(CHECK), "tests/cpp/fdi_deep_check not built (run build())"
    r = subprocess.run([CHECK], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "deep tail mismatches 0" in r.stdout
