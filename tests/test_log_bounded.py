"""The bounded logarithm of the device code (log_bounded in
bns_nurates_amd/csrc/gray_math.h), used for the HR98 bremsstrahlung
logarithms and the neutrino-energy logarithms of the spectral path where the
reference calls log and pow: the host build of the same operations against
the x87 extended-precision logarithm on 27 M samples (the normal range,
around 1, every power of two) must stay within one ulp.  The device path is
covered end to end by the GPU parity tests (HR98 against the reference's
outputs and the oracle)."""
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "tests", "cpp", "log_bounded_check")


def test_log_bounded_within_one_ulp():
    assert os.path.exists(CHECK), "tests/cpp/log_bounded_check not built (run build())"
    r = subprocess.run([CHECK], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    m = re.search(r"max error ([0-9.]+) ulp.*correctly rounded ([0-9.]+)", r.stdout)
    assert m, r.stdout
    assert float(m.group(1)) < 1.0 and float(m.group(2)) > 0.98, r.stdout
