"""The NEPS half-cell visit orders of the resident tables
(bns_nurates_amd/csrc/gray_tables.{h,cpp}): the kernel visits a point's
half-cells in the order of its electron-degeneracy bin, so each bin's order
must be a permutation of the records (a repeated or missing half-cell would
double-count or drop a quadrature cell of the 2D NES/NPS sums,
integration.hpp:687-768).  The host build of the table builder checks, for
quadrature orders 4..32, that every bin's order is a permutation, that the
last bin keeps the stored ascending-w order, that the graded bins group the
half-cells by the Fermi-integral form classes at the bin's central |eta|,
and that the bin index covers its range.  The device results in each bin
are covered by the GPU parity tests (2048 grid points span all bins)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "tests", "cpp", "table_order_check")


def test_table_orders():
    assert os.path.exists(CHECK), "tests/cpp/table_order_check not built (run build())"
    r = subprocess.run([CHECK], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "table order failures 0" in r.stdout
