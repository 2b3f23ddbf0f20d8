"""The raster's k-step advance of the barycentric recurrence (s <- fl(s + D),
SL_TriRasterizer.cpp:857-910) must be bit-identical to stepping: checked on
seeded cases covering binade crossings, ties, zero crossings, subnormals and
steps below half an ulp.  The header is compiled for the host with the
kernels' flags (-ffp-contract=off)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


import pytest


@pytest.mark.parametrize("variant", [[], ["-DSKEWED_RCP"]], ids=["exact", "skewed_rcp"])
def test_recurrence_advance_is_exact(tmp_path, variant):
    exe = tmp_path / "recur_check"
    src = os.path.join(ROOT, "tests", "native", "recur_check.cpp")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", *variant, "-o", str(exe), src], check=True)
    r = subprocess.run([str(exe), "1000000"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout
    n, bad = map(int, r.stdout.strip().split()[-2:])
    assert n == 1000000 and bad == 0
