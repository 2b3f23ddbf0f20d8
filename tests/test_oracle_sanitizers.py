"""The oracle under AddressSanitizer + UndefinedBehaviorSanitizer and under
ThreadSanitizer (its thread pool): `make -C oracle asan tsan` builds the
instrumented libraries, tests/sanitizer_workload.py drives every oracle path
(triangles, points, lines, blending, the differentiable and smoothed-coverage
adjoints, four threads) in a child Python process with the sanitizer's
runtime preloaded; any report fails the test."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _runtime(name):
    try:
        p = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True, check=True).stdout.strip()
    except (OSError, subprocess.CalledProcessError):
        return None
    return p if os.path.isabs(p) and os.path.exists(p) else None


@pytest.fixture(scope="module")
def sanitizer_builds():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "asan", "tsan"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return os.path.join(ROOT, "oracle", "_build")


@pytest.mark.parametrize("kind,runtime,env", [
    ("asan", "libasan.so", {"ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1", "UBSAN_OPTIONS": "halt_on_error=1"}),
    ("tsan", "libtsan.so", {"TSAN_OPTIONS": "halt_on_error=1"}),
])
def test_oracle_is_sanitizer_clean(sanitizer_builds, kind, runtime, env):
    rt = _runtime(runtime)
    if rt is None:
        pytest.skip(f"{runtime} not found")
    lib = os.path.join(sanitizer_builds, f"libslo_{kind}.so")
    e = dict(os.environ, LD_PRELOAD=rt, **env)
    r = subprocess.run([sys.executable, "-m", "tests.sanitizer_workload", lib], cwd=ROOT, env=e,
                       capture_output=True, text=True, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "sanitizer workload: ok" in r.stdout, out[-4000:]
    for needle in ("AddressSanitizer", "runtime error:", "ThreadSanitizer", "LeakSanitizer"):
        assert needle not in out, out[-4000:]
