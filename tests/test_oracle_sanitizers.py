"""The oracle under AddressSanitizer + UndefinedBehaviorSanitizer and under
ThreadSanitizer (oracle/sanitize_main.c): products on four threads sharing the
plan cache (src/fpntt.jl:273-293), each equal to a schoolbook product, with no
sanitizer report."""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


@pytest.fixture(scope="module")
def built():
    if not shutil.which("gcc") or not shutil.which("make"):
        pytest.skip("gcc/make not available")
    r = subprocess.run(["make", "-C", ORACLE, "sanitize"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return ORACLE


@pytest.mark.parametrize("kind", ["asan", "tsan"])
def test_oracle_clean_under_sanitizer(built, kind):
    exe = os.path.join(built, "build", f"oracle_{kind}")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120,
                       env={**os.environ, "TSAN_OPTIONS": "halt_on_error=1", "ASAN_OPTIONS": "detect_leaks=1"})
    out = r.stdout + r.stderr
    assert r.returncode == 0, out
    assert "0 mismatches" in out
    assert "Sanitizer" not in out, out
