"""The library's host code under AddressSanitizer + UndefinedBehaviorSanitizer and under
ThreadSanitizer (tests/cpp/Makefile targets `sanitize` and `tsan`, host-only instrumentation):
tests/cpp/host_sanitize.cpp drives problem preparation and planning for 1 to 20 modes, the
tableaux, the error paths (non-finite field, all-zero or empty coupling, bad sizes and indices,
no device), the request server's copies, records and budget refusal, and planning from eight
threads at once. No device is needed."""
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(REPO, "tests", "cpp")


def _binary(target, name):
    path = os.path.join(CPP, "_build", name)
    r = subprocess.run(["make", "-s", "-j4", "-C", CPP, target], capture_output=True, text=True, timeout=1200)
    if r.returncode != 0:
        pytest.fail(f"sanitizer build failed:\n{r.stdout[-3000:]}\n{r.stderr[-3000:]}")
    return path


@pytest.mark.parametrize("target,name,env", [
    ("sanitize", "host_sanitize_asan", {"ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1",
                                        "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1"}),
    ("tsan", "host_sanitize_tsan", {"TSAN_OPTIONS": "halt_on_error=1"}),
])
def test_host_code_under_sanitizers(target, name, env):
    exe = _binary(target, name)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env={**os.environ, **env})
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-6000:]
    assert "host checks passed" in r.stdout
