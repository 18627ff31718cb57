"""Host sanitizers on the CPU: the FP64 oracle and the C ABI's host code
(ocean_capi.cpp) built with AddressSanitizer + UndefinedBehaviorSanitizer
(tests/sanitize/Makefile, plain gcc/g++), then run over every oracle function
at N=64 and every host-side validation path of the ABI (bad sizes, NaN and
infinite parameters, NULL handles and pointers, bad batch/fit/interop
arguments) and the host utilities at their largest sizes.  Any sanitizer
report (out-of-bounds access, use after free, leak, undefined behaviour)
aborts the run with a non-zero status."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "sanitize")


def test_oracle_and_abi_validation_are_clean_under_asan_and_ubsan():
    if not os.path.exists(os.path.join(SAN, "Makefile")):
        pytest.skip("sanitizer build files not present")
    if shutil.which("g++") is None or not os.path.exists("/opt/rocm/lib/libamdhip64.so"):
        pytest.skip("g++ or the HIP runtime is missing")
    b = subprocess.run(["make", "-s"], cwd=SAN, capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stdout[-3000:] + b.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="abort_on_error=1:detect_leaks=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([os.path.join(SAN, "san_check")], capture_output=True, text=True, timeout=300, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-5000:]
    assert "san_check OK" in r.stdout
    assert "FAIL" not in r.stdout
    assert "runtime error" not in out and "AddressSanitizer" not in out
    assert r.stdout.count("\nok ") >= 100
