"""AddressSanitizer + UndefinedBehaviorSanitizer runs of host code, on the CPU (gcc/g++, no GPU
code): libcfgmm's C-ABI argument validation with the kernel launchers stubbed out, the CPU
oracle against naive loops, and the batched rho's reduction-pass assignment against brute
force.  Sources and the Makefile are in tests/sanitize/."""
import os
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")


@pytest.fixture(scope="module")
def built():
    subprocess.run(["make", "-s", "-j4", "-C", HERE], check=True, capture_output=True, timeout=600)
    return os.path.join(HERE, "_build")


@pytest.mark.parametrize("name", ["capi_checks", "oracle_checks", "waves_checks"])
def test_sanitized(built, name):
    env = dict(os.environ, ASAN_OPTIONS="abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([os.path.join(built, name)], capture_output=True, text=True, timeout=300, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "0 failed" in r.stdout
    assert "runtime error" not in out and "AddressSanitizer" not in out, out[-4000:]
