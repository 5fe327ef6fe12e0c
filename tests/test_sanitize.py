"""The library's host code (spring builders, scene generators, OBJ reader,
mesh set-up) and the oracle under AddressSanitizer and
UndefinedBehaviorSanitizer: a CPU-only build (tests/sanitize/Makefile, g++ /
gcc, no GPU code) of a driver that runs them on inputs reaching every branch
of their loops and compares each library result with the oracle's."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_host_code_and_oracle_under_asan_and_ubsan():
    subprocess.run(["make", "-s", "-C", HERE, "-j4"], check=True, timeout=600)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([os.path.join(HERE, "_build", "sanitize_run")], capture_output=True, text=True,
                       timeout=600, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr
    assert "sanitize_run: ok" in r.stdout
