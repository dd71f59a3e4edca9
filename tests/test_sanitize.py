"""The host-side code under sanitizers: the graph readers, writer and
generators, inversion, the pinned-upload thread team (against a simulated copy
engine) and the oracle, built with AddressSanitizer + UndefinedBehavior-
Sanitizer and with ThreadSanitizer from tests/san (host compilers only)."""
import os
import shutil
import subprocess

import pytest

SAN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "san")


@pytest.fixture(scope="module")
def san_build():
    if not shutil.which("g++"):
        pytest.skip("no host C++ compiler")
    r = subprocess.run(["make", "-C", SAN, "-j2"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    return os.path.join(SAN, "_build")


@pytest.mark.parametrize("variant", ["san_asan", "san_tsan"])
def test_host_code_under_sanitizer(san_build, variant):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1", TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([os.path.join(san_build, variant)], capture_output=True, text=True, timeout=600, env=env)
    report = r.stdout + r.stderr
    for marker in ("ERROR: AddressSanitizer", "ERROR: LeakSanitizer", "runtime error:", "WARNING: ThreadSanitizer"):
        assert marker not in report, report[-4000:]
    assert r.returncode == 0 and "san: ok" in r.stdout, report[-4000:]
