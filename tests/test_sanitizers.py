"""Sanitizer runs of the host-only C++ and of the oracle (CPU only).

tests/sanitize/ builds san_driver.cpp with the library's host-only sources
(cpu_fallback.cpp, host_helpers.cpp) and the oracle's C, twice: under
AddressSanitizer + UndefinedBehaviorSanitizer and under ThreadSanitizer.  The
driver checks the library's host entries against the oracle bit for bit on
small and edge-case inputs; any sanitizer report aborts it.  A planted
overflow and a planted race show that each build's instrumentation is live.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")
ENV = dict(os.environ,
           ASAN_OPTIONS="halt_on_error=1:detect_leaks=1:abort_on_error=0",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
           TSAN_OPTIONS="halt_on_error=1")


@pytest.fixture(scope="module")
def built():
    if not shutil.which("g++") or not shutil.which("gcc") or not shutil.which("make"):
        pytest.skip("g++/gcc/make not available")
    r = subprocess.run(["make", "-C", HERE, "-j4"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return {k: os.path.join(HERE, "_build", "san_" + k) for k in ("asan", "tsan")}


def run(path, *args):
    return subprocess.run([path, *args], capture_output=True, text=True, env=ENV, timeout=300)


@pytest.mark.parametrize("kind", ["asan", "tsan"])
def test_host_code_clean_under_sanitizer(built, kind):
    r = run(built[kind])
    print(r.stdout, r.stderr[-4000:])
    assert r.returncode == 0, r.stderr[-4000:]
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr
    assert ", 0 failed" in r.stdout


def test_asan_build_catches_planted_overflow(built):
    r = run(built["asan"], "--plant-overflow")
    assert r.returncode != 0
    assert "AddressSanitizer: heap-buffer-overflow" in r.stderr


def test_tsan_build_catches_planted_race(built):
    r = run(built["tsan"], "--plant-race")
    assert r.returncode != 0
    assert "ThreadSanitizer: data race" in r.stderr
