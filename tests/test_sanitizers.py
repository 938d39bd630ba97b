"""The C runtime and the Cholesky application under ThreadSanitizer and
under AddressSanitizer + UndefinedBehaviorSanitizer (CPU workers only;
tests/sanitize/Makefile builds them with gcc, tests/sanitize/rt_stress.c
drives the DAG and task cases of tests/test_runtime_cpu.py).  A race, a
use-after-free, a leak or undefined behaviour fails the run."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SAN = os.path.join(HERE, "sanitize")


@pytest.fixture(scope="module")
def built():
    r = subprocess.run(["make", "-s", "-C", SAN, "all"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True)
    assert r.returncode == 0, r.stdout[-4000:]
    return os.path.join(SAN, "_build")


@pytest.mark.parametrize("kind,env", [
    ("tsan", {"TSAN_OPTIONS": "halt_on_error=1 exitcode=66 second_deadlock_stack=1"}),
    ("asan", {"ASAN_OPTIONS": "detect_leaks=1 exitcode=67", "UBSAN_OPTIONS": "print_stacktrace=1"}),
])
def test_runtime_is_sanitizer_clean(built, kind, env):
    e = dict(os.environ, **env)
    r = subprocess.run([os.path.join(built, f"rt_stress_{kind}")], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, env=e, timeout=600)
    out = r.stdout
    assert r.returncode == 0, out[-6000:]
    assert "all checks passed" in out
    for marker in ("WARNING: ThreadSanitizer", "ERROR: AddressSanitizer", "ERROR: LeakSanitizer",
                   "runtime error:"):
        assert marker not in out, out[-6000:]
