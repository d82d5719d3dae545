"""The product's host code (tree build, CPU search, serialization) under AddressSanitizer and
UndefinedBehaviorSanitizer, with the reference's debug cross-checks compiled in (TK_DEBUG_CHECKS:
iterative vs recursive search, src/kdtree.cpp:182-189; leaf offsets, src/kdtree.hpp:517-529), on
clouds with ties, NaN and infinite coordinates and tiny leaves, each result equal to the oracle's.
Host only: built with g++ from tests/sanitize/Makefile."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None, reason="needs g++ and make")
def test_host_code_under_sanitizers_with_debug_checks():
    build = subprocess.run(["make", "-s", "-C", os.path.join(HERE, "sanitize")], capture_output=True, text=True)
    assert build.returncode == 0, build.stdout + build.stderr
    exe = os.path.join(os.path.dirname(HERE), "_build_sanitize", "host_selftest")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
    out = run.stdout + run.stderr
    assert run.returncode == 0, out[-4000:]
    assert "host selftest OK" in run.stdout
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out
