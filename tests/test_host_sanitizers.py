"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU).

Builds tests/host_sanitize/host_check (g++ -fsanitize=address,undefined
-fno-sanitize-recover=all) from the library's host sources (mesh generator,
OBJ reader, geometry, receiver clusters, BVH builds in every layout, single
and multi-body) and the CPU oracle, and runs it: any out-of-bounds access,
use after free, leak or undefined operation ends the run with a report and a
nonzero status.  The driver also checks the product's geometry against the
oracle's bit for bit and the oracle's BVH walks against brute force.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "host_sanitize")
EXE = os.path.join(ROOT, "_build", "host_sanitize", "host_check")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_code_and_oracle_under_asan_ubsan():
    subprocess.run(["make", "-C", SRC], check=True, capture_output=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    p = subprocess.run([EXE, os.path.join(ROOT, "tests", "golden", "two_shapes.obj")],
                       capture_output=True, text=True, env=env, timeout=300)
    out = p.stdout + p.stderr
    assert "AddressSanitizer" not in out and "runtime error" not in out, out
    assert p.returncode == 0, out
    assert "0 failures" in p.stdout, out
