"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer
(host only): a driver runs every oracle entry point on small inputs
(oracle/sanitize_main.cpp); any out-of-bounds access, leak, overflow or
other undefined behaviour aborts the run."""

import os
import shutil
import subprocess
from pathlib import Path

import pytest

ORACLE = Path(__file__).resolve().parents[1] / "oracle"


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_oracle_under_asan_and_ubsan():
    b = subprocess.run(["make", "-s", "-C", str(ORACLE), "sanitize"], capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="2")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([str(ORACLE / "_san" / "oracle_san")], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "sanitized oracle run: ok" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
