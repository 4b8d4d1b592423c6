"""The C oracle under AddressSanitizer and UndefinedBehaviorSanitizer: every
routine runs once on small inputs (oracle/sanitize_check.c), with invariant
checks (Parseval vs real-space energy, FFT round trip, direct-space forces
summing to zero, SPME force = -dE/dx by central difference, the exact
Mehrstellen solve as a fixed point of the relaxation).  CPU only."""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_oracle_under_asan_ubsan(tmp_path):
    exe = tmp_path / "sanitize_check"
    b = subprocess.run(["make", "-s", "-C", ORACLE, "sanitize", f"SAN_OUT={exe}"],
                       capture_output=True, text=True)
    assert b.returncode == 0, b.stdout + b.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all checks passed" in r.stdout
