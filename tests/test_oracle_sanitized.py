"""The oracle built with AddressSanitizer and UndefinedBehaviorSanitizer
(oracle/Makefile target `san`, host code only) runs every entry point on small
2D and 1D problems without a report (UBSan does not recover: any undefined
behaviour aborts the run)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++ with the sanitizer runtimes")
def test_oracle_under_asan_and_ubsan():
    subprocess.run(["make", "-s", "-C", ORACLE, "san"], check=True, timeout=300)
    env = dict(os.environ, OMP_NUM_THREADS="1",
               ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([os.path.join(ORACLE, "fro_san")], env=env, capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0, r.stderr[-4000:]
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
    assert r.stdout.strip().endswith("OK")
