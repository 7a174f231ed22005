"""The CPU side under AddressSanitizer and UndefinedBehaviorSanitizer: the
oracle's restatement (oracle/rd_oracle.cpp) on random and extreme inputs of
every stage plus the committed fxp vectors, and the library's host-side file
readers (rd_io.hip) on the reference's fixtures, on damaged copies of them and
on an iq_ascii_v1 round trip (tests/cpp/san_check.cpp, built by
tests/cpp/Makefile.san). Any sanitizer report aborts the run."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
CPP = ROOT / "tests" / "cpp"


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_cpu_side_under_asan_and_ubsan(tmp_path):
    b = subprocess.run(["make", "-s", "-f", "Makefile.san"], cwd=CPP, capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stdout[-2000:] + b.stderr[-2000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(CPP / "san_check"), str(ROOT / "tests" / "golden" / "ref_vectors"), str(tmp_path)],
                       capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0 and "RESULT: PASS" in r.stdout, r.stdout[-2000:] + r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
