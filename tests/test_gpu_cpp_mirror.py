"""The reference-style C++ harness (tests/cpp/test_mirror.cpp) over the C++
mirror of the reference API (include/rd_amd.hpp), run on the GPU."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "tests" / "cpp" / "test_mirror"

pytestmark = pytest.mark.gpu


def test_cpp_mirror_harness(gpu):
    assert BIN.exists(), "tests/cpp/test_mirror is not built (run __graft_entry__.build())"
    r = subprocess.run([str(BIN), str(ROOT / "tests" / "golden" / "ref_vectors")], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0 and "RESULT: PASS" in r.stdout, r.stdout[-2000:] + r.stderr[-2000:]
