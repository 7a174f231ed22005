"""Shared test setup: repo paths, the gpu marker, seeded stimulus helpers."""
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "agilex-7-radar-detect_amd"
for p in (str(ROOT), str(PKG), str(ROOT / "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP library via its C ABI)")
    config.addinivalue_line("markers", "slow: longer CPU cross-checks")


@pytest.fixture(scope="session")
def orc():
    import oracle

    return oracle.load("oracle")


@pytest.fixture(scope="session")
def ref():
    import oracle

    if not oracle.ref_available():
        pytest.skip("reference models not built here (oracle/_ref absent)")
    return oracle.load("ref")


@pytest.fixture(scope="session")
def gpu():
    """A context on device 0 of the HIP library (loaded through its C ABI).
    Raises if the library is not built; skips only when no GPU is visible."""
    import rd_amd

    lib = rd_amd.load()
    if lib.device_count() < 1:
        pytest.skip("no GPU visible")
    ctx = lib.context(0)
    yield ctx
    ctx.close()
