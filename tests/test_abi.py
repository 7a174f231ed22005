"""The C-ABI library: loads, exports every entry point include/rd_amd.h declares,
and its host-side pieces work without a GPU (no compute calls here)."""
import re
from pathlib import Path

import numpy as np

import rd_amd

ROOT = Path(__file__).resolve().parent.parent


def declared(header="rd_amd.h"):
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rd_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = rd_amd.load()
    names = declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib.lib, n)]
    assert not missing, missing
    assert set(names) == set(rd_amd.EXPORTS)


def test_every_shadow_symbol_is_exported():
    import rd_shadow

    lib = rd_amd.load()
    names = declared("rd_shadow.h")
    missing = [n for n in names if not hasattr(lib.lib, n)]
    assert not missing, missing
    assert set(names) == set(rd_shadow.SH_EXPORTS)
    rd_shadow._lib()  # argtypes bind without a GPU


def test_twiddle_table_matches_oracle(orc):
    lib = rd_amd.load()
    assert np.array_equal(lib.twiddle_table(), orc.twiddle_table(4096))


def test_twiddle_magnitudes_bound_the_fft_products():
    """The FFT kernel keeps twiddle products in 32 bits: with |x.re|, |x.im| <= 32768,
    the partial sum x.re*w.re - x.im*w.im - x.im of the complemented form stays
    within 32768*(|w.re| + |w.im| + 1) < 2^31, which needs |w.re| + |w.im| <= 65534."""
    t = rd_amd.load().twiddle_table()
    re = (t & 0xFFFF).astype(np.int64)
    im = (t >> 16).astype(np.int64)
    re = np.where(re >= 0x8000, re - 0x10000, re)
    im = np.where(im >= 0x8000, im - 0x10000, im)
    assert np.max(np.abs(re) + np.abs(im)) <= 65534
    assert np.max(re * re + im * im) < 32769 * 32769


def test_status_strings_and_no_device_path():
    lib = rd_amd.load()
    assert lib.lib.rd_status_str(0) == b"ok"
    assert lib.lib.rd_status_str(-2) == b"geometry rejected"
    if lib.device_count() == 0:
        import ctypes as C

        h = C.c_void_p()
        assert lib.lib.rd_create(0, C.byref(h)) == rd_amd.RD_E_NODEV
