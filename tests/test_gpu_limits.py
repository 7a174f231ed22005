"""GPU parity at the size limits the C ABI accepts.

The covariance's int8 digit GEMMs accumulate in int32; a frame adds at most
2^15 to one digit accumulator, so the largest frame count (65535) is exact only
if no accumulator sums two digit kinds. These tests drive every accumulator to
its bound: all-(-32768) and all-(-32767) inputs over 32768 and 65535 frames, for
each padded channel count (16 and 32 with an odd bin count: four and two frame
splits per block in the single-CPI launch; 32 and 64 with the bins a multiple
of 64 / CH: the loader-fed kernel, every frame in one block), and through the
batch with one frame split.
Semantics: covar::cross_power summed by an acc_w = 64 block-mode Integrator
(model/cpp/covariance/covar_model.hpp:137-151, 186-262).
"""
import numpy as np
import pytest

import rd_amd
import rd_scenario
import vecio

pytestmark = pytest.mark.gpu


def _eq(a, b, what):
    bad = np.flatnonzero(np.asarray(a) != np.asarray(b))
    assert bad.size == 0, f"{what}: {bad.size} mismatches, first at {bad[:5]}"


@pytest.mark.parametrize("CH,N", [(16, 3), (32, 3), (32, 2), (64, 1)])
@pytest.mark.parametrize("F", [32768, 65535])
def test_cov_full_scale_long_frames(gpu, orc, CH, N, F):
    print(f"cov CH={CH} F={F} N={N}", flush=True)
    for val in (vecio.pack(-32768, -32768), vecio.pack(-32767, -32767), vecio.pack(-32768, 32767)):
        Xa = np.full(F * N * CH, val, np.uint32)
        _eq(gpu.cov(Xa, CH, F, N), orc.cov(Xa, CH, F, N, threads=16)[0], f"R CH={CH} F={F} value {val:#x}")


def test_cov_full_scale_long_frames_random_extreme(gpu, orc):
    rng = np.random.default_rng(65535)
    CH, F, N = 64, 65535, 1
    # mostly -32768 with a few other corner values: the sums stay near their bound
    vals = np.array([vecio.pack(-32768, -32768)] * 12 + [vecio.pack(-32767, 32767), vecio.pack(32767, -32768)],
                    np.uint32)
    Xa = rng.choice(vals, F * N * CH)
    _eq(gpu.cov(Xa, CH, F, N), orc.cov(Xa, CH, F, N, threads=16)[0], "R random extreme")


@pytest.mark.parametrize("autotune", [False, True])
def test_batch_cov_full_scale_long_frames(gpu, orc, autotune):
    """CPIs of 65535 frames whose bin 0 stays near -32767 in every channel and
    frame (a constant full-scale input through identity taps), so every digit
    accumulator of the covariance runs to its bound; the batch runs one frame
    split for 64 and 32 channels (loader-fed) and the autotuner also tries one
    for 16."""
    specs = [(64, 8, 65535), (32, 8, 65535), (16, 8, 65535)]
    chains, exp = [], []
    for CH, N, F in specs:
        T, B = 16, 1
        iq = np.full(CH * N * F, vecio.pack(-32768, -32768), np.uint32)
        taps = np.zeros(N * T, np.uint32)
        taps[::T] = vecio.pack(32767, 0)
        W = np.full(B * CH, vecio.pack(1024, 0), np.uint32)
        cfg = np.array([1, 0, 0, 1, 1, 2, 2, 0x800], np.uint32)
        ch = rd_amd.Chain(gpu, CH, N, T, F, B, cfg, keep_y=False, keep_flags=False)
        ch.load(iq, taps, W)
        chains.append(ch)
        exp.append((CH, N, T, F, B, cfg, iq, taps, W))
    b = rd_amd.Batch.of_chains(gpu, chains, autotune=autotune)
    for ch in chains:
        ch.b["R"].zero()
    b.launch()
    totals = b.events()
    for ch, (CH, N, T, F, B, cfg, iq, taps, W), tot in zip(chains, exp, totals):
        print(f"batch CPI CH={CH} N={N} F={F}", flush=True)
        o = orc.chain(iq, taps, W, CH, N, T, F, B, cfg, rd_scenario.MAX_GUARD, rd_scenario.MAX_REF,
                      keep=("X", "power", "R", "events"), threads=16)
        got = ch.fetch(("X", "power", "R"), n_events=int(tot))
        x0 = got["X"].reshape(CH, F, N)[:, :, 0]
        assert np.all((x0 & 0xFFFF) >= 0x8000), "bin 0 is not full-scale negative"
        _eq(got["R"], o["R"], f"R CH={CH}")
        _eq(got["power"], o["power"], f"power CH={CH}")
        assert int(tot) == o["n_events"]
        _eq(got["events"], o["events"], f"events CH={CH}")
    b.close()


def test_cfar_enqueue_matches_cfar_run(gpu, orc):
    """rd_cfar_enqueue (total left on the device) equals rd_cfar_run and the
    oracle; with no rows it writes a zero total."""
    import ctypes as C

    rng = np.random.default_rng(5)
    n, rows = 1024, 37
    cells = rng.integers(0, 1 << 20, n * rows).astype(np.uint32)
    cfg8 = [1, 0, 0, 4, 4, 16, 16, 0x200]
    ev, rc, tot, st = gpu.cfar(cells, n, cfg8, 31, 63)
    dc = gpu.put(cells)
    cap = rows * (n + 1)
    dev, drc, dtot = gpu.buf(24 * cap), gpu.buf(4 * rows), gpu.buf(8)
    cfg = rd_amd.CfarCfg.of(cfg8)
    gpu.lib.check("rd_cfar_enqueue", gpu.L.rd_cfar_enqueue(gpu.h, dc.ptr, 4, n, rows, C.byref(cfg), 31, 63, None, 0,
                                                            dev.ptr, cap, drc.ptr, dtot.ptr))
    gpu.sync()
    t = int(dtot.download(np.uint64, 1)[0])
    assert t == tot
    _eq(dev.download(np.uint64, 3 * t).reshape(t, 3), ev, "events")
    _eq(drc.download(np.uint32, rows), rc, "row counts")
    o = orc.cfar(cells.astype(np.uint64), n, cfg8, 31, 63)
    _eq(ev, o[0], "events vs oracle")
    dtot.upload(np.array([12345], np.uint64))
    gpu.lib.check("rd_cfar_enqueue", gpu.L.rd_cfar_enqueue(gpu.h, dc.ptr, 4, n, 0, C.byref(cfg), 31, 63, None, 0,
                                                            dev.ptr, cap, drc.ptr, dtot.ptr))
    gpu.sync()
    assert int(dtot.download(np.uint64, 1)[0]) == 0
