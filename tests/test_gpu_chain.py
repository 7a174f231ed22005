"""GPU parity of the composed chain (one CPI through rd_chain_run) against the
oracle's composed chain: every stage output, every flag, the covariance and the
packed event list, bit for bit, on the published workloads.

The large workload (64 ch, 4096-pt, 512 frames) is checked at its full
geometry with 16 frames here (the oracle finishes in seconds); the full frame
count runs in bench.py with size-independent checks.
"""
import numpy as np
import pytest

import rd_amd
import rd_scenario

pytestmark = pytest.mark.gpu

STAGES = ("u", "pfb_flags", "X", "fft_flags", "Xa", "y", "yflags", "power", "R")


def _run_both(gpu, orc, sc, threads=16):
    ch = rd_amd.Chain(gpu, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg(), rd_scenario.MAX_GUARD, rd_scenario.MAX_REF)
    iq, taps, W = sc.iq(), sc.taps(), sc.weights()
    ch.load(iq, taps, W)
    tot = ch.run()
    got = ch.fetch(STAGES, n_events=tot)
    exp = orc.chain(iq, taps, W, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg(), rd_scenario.MAX_GUARD,
                    rd_scenario.MAX_REF, keep=STAGES + ("events",), threads=threads)
    return got, exp, tot


@pytest.mark.parametrize("name,frames", [("tiny", None), ("medium", None), ("large", 16)])
def test_chain_bit_exact(gpu, orc, name, frames):
    sc = rd_scenario.config(name)
    if frames:
        sc.F = frames
    got, exp, tot = _run_both(gpu, orc, sc)
    for k in STAGES + ("row_count",):
        a, b = got[k], exp[k]
        bad = np.flatnonzero(a != b)
        assert bad.size == 0, f"{name}: stage {k}: {bad.size} mismatches (first {bad[:4]})"
    assert tot == exp["n_events"]
    assert np.array_equal(got["events"], exp["events"]), f"{name}: events differ"
    assert tot > sc.B * sc.F  # at least one detection beyond the per-row summaries


def test_chain_saturating_scene(gpu, orc):
    """Full-scale input and max taps drive PFB, FFT and beamformer saturation."""
    sc = rd_scenario.config("tiny")
    rng = np.random.default_rng(3)
    vals = np.array([-32768, 32767, -32767, 0])
    iq = rd_scenario.pack(rng.choice(vals, sc.CH * sc.N * sc.F), rng.choice(vals, sc.CH * sc.N * sc.F))
    taps = rd_scenario.pack(np.full(sc.N * sc.T, -32768), np.full(sc.N * sc.T, -32768))
    W = rd_scenario.pack(np.full(sc.B * sc.CH, -32768), np.full(sc.B * sc.CH, -32768))
    ch = rd_amd.Chain(gpu, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg())
    ch.load(iq, taps, W)
    tot = ch.run()
    got = ch.fetch(STAGES, n_events=tot)
    exp = orc.chain(iq, taps, W, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg(), 31, 63, keep=STAGES + ("events",))
    assert got["pfb_flags"].any() and got["yflags"].any()
    for k in STAGES:
        assert np.array_equal(got[k], exp[k]), k
    assert np.array_equal(got["events"], exp["events"])


@pytest.mark.parametrize("lean", [False, True])
@pytest.mark.parametrize("name,frames", [("tiny", None), ("medium", 8)])
def test_chain_graph_replays(gpu, orc, name, frames, lean):
    """The chain captured as one HIP graph, replayed several times with eager work
    (output clobbering and an eager chain on other data) between the replays;
    every replay must equal the oracle."""
    sc = rd_scenario.config(name)
    if frames:
        sc.F = frames
    taps, W = sc.taps(), sc.weights()
    iq_g, iq_e = sc.iq(seed=11), sc.iq(seed=12)
    ch = rd_amd.Chain(gpu, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg())
    ch.load(iq_e, taps, W)
    src = gpu.put(iq_g)
    g = ch.graph(iq_ptr=src.ptr, lean=lean)
    checked = ("u", "X", "power", "R") if lean else STAGES  # a lean chain writes no aligned copy
    exp = orc.chain(iq_g, taps, W, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg(), 31, 63, keep=STAGES + ("events",))
    exp_e = orc.chain(iq_e, taps, W, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg(), 31, 63, keep=("events",))
    for rep in range(3):
        for k in ("u", "X", "Xa", "y", "power", "R", "events", "row_count"):
            ch.b[k].zero()
        n_e = ch.run()  # eager chain on the other input
        assert n_e == exp_e["n_events"]
        g.launch()
        gpu.sync()
        n = int(ch.total.download(np.uint64, 1)[0])
        got = ch.fetch(checked, n_events=n)
        assert n == exp["n_events"], rep
        for k in checked + ("row_count",):
            assert np.array_equal(got[k], exp[k]), (rep, k)
        assert np.array_equal(got["events"], exp["events"]), rep
    g.close()


def test_chain_fed_from_an_iq_ascii_scene(gpu, orc, tmp_path):
    """A scene written and read back in the reference's iq_ascii_v1 format
    feeds the chain; the result equals the oracle's chain on the original I/Q."""
    sc = rd_scenario.config("tiny")
    iq = sc.iq(seed=21)
    path = tmp_path / "tiny.iq"
    gpu.lib.write_iq_ascii(path, iq, sc.CH, sc.N, sc.F, 2, "tiny")
    back, A, N, H, P = gpu.lib.read_iq_ascii(path)
    assert (A, N, H) == (sc.CH, sc.N, sc.F)
    ch = rd_amd.Chain(gpu, A, N, sc.T, H, sc.B, sc.cfar_cfg())
    ch.load(back, sc.taps(), sc.weights())
    n = ch.run()
    got = ch.fetch(("power", "R"), n_events=n)
    exp = orc.chain(iq, sc.taps(), sc.weights(), sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg(), 31, 63,
                    keep=("power", "R", "events"))
    assert n == exp["n_events"]
    for k in ("power", "R", "events"):
        assert np.array_equal(got[k], exp[k]), k


def test_chain_graph_survives_context_buffer_growth(gpu, orc):
    """A graph keeps the context scratch it captured: a later call that needs a
    larger CFAR scratch on the same context leaves that buffer alive (until
    the context is destroyed), so replays after the growth stay exact."""
    ctx = rd_amd.Ctx(gpu.lib)
    try:
        sc = rd_scenario.config("tiny")
        taps, W = sc.taps(), sc.weights()
        iq = sc.iq(seed=21)
        ch = rd_amd.Chain(ctx, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg())
        ch.load(iq, taps, W)
        g = ch.graph()
        exp = orc.chain(iq, taps, W, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg(), 31, 63, keep=("power", "events"))
        rng = np.random.default_rng(3)
        rows, n = 3000, 64  # far more CFAR rows than the chain's: the scratch grows
        cells = rng.integers(0, 2**31, rows * n).astype(np.uint32)
        cfg = [1, 0, 0, 2, 2, 8, 8, 0x500]
        for rep in range(2):
            g.launch()
            ctx.sync()
            eg, _, tot, _ = ctx.cfar(cells, n, cfg, 31, 63, frame_mod=5)
            eo, _ = orc.cfar(cells.astype(np.uint64), n, cfg, 31, 63, frame_ids=np.arange(rows) % 5)
            assert tot == eo.shape[0] and np.array_equal(eg, eo), rep
            g.launch()
            ctx.sync()
            k = int(ch.total.download(np.uint64, 1)[0])
            got = ch.fetch(("power",), n_events=k)
            assert k == exp["n_events"] and np.array_equal(got["events"], exp["events"]), rep
            assert np.array_equal(got["power"], exp["power"]), rep
        g.close()
    finally:
        ctx.close()


def test_graph_keeps_its_chain_alive(gpu, orc):
    """A graph replays into its chain's device buffers, so the Python graph
    object holds the chain: dropping every other reference to the chain and
    collecting garbage leaves the buffers in place."""
    import gc

    sc = rd_scenario.config("tiny")
    taps, W = sc.taps(), sc.weights()
    iq = sc.iq(seed=31)
    ch = rd_amd.Chain(gpu, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg())
    ch.load(iq, taps, W)
    g = ch.graph()
    total = ch.total
    del ch
    gc.collect()
    g.launch()
    gpu.sync()
    exp = orc.chain(iq, taps, W, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg(), 31, 63, keep=("events",))
    assert int(total.download(np.uint64, 1)[0]) == exp["n_events"]
    got = g._keep.fetch((), n_events=exp["n_events"])
    assert np.array_equal(got["events"], exp["events"])
    g.close()
