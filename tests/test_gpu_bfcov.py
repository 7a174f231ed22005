"""GPU parity of the fused beamformer + covariance stage (rd_bf_cov_stage):
one kernel reading the FFT layout once for CPIs of at most 16 channels with a
bin count that is a multiple of 4, the separate beamformer and covariance
kernels otherwise. Every output (y, flags, power, the aligned copy, R) against
the oracle's alignment, beamformer (beamformer_model.hpp:162-179) and block-mode
covariance (covar_model.hpp:137-151), bit for bit."""
import numpy as np
import pytest

import rd_amd
import stim
import vecio

pytestmark = pytest.mark.gpu

M32768 = vecio.pack(-32768, -32768)


def _eq(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    bad = np.flatnonzero(a != b)
    assert bad.size == 0, f"{what}: {bad.size} mismatches, first at {bad[:5]}: {a[bad[:5]]} vs {b[bad[:5]]}"


def _check(gpu, orc, X, W, CH, B, F, N, outputs=("y", "flags", "power", "Xa"), what="", x_layout=0):
    y, fl, pw, xa, R = gpu.bf_cov_stage(X, W, CH, B, F, N, outputs=outputs, x_layout=x_layout)
    oxa = orc.align(X, CH, F, N)
    oy, ofl, opw, _ = orc.bf(oxa, W, CH, B, want_acc=False)
    _eq(R, orc.cov(oxa, CH, F, N)[0], f"R {what}")
    if "power" in outputs:
        _eq(pw, opw, f"power {what}")
    if "y" in outputs:
        _eq(y, oy, f"y {what}")
    if "flags" in outputs:
        _eq(fl, ofl, f"flags {what}")
    if "Xa" in outputs:
        _eq(xa, oxa, f"aligned copy {what}")


def _inputs(rng, CH, B, F, N):
    S = CH * F * N
    return [("random", stim.rand_cx(rng, S), stim.rand_cx(rng, B * CH)),
            ("extreme", stim.rand_extreme(rng, S), stim.rand_extreme(rng, B * CH)),
            ("all -32768", np.full(S, M32768, np.uint32), np.full(B * CH, M32768, np.uint32)),
            ("-32768 x random", np.full(S, M32768, np.uint32), stim.rand_cx(rng, B * CH))]


# (CH, B, F, N): one and several 64-frame steps, a partial last step, one
# frame, 4 bins (one block), odd channel counts, 1..8 beam tiles
FUSED = [(16, 16, 5, 1024), (16, 16, 256, 64), (16, 16, 65, 8), (1, 1, 1, 4), (3, 2, 7, 12), (8, 8, 64, 4),
         (12, 17, 33, 16), (16, 128, 3, 8), (16, 40, 130, 4), (5, 16, 64, 256), (16, 16, 128, 32), (9, 31, 200, 20)]


@pytest.mark.parametrize("CH,B,F,N", FUSED)
def test_fused_stage_against_oracle(gpu, orc, CH, B, F, N):
    rng = np.random.default_rng(CH * 1000 + B * 17 + F * 3 + N)
    for name, X, W in _inputs(rng, CH, B, F, N):
        _check(gpu, orc, X, W, CH, B, F, N, what=f"{name} CH={CH} B={B} F={F} N={N}")


@pytest.mark.parametrize("CH,B,F,N", [(16, 16, 5, 1024), (16, 16, 256, 64), (16, 16, 65, 8), (1, 1, 1, 4),
                                      (12, 17, 33, 16), (16, 128, 3, 8), (9, 31, 200, 20)])
def test_fused_stage_blocked_layout(gpu, orc, CH, B, F, N):
    """X given bin-quad blocked [CH][N/4][F][4] (x_layout 1): the same bits
    as from the natural layout, every output including the aligned copy."""
    rng = np.random.default_rng(CH * 7 + B * 3 + F + N)
    for name, X, W in _inputs(rng, CH, B, F, N):
        _check(gpu, orc, X, W, CH, B, F, N, what=f"blocked {name} CH={CH} B={B} F={F} N={N}", x_layout=1)


@pytest.mark.parametrize("CH,B,F,N", [(16, 16, 256, 1024), (16, 16, 70, 8), (7, 20, 9, 12)])
def test_fused_stage_power_only(gpu, orc, CH, B, F, N):
    """The lean chain's output set: power and R, no beam samples, flags or
    aligned copy (the compile-time power-only kernel)."""
    rng = np.random.default_rng(CH + B + F + N)
    for name, X, W in _inputs(rng, CH, B, F, N)[:3]:
        _check(gpu, orc, X, W, CH, B, F, N, outputs=("power",), what=f"{name} CH={CH} B={B} F={F} N={N}")


@pytest.mark.parametrize("F", [32768, 65535])
def test_fused_accumulators_at_the_frame_limit(gpu, orc, F):
    """Full-scale input over up to 65535 frames: every int32 digit
    accumulator of the covariance at its bound (a frame adds at most 2^15)."""
    CH, B, N = 16, 16, 4
    rng = np.random.default_rng(F)
    X = np.full(CH * F * N, M32768, np.uint32)
    X[rng.integers(0, X.size, 64)] = vecio.pack(32767, 32767)
    W = stim.rand_extreme(rng, B * CH)
    _check(gpu, orc, X, W, CH, B, F, N, outputs=("power",), what=f"F={F}")


@pytest.mark.parametrize("CH,B,F,N", [(17, 16, 9, 64), (32, 32, 5, 16), (8, 8, 6, 6), (64, 64, 2, 128)])
def test_unfused_geometries_take_the_two_kernels(gpu, orc, CH, B, F, N):
    """More than 16 channels or bins not a multiple of 4: the separate
    beamformer and covariance, which need the aligned copy."""
    rng = np.random.default_rng(CH * B + F * N)
    name, X, W = _inputs(rng, CH, B, F, N)[1]
    _check(gpu, orc, X, W, CH, B, F, N, what=f"{name} CH={CH} B={B} F={F} N={N}")
    with pytest.raises(RuntimeError):
        gpu.bf_cov_stage(X, W, CH, B, F, N, outputs=("power",))
    if N % 4 == 0:  # the blocked layout is the one-kernel form's only
        with pytest.raises(RuntimeError):
            gpu.bf_cov_stage(X, W, CH, B, F, N, x_layout=1)


def test_fused_refuses_a_misaligned_x(gpu):
    """The one-kernel form moves X in 16-byte pieces: an X that is not 16-byte
    aligned is refused (RD_E_ARG, nothing launched)."""
    CH, B, F, N = 16, 16, 4, 8
    dX = gpu.put(np.zeros(CH * F * N + 4, np.uint32))
    dW = gpu.put(np.zeros(B * CH, np.uint32))
    dR = gpu.buf(16 * N * CH * (CH + 1) // 2)
    st = gpu.L.rd_bf_cov_stage(gpu.h, dX.ptr + 4, dW.ptr, CH, B, F, N, None, None, None, None, dR.ptr, 0)
    assert st == -1, st  # RD_E_ARG
    gpu.lib.check("rd_bf_cov_stage", gpu.L.rd_bf_cov_stage(gpu.h, dX.ptr + 16, dW.ptr, CH, B, F, N, None, None,
                                                           None, None, dR.ptr, 0))
    gpu.sync()


def test_fused_equals_the_two_kernels(gpu):
    """The fused kernel and rd_bf_stage + rd_cov_run give the same bits on a
    medium-sized CPI."""
    CH, B, F, N = 16, 16, 256, 1024
    rng = np.random.default_rng(77)
    X, W = stim.rand_cx(rng, CH * F * N), stim.rand_cx(rng, B * CH)
    y, fl, pw, xa, R = gpu.bf_cov_stage(X, W, CH, B, F, N)
    y2, fl2, pw2, xa2 = gpu.bf_stage(X, W, CH, B, F, N)
    _eq(xa, xa2, "aligned copy")
    _eq(y, y2, "y")
    _eq(fl, fl2, "flags")
    _eq(pw, pw2, "power")
    _eq(R, gpu.cov(xa2, CH, F, N), "R")


def test_lean_chain_without_aligned_copy(gpu, orc):
    """The chain with no aligned-copy buffer (Xa NULL): fused for 16 channels,
    the context's own copy for 32; both equal the oracle, and the
    allocation-free entry point refuses the unfused case without the buffer."""
    import rd_scenario

    for name, CH in (("medium", 16), ("medium", 32)):
        sc = rd_scenario.config(name)
        sc.CH, sc.F = CH, 6
        taps, W = sc.taps(), sc.weights()
        iq = sc.iq(seed=5)
        ch = rd_amd.Chain(gpu, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg())
        ch.load(iq, taps, W)
        exp = orc.chain(iq, taps, W, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg(), 31, 63,
                        keep=("power", "R", "events"))
        ch.enqueue(lean=True)
        gpu.sync()
        n = int(ch.total.download(np.uint64, 1)[0])
        got = ch.fetch(("power", "R"), n_events=n)
        assert n == exp["n_events"], CH
        _eq(got["power"], exp["power"], f"power CH={CH}")
        _eq(got["R"], exp["R"], f"R CH={CH}")
        _eq(got["events"], exp["events"], f"events CH={CH}")
        if CH > 16:
            with pytest.raises(RuntimeError):
                ch.enqueue(lean=True, static=True)


@pytest.mark.parametrize("name,CH,F", [("medium", 16, 6), ("medium", 8, 20), ("tiny", 4, 16)])
def test_chain_with_blocked_x(gpu, orc, name, CH, F):
    """The chain with X kept bin-quad blocked (x_layout 1): power, R and the
    events equal the oracle, and X read back in natural order equals the
    oracle's FFT output; full outputs, the lean graph and the allocation-free
    entry point alike."""
    import rd_scenario

    sc = rd_scenario.config(name)
    sc.CH, sc.F = CH, F
    taps, W = sc.taps(), sc.weights()
    iq = sc.iq(seed=9)
    exp = orc.chain(iq, taps, W, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg(), 31, 63,
                    keep=("X", "power", "R", "events"))
    ch = rd_amd.Chain(gpu, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg(), x_layout=1)
    ch.load(iq, taps, W)
    for how in ("run", "lean graph", "static"):
        ch.b["power"].upload(np.zeros(sc.B * sc.F * sc.N, np.uint32))
        if how == "run":
            n = ch.run()
        else:
            if how == "lean graph":
                g = ch.graph(lean=True)
                g.launch()
            else:
                ch.enqueue(static=True)
            gpu.sync()
            n = int(ch.total.download(np.uint64, 1)[0])
        got = ch.fetch(("X", "power", "R"), n_events=n)
        assert n == exp["n_events"], how
        for k in ("X", "power", "R", "events"):
            _eq(got[k], exp[k], f"{k} ({how}) CH={CH}")
        if how == "lean graph":
            g.close()


def test_blocked_x_only_where_the_chain_fuses(gpu):
    """x_layout 1 is refused where the chain does not fuse (more than 16
    channels), by the plain chain, and for layouts other than 0 and 1."""
    import rd_scenario

    sc = rd_scenario.config("medium")
    sc.F = 4
    for CH, layout, plain in ((32, 1, False), (16, 1, True), (16, 2, False)):
        ch = rd_amd.Chain(gpu, CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg(), x_layout=layout)
        ch.load(np.zeros(CH * sc.N * sc.F, np.uint32), sc.taps(), np.zeros(sc.B * CH, np.uint32))
        with pytest.raises(RuntimeError):
            ch.enqueue(plain=plain)
