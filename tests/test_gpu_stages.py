"""GPU parity: every HIP stage, called through the C ABI, against the oracle.

Bit-exact on every output (values, flags, exact accumulators, packed events),
on the reference's golden vectors, seeded random data, corner values that drive
every saturation path, and the edge geometries the reference tests.
"""
import numpy as np
import pytest

import stim
import vecio

pytestmark = pytest.mark.gpu


def _eq(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    bad = np.flatnonzero(a != b)
    assert bad.size == 0, f"{what}: {bad.size} mismatches, first at {bad[:5]}: {a[bad[:5]]} vs {b[bad[:5]]}"


# ---------------------------------------------------------------------------- PFB
@pytest.mark.parametrize("set_name,geom", vecio.PFB_SETS)
def test_pfb_golden(gpu, set_name, geom):
    P, T, coeff, v = vecio.load_pfb(set_name, geom)
    y, fl = gpu.pfb(v["x"], coeff, P, T, 1, v["beats"])
    _eq(y, v["y"], "y")
    _eq(fl & 3, v["f_re"], "f_re")
    _eq(fl >> 2, v["f_im"], "f_im")
    _eq(gpu.pfb(v["x"], coeff, P, T, 1, v["beats"], flags=False)[0], v["y"], "y without the flag buffer")


@pytest.mark.parametrize("P,T,nb,ch", [(64, 8, 16, 4), (1024, 16, 40, 3), (4096, 16, 20, 2), (256, 4, 33, 2),
                                       (128, 5, 17, 2), (64, 1, 9, 1), (1000, 16, 70, 2), (32, 32, 12, 2)])
def test_pfb_random_and_extreme(gpu, orc, P, T, nb, ch):
    rng = np.random.default_rng(P + T + nb)
    for x, h in [(stim.rand_cx(rng, ch * nb * P), stim.rand_cx(rng, P * T)),
                 (stim.rand_extreme(rng, ch * nb * P), stim.rand_extreme(rng, P * T))]:
        yg, fg = gpu.pfb(x, h, P, T, ch, nb)
        yo, fo = orc.pfb(x, h, P, T, ch, nb)
        _eq(yg, yo, "y")
        _eq(fg, fo, "flags")
        _eq(gpu.pfb(x, h, P, T, ch, nb, flags=False)[0], yo, "y without the flag buffer")


@pytest.mark.parametrize("P,T,nb,ch", [(32, 16, 257, 2), (64, 15, 255, 1), (96, 16, 513, 1), (32, 1, 3, 3),
                                       (128, 9, 256, 2), (2048, 16, 7, 1)])
def test_pfb_beat_counts_and_history(gpu, orc, P, T, nb, ch):
    """Beat counts around powers of two (255/256/257/513), a run shorter than
    T, one tap, random and extreme values with and without the flag buffer,
    and a random extreme imported history whose final state is compared too."""
    rng = np.random.default_rng(P * 7 + T + nb)
    for x, h in [(stim.rand_cx(rng, ch * nb * P), stim.rand_cx(rng, P * T)),
                 (stim.rand_extreme(rng, ch * nb * P), stim.rand_extreme(rng, P * T))]:
        yg, fg = gpu.pfb(x, h, P, T, ch, nb)
        yo, fo = orc.pfb(x, h, P, T, ch, nb)
        _eq(yg, yo, "y")
        _eq(fg, fo, "flags")
        _eq(gpu.pfb(x, h, P, T, ch, nb, flags=False)[0], yo, "y without the flag buffer")
        if T > 1:
            hist0 = stim.rand_extreme(rng, ch * (T - 1) * P)
            hg, ho = hist0.copy(), hist0.copy()
            yg, fg = gpu.pfb(x, h, P, T, ch, nb, hist=hg)
            yo, fo = orc.pfb(x, h, P, T, ch, nb, hist=ho)
            _eq(yg, yo, "y with history")
            _eq(fg, fo, "flags with history")
            _eq(hg, ho, "history out")


def test_pfb_long_run_chunks(gpu, orc):
    """Many beats per channel so the launch splits beats into chunks with warm-up."""
    rng = np.random.default_rng(5)
    P, T, nb, ch = 256, 16, 600, 1
    x, h = stim.rand_cx(rng, ch * nb * P), stim.rand_cx(rng, P * T)
    _eq(gpu.pfb(x, h, P, T, ch, nb)[0], orc.pfb(x, h, P, T, ch, nb)[0], "y")


@pytest.mark.parametrize("P,T,nb,ch", [(1024, 16, 40, 2), (256, 8, 33, 2), (128, 4, 257, 1), (2048, 16, 600, 1),
                                       (96, 16, 7, 2)])
@pytest.mark.parametrize("mix", ["real", "mixed"])
def test_pfb_real_taps(gpu, orc, P, T, nb, ch, mix):
    """Taps with zero imaginary parts take the paired-tap window per wave; in
    "mixed" a few branches keep complex taps, so their waves take the complex
    window while the others stay real. Random and extreme values (-32768 in
    taps and samples), with and without flags, with an imported history, and
    a beat count that splits into chunks."""
    rng = np.random.default_rng(P * 3 + T + nb)
    for x, h in [(stim.rand_cx(rng, ch * nb * P), stim.rand_cx(rng, P * T)),
                 (stim.rand_extreme(rng, ch * nb * P), stim.rand_extreme(rng, P * T))]:
        h = h & np.uint32(0xFFFF)  # real taps
        if mix == "mixed":
            hb = h.reshape(P, T)
            hb[::37] = stim.rand_extreme(rng, hb[::37].size).reshape(hb[::37].shape)
        yg, fg = gpu.pfb(x, h, P, T, ch, nb)
        yo, fo = orc.pfb(x, h, P, T, ch, nb)
        _eq(yg, yo, "y")
        _eq(fg, fo, "flags")
        _eq(gpu.pfb(x, h, P, T, ch, nb, flags=False)[0], yo, "y without the flag buffer")
        hist0 = stim.rand_extreme(rng, ch * (T - 1) * P)
        hg, ho = hist0.copy(), hist0.copy()
        yg, fg = gpu.pfb(x, h, P, T, ch, nb, hist=hg)
        yo, fo = orc.pfb(x, h, P, T, ch, nb, hist=ho)
        _eq(yg, yo, "y with history")
        _eq(hg, ho, "final history")


def _tap_magnitudes(rng, P, T, total):
    """P rows of T magnitudes in [1, 32767], each row summing to `total`."""
    m = np.full((P, T), total // T, np.int64)
    m[:, : total % T] += 1
    for _ in range(4 * T):
        i, j = rng.integers(0, T, 2)
        if i == j:
            continue
        d = np.minimum(np.minimum(rng.integers(0, 2000, P), m[:, i] - 1), 32767 - m[:, j])
        m[:, i] -= d
        m[:, j] += d
    assert (m.sum(1) == total).all() and m.min() >= 1 and m.max() <= 32767
    return m


@pytest.mark.parametrize("T", [16, 8, 4])
def test_pfb_direct_window_at_its_bound(gpu, orc, T):
    """Real taps whose magnitudes sum to exactly 65535 on every branch of the
    first wave (the direct window: no digit split, int32 sums up to
    65535 * 32768 = 2^31 - 2^15 in magnitude) and to 65536 on one branch of
    the second wave (the whole wave keeps the digit window). Taps all
    positive, all negative and of random signs against samples all -32768
    (the largest sums, saturating), all 32767, extreme and random; with and
    without flags, with an imported history, beats split into chunks."""
    P, nb, ch = 128, 70, 2
    rng = np.random.default_rng(T * 11)
    m = _tap_magnitudes(rng, P, T, 65535)
    m[100, int(np.argmin(m[100]))] += 1  # branch 100 (second wave): 65536
    assert m[:64].sum(1).max() == 65535 and m[100].sum() == 65536
    signs = {"positive": np.ones((P, T), np.int64), "negative": -np.ones((P, T), np.int64),
             "random": rng.choice([-1, 1], (P, T))}
    S = ch * nb * P
    samples = {"all -32768": np.full(S, vecio.pack(-32768, -32768), np.uint32),
               "all 32767": np.full(S, vecio.pack(32767, 32767), np.uint32),
               "extreme": stim.rand_extreme(rng, S), "random": stim.rand_cx(rng, S)}
    for sn, sg in signs.items():
        h = ((m * sg).reshape(-1) & 0xFFFF).astype(np.uint32)  # real taps
        for xn, x in samples.items():
            yg, fg = gpu.pfb(x, h, P, T, ch, nb)
            yo, fo = orc.pfb(x, h, P, T, ch, nb)
            _eq(yg, yo, f"y taps {sn} samples {xn}")
            _eq(fg, fo, f"flags taps {sn} samples {xn}")
            _eq(gpu.pfb(x, h, P, T, ch, nb, flags=False)[0], yo, f"y without flags, taps {sn} samples {xn}")
        hist0 = stim.rand_extreme(rng, ch * (T - 1) * P)
        hg, ho = hist0.copy(), hist0.copy()
        x = samples["extreme"]
        _eq(gpu.pfb(x, h, P, T, ch, nb, hist=hg)[0], orc.pfb(x, h, P, T, ch, nb, hist=ho)[0], f"y with history {sn}")
        _eq(hg, ho, f"final history {sn}")


@pytest.mark.parametrize("T", [16, 5])
def test_pfb_history_carry(gpu, orc, T):
    rng = np.random.default_rng(T)
    P, nb, ch = 128, 30, 3
    x, h = stim.rand_cx(rng, ch * nb * P), stim.rand_cx(rng, P * T)
    y_all, f_all = orc.pfb(x, h, P, T, ch, nb)
    xs = x.reshape(ch, nb * P)
    hist = np.zeros(ch * (T - 1) * P, np.uint32)
    outs = []
    for lo, hi in [(0, 2), (2, 13), (13, 30)]:  # includes a call shorter than T-1 beats
        seg = np.ascontiguousarray(xs[:, lo * P: hi * P]).reshape(-1)
        y, _ = gpu.pfb(seg, h, P, T, ch, hi - lo, hist=hist)
        outs.append(y.reshape(ch, -1))
    _eq(np.concatenate(outs, axis=1).reshape(-1), y_all, "y across calls")


# ---------------------------------------------------------------------------- FFT
def test_fft_golden(gpu):
    hdr, recs = vecio.load_fft()
    for r in recs:
        out, bins, sf = gpu.fft(r["inp"], hdr["n"], hdr["spc"], r["sched"], r["reorder"])
        _eq(out, r["out"], r["id"])
        assert int(sf[0]) == r["flags"], r["id"]


@pytest.mark.parametrize("n", [8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096])
@pytest.mark.parametrize("spc", [1, 2])
def test_fft_random(gpu, orc, n, spc):
    if n // spc < 4:
        pytest.skip("geometry rejected by the reference")
    rng = np.random.default_rng(n * 7 + spc)
    frames = max(3, 4096 // n)
    xs = [stim.rand_cx(rng, frames * n), stim.rand_extreme(rng, frames * n), stim.rand_cx(rng, frames * n, 2000)]
    for x in xs:
        for sched in (0xFFFFFFFF, 0, int(rng.integers(0, 2**32))):
            for reorder in (1, 0):
                g = gpu.fft(x, n, spc, sched, reorder)
                o = orc.fft(x, n, spc, sched, reorder)
                for a, b, w in zip(g, o, ("out", "bins", "stage_flags")):
                    _eq(a, b, f"{w} n={n} spc={spc} sched={sched:#x} reorder={reorder}")
                # without a stage-flag buffer (the all-ones path then forms no flags)
                g = gpu.fft(x, n, spc, sched, reorder, flags=False)
                for a, b, w in zip(g[:2], o[:2], ("out", "bins")):
                    _eq(a, b, f"{w} without flags n={n} spc={spc} sched={sched:#x} reorder={reorder}")


@pytest.mark.parametrize("n,spc,CH,F", [(8, 2, 3, 5), (16, 1, 2, 7), (64, 2, 4, 9), (1024, 2, 3, 6),
                                         (4096, 2, 2, 3), (256, 1, 5, 2)])
def test_fft_stage_layouts(gpu, orc, n, spc, CH, F):
    """rd_fft_stage: X in natural order equals rd_fft_run's natural bins, and
    the bin-quad blocked X [CH][n/4][F][4] holds the same words; under the
    all-ones schedule (16 points per thread from 64 points) and others (4
    points per thread), with and without stage flags."""
    import rd_amd

    rng = np.random.default_rng(n + CH * 31 + F)
    x = stim.rand_extreme(rng, CH * F * n)
    for sched in (0xFFFFFFFF, int(rng.integers(0, 2**32))):
        o = orc.fft(x, n, spc, sched, 1)
        for layout in (0, 1):
            for flags in (True, False):
                X, sf = gpu.fft_stage(x, n, CH, F, x_layout=layout, spc=spc, sched=sched, flags=flags)
                _eq(X, o[1], f"X layout={layout} n={n} spc={spc} sched={sched:#x} flags={flags}")
                if flags:
                    _eq(sf, o[2], f"stage flags layout={layout} n={n}")
        Xb = rd_amd.natural_x(gpu.fft_stage(x, n, CH, F, x_layout=0, spc=spc, sched=sched)[0], CH, F, n)
        _eq(rd_amd.blocked_x(Xb, CH, F, n), o[1], "layout helpers round trip")


def twiddle_saturation_frame(n):
    """A frame whose first radix-4 pass under the all-ones schedule leaves
    (-32768, -32768) at quarter 1 of every group, so the twiddles near 45 deg
    saturate their products negatively: quarters 0 and 2 hold (-32768, -32768),
    quarters 1 and 3 hold (32767, 32767) (for spc = 2 each lane's quarters are
    the frame's quarters)."""
    q = (np.arange(n) // (n // 4)) & 1
    re = np.where(q == 0, -32768, 32767)
    return stim.pack(re, re)


@pytest.mark.parametrize("n,spc", [(16, 1), (64, 2), (1024, 2), (1024, 1), (4096, 2)])
def test_fft_twiddle_saturation(gpu, orc, n, spc):
    x = np.concatenate([twiddle_saturation_frame(n), stim.rand_extreme(np.random.default_rng(n), n)])
    o = orc.fft(x, n, spc, 0xFFFFFFFF, 1)
    assert (int(o[2][0]) >> 2) & 1, "sub-stage 1 must saturate negatively (a twiddle product)"
    g = gpu.fft(x, n, spc, 0xFFFFFFFF, 1)
    for a, b, w in zip(g, o, ("out", "bins", "stage_flags")):
        _eq(a, b, f"{w} n={n} spc={spc}")
    g = gpu.fft(x, n, spc, 0xFFFFFFFF, 1, flags=False)
    for a, b, w in zip(g[:2], o[:2], ("out", "bins")):
        _eq(a, b, f"{w} without flags n={n} spc={spc}")


def test_fft_rejects_invalid_geometry(gpu):
    import rd_amd

    for n, spc in [(4, 1), (8, 4), (6, 1), (8192, 2)]:
        with pytest.raises(rd_amd.RdError) as e:
            gpu.fft(np.zeros(max(n, 8) * 2, np.uint32), n, spc)
        assert e.value.code == rd_amd.RD_E_GEOM


def test_fft_rejects_misaligned_buffers(gpu):
    import rd_amd

    d = gpu.buf(4 * 1024 + 64)
    for i, o, b in [(d.ptr + 4, d.ptr, d.ptr), (d.ptr, d.ptr + 4, None), (d.ptr, None, d.ptr + 8)]:
        rc = gpu.L.rd_fft_run(gpu.h, i, 256, 2, 0xFFFFFFFF, 1, 4, o, b, None)
        assert rc == rd_amd.RD_E_ARG, (i - d.ptr, o and o - d.ptr, b and b - d.ptr)


# ---------------------------------------------------------------------------- align
def test_align(gpu, orc):
    rng = np.random.default_rng(1)
    for CH, F, N in [(4, 16, 64), (16, 3, 1024), (7, 5, 100), (64, 2, 4096)]:
        X = stim.rand_cx(rng, CH * F * N)
        _eq(gpu.align(X, CH, F, N), orc.align(X, CH, F, N), f"align {CH}x{F}x{N}")


# ---------------------------------------------------------------------------- beamformer
@pytest.mark.parametrize("CH,B", [(4, 4), (16, 16), (64, 64), (7, 3), (33, 17), (64, 128), (1, 1), (48, 20),
                                  (32, 32), (17, 5), (32, 128)])
def test_bf_random_and_extreme(gpu, orc, CH, B):
    rng = np.random.default_rng(CH * 131 + B)
    cols = 1000
    for Xa, W in [(stim.rand_cx(rng, cols * CH), stim.rand_cx(rng, B * CH)),
                  (stim.rand_extreme(rng, cols * CH), stim.rand_extreme(rng, B * CH)),
                  (np.full(cols * CH, vecio.pack(-32768, -32768), np.uint32),
                   np.full(B * CH, vecio.pack(-32768, -32768), np.uint32))]:
        g = gpu.bf(Xa, W, CH, B)
        o = orc.bf(Xa, W, CH, B)
        for a, b, w in zip(g, o, ("y", "flags", "power", "acc")):
            _eq(a, b, f"{w} CH={CH} B={B}")


def test_bf_unit_weights_passthrough(gpu):
    rng = np.random.default_rng(2)
    CH, B, cols = 16, 16, 512
    Xa = stim.rand_cx(rng, cols * CH)
    W = np.zeros(B * CH, np.uint32)
    for b in range(B):
        W[b * CH + (b * 5) % CH] = vecio.pack(32767, 0)
    y, fl, _, _ = gpu.bf(Xa, W, CH, B)
    x = Xa.reshape(cols, CH)
    for b in range(B):
        col = x[:, (b * 5) % CH]
        exp = [vecio.pack(stim.rne_int(vecio.unpack(v)[0] * 32767, 15), stim.rne_int(vecio.unpack(v)[1] * 32767, 15))
               for v in col.tolist()]
        _eq(y[b * cols:(b + 1) * cols], np.array(exp, np.uint32), f"beam {b}")
    assert not fl.any()


# (64, 64, 2, 4096), (48, 32, 3, 256), (64, 16, 1, 128), (64, 16, 64, 4096)
# and (64, 128, 2, 512) take the loader-fed kernel (48 or 64 channels, beams a
# multiple of 16, columns a multiple of 128): one tile, a few, eight tiles per
# block, and 128 beams (eight beam tiles held in LDS)
@pytest.mark.parametrize("CH,B,F,N", [(16, 16, 5, 1024), (64, 64, 2, 4096), (4, 4, 16, 64), (7, 3, 3, 100),
                                      (33, 20, 2, 512), (48, 17, 1, 256), (32, 24, 3, 256), (16, 16, 2, 64),
                                      (48, 32, 3, 256), (64, 16, 1, 128), (64, 16, 64, 4096), (32, 32, 4, 256),
                                      (64, 128, 2, 512)])
def test_bf_stage_matches_align_then_bf(gpu, orc, CH, B, F, N):
    rng = np.random.default_rng(CH + B + F + N)
    for X, W in [(stim.rand_cx(rng, CH * F * N), stim.rand_cx(rng, B * CH)),
                 (stim.rand_extreme(rng, CH * F * N), stim.rand_extreme(rng, B * CH)),
                 (np.full(CH * F * N, vecio.pack(-32768, -32768), np.uint32),
                  np.full(B * CH, vecio.pack(-32768, -32768), np.uint32))]:
        y, fl, pw, Xa = gpu.bf_stage(X, W, CH, B, F, N)
        xa = orc.align(X, CH, F, N)
        _eq(Xa, xa, "aligned copy")
        oy, ofl, opw, _ = orc.bf(xa, W, CH, B, want_acc=False)
        _eq(y, oy, "y")
        _eq(fl, ofl, "flags")
        _eq(pw, opw, "power")


# ---------------------------------------------------------------------------- covariance
# 32 and 64 channels with the bin count a multiple of 64 / CH take the
# loader-fed kernel: (64, 20, 5), (64, 512, 2), (32, 40, 6), (32, 1, 2)
@pytest.mark.parametrize("CH,F,N", [(4, 16, 64), (16, 40, 33), (64, 20, 5), (1, 7, 9), (20, 33, 8), (33, 17, 6),
                                    (64, 512, 2), (16, 256, 12), (32, 40, 6), (16, 33, 4), (32, 1, 2)])
def test_cov_random_and_extreme(gpu, orc, CH, F, N):
    rng = np.random.default_rng(CH * 7 + F + N)
    for Xa in (stim.rand_cx(rng, F * N * CH), stim.rand_extreme(rng, F * N * CH),
               np.full(F * N * CH, vecio.pack(-32768, -32768), np.uint32)):
        _eq(gpu.cov(Xa, CH, F, N), orc.cov(Xa, CH, F, N)[0], f"R CH={CH} F={F} N={N}")


# ---------------------------------------------------------------------------- CFAR
CFAR_CFGS = [
    [1, 0, 0, 2, 2, 8, 8, 0x800],
    [1, 1, 0, 2, 3, 8, 4, 0x300],
    [1, 0, 1, 1, 1, 4, 4, 0x200],
    [1, 1, 1, 0, 0, 1, 0, 0x100],
    [0, 0, 0, 2, 2, 8, 8, 0x800],
    [1, 0, 0, 40, 2, 80, 8, 0xFFFF],
    [1, 0, 0, 4, 4, 32, 32, 0x0800],
    [1, 0, 0, 0, 0, 0, 0, 0x0800],
    [1, 0, 1, 4, 4, 16, 16, 0x0800],
    # alpha 1.0 in the events mode: about a third of the bins detect, so waves
    # of 128 bins and more emit over 64 events (the chunk-by-chunk emission)
    [1, 0, 0, 1, 1, 4, 4, 0x100],
    [1, 1, 0, 1, 1, 4, 4, 0x100],
]


@pytest.mark.parametrize("cfg", CFAR_CFGS)
@pytest.mark.parametrize("n", [64, 1024, 4096, 100, 1, 7, 2049, 5000, 8191])
def test_cfar_u64(gpu, orc, cfg, n):
    rng = np.random.default_rng(n + sum(cfg))
    rows = 5
    cells = rng.integers(0, 2**31, rows * n).astype(np.uint64)
    cells[rng.integers(0, rows * n, 30)] = 2**39
    cells[rng.integers(0, rows * n, 5)] = (1 << 63) + 77
    fid = (np.arange(rows) * 13 + 2).astype(np.uint32)
    eg, rcg, tot, _ = gpu.cfar(cells, n, cfg, 31, 63, frame_ids=fid)
    eo, rco = orc.cfar(cells, n, cfg, 31, 63, frame_ids=fid)
    _eq(rcg, rco, "row_count")
    assert tot == eo.shape[0]
    _eq(eg, eo, "events")


def test_cfar_many_rows_across_calls(gpu, orc):
    """Thousands of rows (hundreds of row tiles, whose event offsets chain from
    tile to tile) in consecutive calls of different sizes on one context, so
    each call starts from the state the previous one left."""
    rng = np.random.default_rng(11)
    cfg = [1, 0, 0, 2, 2, 8, 8, 0x500]
    for rows, n in [(5000, 100), (1200, 64), (5003, 100), (300, 4096), (4096, 1024)]:
        cells = rng.integers(0, 2**31, rows * n).astype(np.uint32)
        eg, rcg, tot, _ = gpu.cfar(cells, n, cfg, 31, 63, frame_mod=7)
        eo, rco = orc.cfar(cells.astype(np.uint64), n, cfg, 31, 63, frame_ids=np.arange(rows) % 7)
        _eq(rcg, rco, f"row_count rows={rows} n={n}")
        assert tot == eo.shape[0]
        _eq(eg, eo, f"events rows={rows} n={n}")


@pytest.mark.parametrize("cfg", [CFAR_CFGS[0], CFAR_CFGS[5], CFAR_CFGS[6], CFAR_CFGS[8], CFAR_CFGS[9]])
@pytest.mark.parametrize("n", [64, 1024, 4096, 100, 5000, 2])
def test_cfar_u32_rows_either_side_of_2_32(gpu, orc, cfg, n):
    """u32 cells in rows whose total is below 2^32, exactly 2^32 - 1, exactly
    2^32 and far above, with planted targets, against the oracle: the 64-bit
    prefix carries past 2^32 inside a row."""
    rng = np.random.default_rng(n * 31 + sum(cfg))
    rows = 16
    cells = np.zeros((rows, n), np.uint64)
    for r in range(rows):
        kind = r % 4
        if kind == 3:
            cells[r] = rng.integers(0, 2**31, n)
            continue
        cells[r] = rng.integers(0, 2**16, n)
        cells[r, rng.integers(0, n, 3)] = rng.integers(2**22, 2**24, 3)  # targets
        if kind == 1:  # total 2^32 - 1
            cells[r, -1] = 0
            cells[r, -1] = 2**32 - 1 - int(cells[r].sum())
        elif kind == 2:  # total 2^32: the last two cells carry it
            cells[r, -2:] = 0
            rest = 2**32 - int(cells[r].sum())
            cells[r, -2], cells[r, -1] = rest // 2, rest - rest // 2
    assert all(int(cells[r].sum()) == 2**32 - 1 for r in range(1, rows, 4))
    assert all(int(cells[r].sum()) == 2**32 for r in range(2, rows, 4))
    c32 = cells.reshape(-1).astype(np.uint32)
    assert np.array_equal(c32.astype(np.uint64), cells.reshape(-1))
    eg, rcg, tot, _ = gpu.cfar(c32, n, cfg, 31, 63, frame_mod=5)
    eo, rco = orc.cfar(cells.reshape(-1), n, cfg, 31, 63, frame_ids=np.arange(rows) % 5)
    _eq(rcg, rco, "row_count")
    assert tot == eo.shape[0]
    _eq(eg, eo, "events")


def test_cfar_u32_frame_mod_and_capacity(gpu, orc):
    import rd_amd

    rng = np.random.default_rng(4)
    n, rows, F = 1024, 12, 4
    cells32 = rng.integers(0, 2**31, rows * n).astype(np.uint32)
    cells32[rng.integers(0, rows * n, 50)] = 2**31
    cfg = [1, 0, 0, 4, 4, 16, 16, 0x600]
    eg, rcg, tot, rc = gpu.cfar(cells32, n, cfg, 31, 63, frame_mod=F)
    eo, rco = orc.cfar(cells32.astype(np.uint64), n, cfg, 31, 63, frame_ids=np.arange(rows) % F)
    _eq(eg, eo, "events")
    assert rc == rd_amd.RD_OK
    cap = tot - 3
    eg2, _, tot2, rc2 = gpu.cfar(cells32, n, cfg, 31, 63, frame_mod=F, cap=cap)
    assert rc2 == rd_amd.RD_E_CAPACITY and tot2 == tot
    _eq(eg2, eo[:cap], "events under capacity")


@pytest.mark.parametrize("set_name,geom", vecio.PFB_SETS)
def test_pfb_fed_from_reference_files(gpu, set_name, geom):
    """The GPU path ingests the reference's .coeff/.vec files through the
    library's own readers and reproduces the committed outputs."""
    lib = gpu.lib
    coeff, P, T = lib.read_coeff(vecio.VEC_DIR / f"pfb_{set_name}_{geom}.coeff")
    v = lib.read_pfb_vec(vecio.VEC_DIR / f"pfb_{set_name}_{geom}.vec")
    y, fl = gpu.pfb(v["x"], coeff, P, T, 1, v["beats"])
    _eq(y, v["y"], "y")
    _eq(fl & 3, v["f_re"], "f_re")
    _eq(fl >> 2, v["f_im"], "f_im")


@pytest.mark.parametrize("mode,exp_k,acc_w", [(1, 0, 64), (1, 4, 64), (1, 15, 40), (1, 2, 24), (0, 0, 36), (0, 0, 64)])
@pytest.mark.parametrize("CH,F,N", [(4, 40, 16), (17, 9, 5), (64, 3, 2)])
def test_cov_integrate(gpu, orc, mode, exp_k, acc_w, CH, F, N):
    rng = np.random.default_rng(CH + F + N + exp_k + acc_w)
    for Xa in (stim.rand_cx(rng, F * N * CH), stim.rand_extreme(rng, F * N * CH)):
        R, fl = gpu.cov_integrate(Xa, CH, F, N, mode, exp_k, acc_w)
        oR, ofl = orc.cov_integrate(Xa, CH, F, N, mode, exp_k, acc_w)
        _eq(R, oR, "R")
        _eq(fl, ofl, "sat flags")
    if mode == 0 and acc_w == 64:
        _eq(R, gpu.cov(Xa, CH, F, N), "block mode = the MFMA contraction")
