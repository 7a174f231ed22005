"""The CPU restatement against the reference's own header models (oracle/_ref).

Skipped where the reference tree was absent at build time (a machine without
a prebuilt oracle/_ref).  Covers every stage on seeded random
and corner-value inputs inside the reference's domain (FFT N <= 1024), and the
composed chain on the tiny workload and a frame-reduced medium workload.
"""
import numpy as np
import pytest

import rd_scenario
import stim


def test_fxp_ops_random(orc, ref):
    rng = np.random.default_rng(21)
    ops = ["sat", "trunc", "round_even", "round_half_up", "round", "round_sat", "add_sat", "sub_sat", "neg_sat",
           "mul_q15", "mul_q15_rs", "cmul_re", "cmul_im", "acc_w", "growth_bits"]
    for op in ops:
        for _ in range(300):
            a = int(rng.integers(-(2**40), 2**40))
            b = int(rng.integers(-(2**40), 2**40))
            sh = int(rng.integers(0, 40))
            w = int(rng.integers(2, 64))
            assert orc.fxp_op(op, a, b, sh, w) == ref.fxp_op(op, a, b, sh, w), (op, a, b, sh, w)


def test_cmult_random_and_corners(orc, ref):
    rng = np.random.default_rng(22)
    vals = np.concatenate([stim.rand_cx(rng, 2000), stim.rand_extreme(rng, 2000)])
    for a, b in zip(vals[::2].tolist(), vals[1::2].tolist()):
        assert orc.cmult(a, b) == ref.cmult(a, b)


@pytest.mark.parametrize("P,T", [(4, 8), (64, 8), (1024, 16), (16, 1), (8, 3)])
def test_pfb_vs_ref(orc, ref, P, T):
    rng = np.random.default_rng(P * 100 + T)
    nb = 20
    for x, h in [(stim.rand_cx(rng, 2 * nb * P), stim.rand_cx(rng, P * T)),
                 (stim.rand_extreme(rng, 2 * nb * P), stim.rand_extreme(rng, P * T))]:
        yo, fo = orc.pfb(x, h, P, T, 2, nb)
        yr, fr = ref.pfb(x, h, P, T, 2, nb)
        assert np.array_equal(yo, yr) and np.array_equal(fo, fr)


@pytest.mark.parametrize("n", [8, 16, 32, 64, 128, 256, 512, 1024])
@pytest.mark.parametrize("spc", [1, 2])
def test_fft_vs_ref(orc, ref, n, spc):
    if n // spc < 4:
        pytest.skip("invalid geometry")
    rng = np.random.default_rng(n * 3 + spc)
    frames = 6
    xs = [stim.rand_cx(rng, frames * n), stim.rand_extreme(rng, frames * n), stim.rand_cx(rng, frames * n, 3000)]
    for x in xs:
        for sched in (0xFFFFFFFF, 0, int(rng.integers(0, 2**32)), 0x3):
            for reorder in (0, 1):
                o = orc.fft(x, n, spc, sched, reorder)
                r = ref.fft(x, n, spc, sched, reorder)
                for a, b in zip(o, r):
                    assert np.array_equal(a, b), (n, spc, hex(sched), reorder)


@pytest.mark.parametrize("CH,B", [(4, 4), (16, 16), (7, 3), (64, 8)])
def test_bf_vs_ref(orc, ref, CH, B):
    rng = np.random.default_rng(CH * 10 + B)
    cols = 96
    for Xa, W in [(stim.rand_cx(rng, cols * CH), stim.rand_cx(rng, B * CH)),
                  (stim.rand_extreme(rng, cols * CH), stim.rand_extreme(rng, B * CH))]:
        o = orc.bf(Xa, W, CH, B)
        r = ref.bf(Xa, W, CH, B)
        for a, b in zip(o, r):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("acc_w", [64, 40, 36])
def test_cov_vs_ref(orc, ref, acc_w):
    rng = np.random.default_rng(acc_w)
    CH, F, N = 5, 40, 6
    for Xa in (stim.rand_cx(rng, F * N * CH), stim.rand_extreme(rng, F * N * CH)):
        Ro, so = orc.cov(Xa, CH, F, N, acc_w)
        Rr, sr = ref.cov(Xa, CH, F, N, acc_w)
        assert np.array_equal(Ro, Rr) and np.array_equal(so, sr)


def test_power_cross_vs_ref(orc, ref):
    rng = np.random.default_rng(23)
    v = np.concatenate([stim.rand_cx(rng, 500), stim.rand_extreme(rng, 500)]).tolist()
    for a, b in zip(v[::2], v[1::2]):
        assert orc.power(a) == ref.power(a)
        assert orc.cross_power(a, b) == ref.cross_power(a, b)


CFAR_CFGS = [
    [1, 0, 0, 2, 2, 8, 8, 0x800],
    [1, 1, 0, 2, 3, 8, 4, 0x300],   # GO, asymmetric
    [1, 0, 1, 1, 1, 4, 4, 0x200],   # dense
    [1, 1, 1, 0, 0, 1, 0, 0x100],   # GO with an empty side: unusable geometry
    [0, 0, 0, 2, 2, 8, 8, 0x800],   # disabled
    [1, 0, 0, 40, 2, 80, 8, 0xFFFF],  # clamped geometry
    [1, 0, 0, 4, 4, 32, 32, 0x0800],
]


@pytest.mark.parametrize("cfg", CFAR_CFGS)
def test_cfar_vs_ref(orc, ref, cfg):
    rng = np.random.default_rng(sum(cfg))
    n, rows = 256, 6
    bins = rng.integers(0, 2**31, rows * n).astype(np.uint64)
    bins[rng.integers(0, rows * n, 40)] = 2**39
    bins[rng.integers(0, rows * n, 10)] = (1 << 63) + 12345  # u64 wrap-around paths
    eo, ro = orc.cfar(bins, n, cfg, 31, 63, frame_ids=np.arange(rows) * 7)
    er, rr = ref.cfar(bins, n, cfg, 31, 63, frame_ids=np.arange(rows) * 7)
    assert np.array_equal(ro, rr) and np.array_equal(eo, er)


@pytest.mark.parametrize("name,frames", [("tiny", None), ("medium", 12)])
def test_chain_vs_ref(orc, ref, name, frames):
    sc = rd_scenario.config(name)
    F = frames or sc.F
    sc.F = F
    iq = sc.iq()
    args = (iq, sc.taps(), sc.weights(), sc.CH, sc.N, sc.T, F, sc.B, sc.cfar_cfg(), rd_scenario.MAX_GUARD,
            rd_scenario.MAX_REF)
    keep = ("u", "pfb_flags", "X", "fft_flags", "Xa", "y", "yflags", "power", "R", "events")
    o = orc.chain(*args, keep=keep, threads=8)
    r = ref.chain(*args, keep=keep, threads=8)
    for k in keep + ("row_count",):
        assert np.array_equal(o[k], r[k]), k
    assert o["n_events"] == r["n_events"] >= sc.B * F


@pytest.mark.parametrize("packet_w", [36, 64, 100, 128, 176, 177, 256, 512])
def test_event_flits_vs_ref(orc, ref, packet_w):
    """Detection packets: the restatement against packet_model::hdr_pack /
    make_flit, on random events with the sof/eof bits above bit 176 set (they
    are not payload) and a sequence number that wraps."""
    rng = np.random.default_rng(packet_w)
    ev = rng.integers(0, 2**63, size=(40, 3), dtype=np.uint64) | (
        rng.integers(0, 2, size=(40, 3), dtype=np.uint64) << np.uint64(63))
    ev[:, 2] |= np.uint64(3) << np.uint64(48)
    for dest, src, vc, seq0 in [(0, 0, 0, 0), (15, 31, 3, 65530), (9, 17, 2, 1234)]:
        a = orc.event_flits(ev, packet_w, dest, src, vc, seq0)
        b = ref.event_flits(ev, packet_w, dest, src, vc, seq0)
        assert np.array_equal(a, b), (packet_w, dest, src, vc, seq0)


@pytest.mark.parametrize("packet_w", [36, 64, 177, 512])
def test_event_writeback_vs_scoreboard(orc, ref, packet_w):
    """The writeback's receive checks against packet_model::Scoreboard
    (packet_model.hpp:360-429) at egress `port`: a clean stream is delivered
    whole with no error of any kind, and for each single fault whose count
    the Scoreboard takes per packet (a payload bit, the parity bit, the
    header's type, vc, dest, length or seq) the parity, framing and misrouted
    counts agree, and a broken seq is the Scoreboard's unknown-seq count.
    Faults in sof/eof/vc bits make the Scoreboard re-frame the stream, which
    fixed packet slots do not, so they are pinned by the closed-form counts
    instead."""
    import flitfault

    rng = np.random.default_rng(packet_w + 11)
    n, port, src, vc, seq0 = 12, 6, 19, 2, 65533
    ev = rng.integers(0, 2**63, size=(n, 3), dtype=np.uint64)
    fl = orc.event_flits(ev, packet_w, dest=port, src=src, vc=vc, seq0=seq0)
    sb = ref.scoreboard(fl, packet_w, port, ev, dest=port, src=src, vc=vc, seq0=seq0)
    ring, st = orc.event_writeback(fl, packet_w, port=port, seq0=seq0)
    assert sb["delivered"] == sb["injected"] == n and sum(sb.values()) == 2 * n
    assert st.tolist() == [n, 0, 0, 0, 0]
    assert np.array_equal(ring, flitfault.ring_of_events(ev))
    for kind in flitfault.AGREEING:
        bad = flitfault.plant(fl, 5, kind, packet_w)
        sb = ref.scoreboard(bad, packet_w, port, ev, dest=port, src=src, vc=vc, seq0=seq0)
        _, st = orc.event_writeback(bad, packet_w, port=port, seq0=seq0)
        got = (int(st[1]), int(st[2]), int(st[3]), int(st[4]))
        assert got == (sb["parity"], sb["framing"], sb["misrouted"], sb["duplicated"]), (kind, sb)
        assert sum(got) == 1, kind


@pytest.mark.parametrize("mode,exp_k,acc_w", [(1, 0, 64), (1, 1, 64), (1, 5, 40), (1, 15, 64), (1, 3, 20),
                                              (0, 0, 33), (0, 0, 64)])
def test_cov_integrate_vs_ref(orc, ref, mode, exp_k, acc_w):
    rng = np.random.default_rng(mode * 100 + exp_k + acc_w)
    CH, F, N = 5, 23, 7
    for Xa in (stim.rand_cx(rng, F * N * CH), stim.rand_extreme(rng, F * N * CH)):
        a = orc.cov_integrate(Xa, CH, F, N, mode, exp_k, acc_w)
        b = ref.cov_integrate(Xa, CH, F, N, mode, exp_k, acc_w)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("n,spc", [(16, 1), (64, 2), (256, 1), (1024, 2)])
def test_fft_twiddle_saturation_vs_ref(orc, ref, n, spc):
    """Twiddle products that saturate under the all-ones schedule (quarters
    0/2 at (-32768, -32768), 1/3 at (32767, 32767)): sat_neg of sub-stage 1
    can only come from a twiddle product there."""
    q = (np.arange(n) // (n // 4)) & 1
    v = np.where(q == 0, -32768, 32767)
    x = stim.pack(v, v)
    o = orc.fft(x, n, spc, 0xFFFFFFFF, 1)
    r = ref.fft(x, n, spc, 0xFFFFFFFF, 1)
    assert (int(o[2][0]) >> 2) & 1
    for a, b in zip(o, r):
        assert np.array_equal(a, b)
