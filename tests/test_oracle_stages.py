"""Oracle stages pinned by the reference's vectors, closed forms and invariants.

Mirrors: the PFB vector check (pfb_model.hpp loaders + test_pfb_bank.cpp),
model/cpp/test/test_fft_ref.cpp (vectors, permutation property, axis twiddles),
test_beamformer.cpp's directed cases (unit/zero/permutation), the covariance
identities of DECISIONS.md (X=Y, swap==direct, Hermitian), and test_cfar.cpp's
flat-spectrum and edge-suppression cases.
"""
import numpy as np
import pytest

import flitfault
import stim
import vecio


@pytest.mark.parametrize("set_name,geom", vecio.PFB_SETS)
def test_pfb_golden_vectors(orc, set_name, geom):
    P, T, coeff, v = vecio.load_pfb(set_name, geom)
    y, fl = orc.pfb(v["x"], coeff, P, T, 1, v["beats"])
    assert np.array_equal(y, v["y"])
    assert np.array_equal(fl & 3, v["f_re"])
    assert np.array_equal(fl >> 2, v["f_im"])


def test_pfb_planted_tap_order_fault_is_caught(orc):
    P, T, coeff, v = vecio.load_pfb("random", "p8t16")
    orc.lib.rdo_set_fault(4)
    try:
        y, _ = orc.pfb(v["x"], coeff, P, T, 1, v["beats"])
    finally:
        orc.lib.rdo_set_fault(0)
    assert not np.array_equal(y, v["y"])


def test_pfb_history_split_equals_one_run(orc):
    rng = np.random.default_rng(3)
    P, T, nb = 16, 8, 24
    x = stim.rand_cx(rng, nb * P)
    h = stim.rand_cx(rng, P * T)
    y1, f1 = orc.pfb(x, h, P, T, 1, nb)
    hist = np.zeros((T - 1) * P, np.uint32)
    ya, fa = orc.pfb(x[: 10 * P], h, P, T, 1, 10, hist=hist)
    yb, fb = orc.pfb(x[10 * P:], h, P, T, 1, nb - 10, hist=hist)
    assert np.array_equal(np.concatenate([ya, yb]), y1)
    assert np.array_equal(np.concatenate([fa, fb]), f1)


def test_fft_golden_vectors(orc):
    hdr, recs = vecio.load_fft()
    assert len(recs) == 36 and hdr["n"] == 64 and hdr["spc"] == 2
    for r in recs:
        out, bins, sf = orc.fft(r["inp"], hdr["n"], hdr["spc"], r["sched"], r["reorder"])
        assert np.array_equal(out, r["out"]), r["id"]
        assert int(sf[0]) == r["flags"], r["id"]


@pytest.mark.parametrize("fault", [1, 2, 3])
def test_fft_planted_faults_are_caught(orc, fault):
    hdr, recs = vecio.load_fft()
    orc.lib.rdo_set_fault(fault)
    try:
        bad = 0
        for r in recs:
            out, _, sf = orc.fft(r["inp"], hdr["n"], hdr["spc"], r["sched"], r["reorder"])
            bad += (not np.array_equal(out, r["out"])) or int(sf[0]) != r["flags"]
    finally:
        orc.lib.rdo_set_fault(0)
    assert bad > 0


@pytest.mark.parametrize("n,spc", [(64, 2), (256, 1), (1024, 2), (4096, 2), (4096, 1)])
def test_fft_permutation_and_bins(orc, n, spc):
    rng = np.random.default_rng(n + spc)
    x = stim.rand_cx(rng, n)
    nat, bins_a, _ = orc.fft(x, n, spc, reorder=1)
    rev, bins_b, _ = orc.fft(x, n, spc, reorder=0)
    m, nl = n // spc, (n // spc).bit_length() - 1
    br = np.array([int(format(j, f"0{nl}b")[::-1], 2) for j in range(m)])
    for q in range(spc):
        assert np.array_equal(nat[np.arange(m) * spc + q], rev[br * spc + q])
        assert np.array_equal(bins_a[np.arange(m) + q * m], nat[np.arange(m) * spc + q])
    assert np.array_equal(bins_a, bins_b)


@pytest.mark.parametrize("n,spc", [(64, 2), (1024, 2), (1024, 1), (4096, 2), (4096, 1)])
def test_fft_impulse_and_dc_closed_forms(orc, n, spc):
    """x = 32767*delta -> every bin 32767/N (scaled 1/N by the all-ones schedule),
    rounded; DC -> only bin 0."""
    x = np.zeros(n, np.uint32)
    x[0] = vecio.pack(32767, 0)
    _, bins, sf = orc.fft(x, n, spc)
    vals = {vecio.unpack(int(b)) for b in bins}
    assert len(vals) == 1 and sf[0] == 0
    (re, im), = vals
    assert im == 0 and abs(re - 32767 / n) <= 1.0
    d = np.full(n, vecio.pack(16384, 0), np.uint32)
    _, bins, _ = orc.fft(d, n, spc)
    b = np.array([vecio.unpack(int(v)) for v in bins])
    assert abs(b[0, 0] - 16384 * 0.999) < 64 and np.all(np.abs(b[1:]) <= 2)


@pytest.mark.parametrize("n,spc", [(64, 2), (1024, 2), (4096, 2), (4096, 1)])
def test_fft_float_dft_bound(orc, n, spc):
    """|fixed - DFT/N| within 16 LSB on non-saturating random records (the reference's
    NumPy cross-check gate, DECISIONS.md:1886-1892) — the only pin for N = 4096."""
    rng = np.random.default_rng(11)
    for _ in range(4):
        x = stim.rand_cx(rng, n, scale=12000)
        _, bins, sf = orc.fft(x, n, spc)
        assert sf[0] == 0
        xc = stim.to_complex(x)
        ref = np.fft.fft(xc) / n
        got = stim.to_complex(bins)
        assert np.max(np.abs(got - ref)) <= 16.0


def test_bf_unit_zero_and_permutation(orc):
    rng = np.random.default_rng(5)
    CH, B, cols = 8, 8, 64
    Xa = stim.rand_cx(rng, cols * CH)
    sel = [3, 1, 4, 1, 5, 2, 6, 7]
    Wu = np.zeros(B * CH, np.uint32)
    for b in range(B):
        Wu[b * CH + sel[b]] = vecio.pack(32767, 0)
    y, fl, pw, acc = orc.bf(Xa, Wu, CH, B)
    for b in range(B):
        for c in range(cols):
            xr, xi = vecio.unpack(int(Xa[c * CH + sel[b]]))
            exp = vecio.pack(stim.rne_int(xr * 32767, 15), stim.rne_int(xi * 32767, 15))
            assert y[b * cols + c] == exp
    y0, fl0, pw0, _ = orc.bf(Xa, np.zeros(B * CH, np.uint32), CH, B)
    assert not y0.any() and not fl0.any() and not pw0.any()
    W = stim.rand_cx(rng, B * CH)
    perm = rng.permutation(CH)
    Wp = np.zeros_like(W)
    Xp = np.zeros_like(Xa)
    for b in range(B):
        Wp[b * CH + perm] = W[b * CH: (b + 1) * CH]
    for c in range(cols):
        Xp[c * CH + perm] = Xa[c * CH: (c + 1) * CH]
    ya, fa, _, aa = orc.bf(Xa, W, CH, B)
    yb, fb, _, ab = orc.bf(Xp, Wp, CH, B)
    assert np.array_equal(ya, yb) and np.array_equal(aa, ab) and np.array_equal(fa, fb)


def test_bf_float_cross_check(orc):
    """bf::dot_float analogue: the exact accumulator equals the double-precision dot."""
    rng = np.random.default_rng(6)
    CH, B, cols = 16, 4, 32
    Xa = stim.rand_cx(rng, cols * CH)
    W = stim.rand_cx(rng, B * CH)
    _, _, _, acc = orc.bf(Xa, W, CH, B)
    x = stim.to_complex(Xa).reshape(cols, CH)
    w = stim.to_complex(W).reshape(B, CH)
    ref = w @ x.T  # [B][cols]
    acc = acc.reshape(B, cols, 2)
    assert np.array_equal(acc[..., 0], ref.real.astype(np.int64))
    assert np.array_equal(acc[..., 1], ref.imag.astype(np.int64))


def test_power_and_cross_power_identities(orc):
    m = vecio.pack(-32768, -32768)
    assert orc.power(m) == 2**31
    rng = np.random.default_rng(8)
    for v in stim.rand_cx(rng, 300).tolist() + [m, vecio.pack(32767, -32768)]:
        sre, sim, dre, dim = orc.cross_power(v, v)
        assert (sre, sim) == (dre, dim) == (orc.power(v), 0)
    a, b = stim.rand_cx(rng, 200), stim.rand_cx(rng, 200)
    for x, y in zip(a.tolist(), b.tolist()):
        sre, sim, dre, dim = orc.cross_power(x, y)
        assert (sre, sim) == (dre, dim)
        xc, yc = stim.to_complex(np.array([x], np.uint32))[0], stim.to_complex(np.array([y], np.uint32))[0]
        z = xc * np.conj(yc)
        assert (sre, sim) == (int(z.real), int(z.imag))


def test_cov_brute_force_and_integrator_width(orc):
    rng = np.random.default_rng(9)
    CH, F, N = 4, 12, 8
    Xa = stim.rand_cx(rng, F * N * CH)
    R, sf = orc.cov(Xa, CH, F, N, acc_w=64)
    x = stim.to_complex(Xa).reshape(F, N, CH)
    R = R.reshape(N, -1, 2)
    for k in range(N):
        G = x[:, k, :].T @ np.conj(x[:, k, :])  # G[i][j] = sum_m x_i conj(x_j)
        p = 0
        for i in range(CH):
            for j in range(i, CH):
                assert R[k, p, 0] == int(G[i, j].real) and R[k, p, 1] == int(G[i, j].imag)
                p += 1
    assert not sf.any()
    # 255 extreme terms are exact at acc_w = 40; the 256th saturates (DECISIONS.md:2030-2041)
    m = np.full(256, vecio.pack(-32768, -32768), np.uint32)
    R40, sf40 = orc.cov(m.reshape(256, 1, 1), 1, 256, 1, acc_w=40)
    assert R40[0] == 2**39 - 1 and (sf40[0] & 3) == 2
    R40b, sf40b = orc.cov(m[:255].reshape(255, 1, 1), 1, 255, 1, acc_w=40)
    assert R40b[0] == 255 * 2**31 and sf40b[0] == 0


CA = [1, 0, 0, 2, 2, 8, 8, 0x100]


def test_cfar_flat_spectrum_alpha_edges(orc):
    n = 64
    flat = np.full(n, 1000, np.uint64)
    cfg = list(CA)
    ev, rc = orc.cfar(flat, n, cfg, 31, 63)
    assert rc[0] == 1  # summary only
    cfg[7] = 0xFF
    ev, rc = orc.cfar(flat, n, cfg, 31, 63)
    evaluable = n - 2 * (2 + 8)
    assert rc[0] == evaluable + 1
    zero = np.zeros(n, np.uint64)
    for a in (0, 1, 0xFFFF):
        cfg[7] = a
        _, rc = orc.cfar(zero, n, cfg, 31, 63)
        assert rc[0] == 1


def test_cfar_planted_side_swap_is_caught(orc, ref):
    rng = np.random.default_rng(12)
    n = 128
    bins = rng.integers(0, 2**20, n).astype(np.uint64)
    bins[[10, 40, 90]] = 2**28
    cfg = [1, 1, 0, 2, 3, 8, 4, 0x300]  # GO, asymmetric geometry
    e_ref, _ = ref.cfar(bins, n, cfg, 31, 63)
    orc.lib.rdo_set_fault(5)
    try:
        e_bad, _ = orc.cfar(bins, n, cfg, 31, 63)
    finally:
        orc.lib.rdo_set_fault(0)
    e_ok, _ = orc.cfar(bins, n, cfg, 31, 63)
    assert np.array_equal(e_ok, e_ref)
    assert e_bad.shape != e_ref.shape or not np.array_equal(e_bad, e_ref)


def test_event_pack_round_trip(orc):
    rng = np.random.default_rng(13)
    widths = [2, 16, 16, 7, 16, 16, 16, 40, 47]
    lsbs = [0, 2, 18, 34, 41, 57, 73, 89, 129]
    assert lsbs[-1] + widths[-1] == 176
    for _ in range(200):
        f = [int(rng.integers(0, 2**w)) for w in widths] + [int(rng.integers(0, 2)), int(rng.integers(0, 2))]
        w = orc.event_pack(np.array(f, np.uint64))
        big = int(w[0]) | (int(w[1]) << 64) | (int(w[2]) << 128)
        for v, lsb, wd in zip(f, lsbs, widths):
            assert (big >> lsb) & ((1 << wd) - 1) == v
        assert (big >> 176) & 1 == f[9] and (big >> 177) & 1 == f[10]
        assert big >> 178 == 0


def test_event_detect_recheck(orc):
    """Re-running C*N*2^8 > A*S on unpacked DETECT fields reproduces the decision."""
    rng = np.random.default_rng(14)
    n = 256
    bins = rng.integers(0, 2**24, n).astype(np.uint64)
    bins[rng.integers(0, n, 12)] = 2**30
    ev, rc = orc.cfar(bins, n, [1, 0, 0, 2, 2, 16, 16, 0x400], 31, 63)
    for w in ev:
        big = int(w[0]) | (int(w[1]) << 64) | (int(w[2]) << 128)
        kind = big & 3
        if kind != 0:
            continue
        refc = (big >> 34) & 0x7F
        alpha = (big >> 41) & 0xFFFF
        cut = (big >> 89) & ((1 << 40) - 1)
        s = (big >> 129) & ((1 << 47) - 1)
        assert (cut * refc) << 8 > alpha * s


def _flit_bits(words, nbits):
    return [(int(words[i // 32]) >> (i % 32)) & 1 for i in range(nbits)]


@pytest.mark.parametrize("packet_w", [36, 64, 128, 256, 512])
def test_event_flits_decode(orc, packet_w):
    """Field-level check of the packet layout from the prose definition
    (ARCHITECTURE.md:1254-1330): odd parity, sof/eof framing, header fields,
    length = total flits, the payload reassembles the low 176 event bits."""
    rng = np.random.default_rng(7)
    ev = rng.integers(0, 2**63, size=(5, 3), dtype=np.uint64)
    fl = orc.event_flits(ev, packet_w, dest=5, src=20, vc=2, seq0=65534)
    L = 1 + -(-176 // packet_w)
    assert fl.shape[1] == L
    for e in range(5):
        payload = []
        for f in range(L):
            bits = _flit_bits(fl[e, f], packet_w + 5)
            assert sum(bits) % 2 == 1
            assert bits[1] + 2 * bits[2] == 2
            assert bits[3] == (f == 0) and bits[4] == (f == L - 1)
            data = bits[5:]
            if f == 0:
                h = sum(b << i for i, b in enumerate(data[:36]))
                assert not any(data[36:])
                assert (h & 15, (h >> 4) & 31, (h >> 9) & 3, (h >> 11) & 7, (h >> 14) & 63, h >> 20) == \
                    (5, 20, 2, 0, L, (65534 + e) & 0xFFFF)
            else:
                payload += data
        want = [(int(ev[e, i // 64]) >> (i % 64)) & 1 for i in range(176)]
        assert payload[:176] == want and not any(payload[176:])


@pytest.mark.parametrize("packet_w", [36, 64, 100, 176, 512])
def test_event_writeback_round_trip(orc, packet_w):
    """Flits -> writeback is the identity on the low 176 event bits: every
    packet lands zero-extended in its 64-byte word, the counters show only the
    packets written; a ring shorter than the stream wraps (later packets
    overwrite earlier ones)."""
    rng = np.random.default_rng(packet_w + 5)
    ev = rng.integers(0, 2**63, size=(20, 3), dtype=np.uint64)
    ev[:, 2] |= np.uint64(3) << np.uint64(48)  # the stream's sof/eof bits are not payload
    fl = orc.event_flits(ev, packet_w, dest=4, src=9, vc=3, seq0=65530)
    ring, st = orc.event_writeback(fl, packet_w, port=4, seq0=65530)
    assert np.array_equal(ring, flitfault.ring_of_events(ev))
    assert st.tolist() == [20, 0, 0, 0, 0]
    ring, st = orc.event_writeback(fl, packet_w, port=4, seq0=65530, ring_words=7, addr0=5)
    assert np.array_equal(ring, flitfault.ring_of_events(ev, 7, 5))


# (parity, framing, misrouted, sequence) each planted fault must add: one
# packet-local violation apiece, derived from the checks' definitions.
_FAULT_COUNTS = {"data": (1, 0, 0, 0), "parity": (1, 0, 0, 0), "type": (0, 1, 0, 0), "hvc": (0, 1, 0, 0),
                 "dest": (0, 0, 1, 0), "seq": (0, 0, 0, 1), "len": (0, 1, 0, 0), "sof_body": (0, 1, 0, 0),
                 "eof_clear": (0, 1, 0, 0), "flit_vc": (0, 1, 0, 0), "hdr_eof": (0, 1, 0, 0), "hdr_sof": (0, 1, 0, 0)}


@pytest.mark.parametrize("kind", flitfault.ALL)
def test_event_writeback_counts_planted_faults(orc, kind):
    """Each planted fault adds exactly its one count, and only a payload fault
    changes what lands in the ring (the one flipped bit)."""
    pw = 64
    rng = np.random.default_rng(3)
    ev = rng.integers(0, 2**63, size=(6, 3), dtype=np.uint64)
    fl = flitfault.plant(orc.event_flits(ev, pw, dest=2, src=1, vc=1, seq0=10), 4, kind, pw)
    ring, st = orc.event_writeback(fl, pw, port=2, seq0=10)
    assert tuple(int(x) for x in st[1:]) == _FAULT_COUNTS[kind] and st[0] == 6
    want = flitfault.ring_of_events(ev)
    if kind == "data":
        want[4, 0] ^= np.uint32(1 << 3)
    assert np.array_equal(ring, want)


@pytest.mark.parametrize("H,spc,shift", [(16, 2, False), (64, 1, True), (256, 2, True)])
def test_doppler_tone_lands_in_its_bin(orc, H, spc, shift):
    """Closed form for the slow-time transform: a per-bin phase ramp across
    frames exp(2 pi i k_d f / H) of amplitude A transforms (every stage scaled,
    so DFT / H) to A in Doppler bin k_d and ~0 elsewhere, within the rounding
    bound of the fixed-point stages; with the shift, bin k_d sits at row
    (k_d + H/2) mod H (range_doppler.py:254-275). The ring's window starts
    after `newest`, so the frames are taken oldest first across the wrap."""
    rows, n_bins, depth = 2, 5, H + 3
    newest = 4  # window wraps around the end of the ring
    kd = np.array([[3, H // 2 + 1, 0, H - 1, 7 % H], [1, 2, H // 4, 5, H - 3]])
    A = 12000.0
    ring = np.zeros((rows, depth, n_bins), np.uint32)
    for f in range(H):
        slot = (newest - (H - 1) + f) % depth
        z = A * np.exp(2j * np.pi * kd * f / H)
        ring[:, slot, :] = rd_pack(np.rint(z.real), np.rint(z.imag))
    out, _ = orc.doppler(ring, rows, depth, n_bins, newest, H, spc=spc, shift=shift)
    spec = unpack_cx(out).reshape(rows, H, n_bins)
    bound = np.log2(H) + 2  # LSB: one rounding per scaled stage plus the input quantisation
    for r in range(rows):
        for k in range(n_bins):
            row = (kd[r, k] + H // 2) % H if shift else kd[r, k]
            want = np.zeros(H, complex)
            want[row] = A
            assert np.abs(spec[r, :, k] - want).max() <= bound, (r, k)
    # the same against the float range-Doppler picture, scaled by 1/H
    flt = np.fft.fft(unpack_cx(ring[:, (newest - (H - 1) + np.arange(H)) % depth, :].reshape(-1)).reshape(
        rows, H, n_bins), axis=1) / H
    if shift:
        flt = np.fft.fftshift(flt, axes=1)
    assert np.abs(spec - flt).max() <= bound


def rd_pack(re, im):
    return ((np.asarray(im, np.int64) & 0xFFFF) << 16 | (np.asarray(re, np.int64) & 0xFFFF)).astype(np.uint32)


def unpack_cx(w):
    w = np.asarray(w, np.uint32).astype(np.int64)
    re = (w & 0xFFFF) - ((w & 0x8000) << 1)
    im = (w >> 16) - (((w >> 16) & 0x8000) << 1)
    return re + 1j * im


def test_cov_exp_mode_closed_forms(orc):
    """Exponential mode: k = 0 passes the last cross power through; for a
    constant input the accumulator settles in (x - 2^k, x], approached from
    below with the truncating dead band (covar_model.hpp:206-214)."""
    rng = np.random.default_rng(3)
    CH, N = 3, 4
    F = 9
    Xa = stim.rand_cx(rng, F * N * CH)
    R, _ = orc.cov_integrate(Xa, CH, F, N, 1, 0, 64)
    last, _ = orc.cov(Xa[(F - 1) * N * CH:], CH, 1, N)
    assert np.array_equal(R, last)
    for k in (1, 3, 4):
        F = 2 ** k * 24
        x = np.tile(stim.rand_cx(rng, N * CH), F)  # the same frame every time
        R, _ = orc.cov_integrate(x, CH, F, N, 1, k, 64)
        one, _ = orc.cov(x[: N * CH], CH, 1, N)
        d = one - R
        assert np.all((d >= 0) & (d < 2 ** k)), k
