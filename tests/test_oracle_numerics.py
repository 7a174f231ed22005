"""Oracle numerics pinned against the reference's committed vectors.

Reference tests mirrored: model/cpp/test/test_fxp_vectors.cpp (ops, accum,
cmult, RNE-bias and saturation properties) and the twiddle-table rule of
model/python/gen_fft_twiddles.py (digest, axis points).
"""
import hashlib
from decimal import Decimal, getcontext

import numpy as np
import pytest

import vecio


def test_fxp_ops_vectors(orc):
    recs = vecio.load_ops()
    assert len(recs) == 1457
    bad = []
    for r in recs:
        y, fl = orc.fxp_op(r["op"], r["a"], r["b"], r["sh"], r["w"])
        if (y, fl) != (r["y"], r["flags"]):
            bad.append((r["id"], y, fl, r["y"], r["flags"]))
    assert not bad, bad[:5]


def test_fxp_accum_vectors(orc):
    seqs = vecio.load_accum()
    assert sum(len(s) for s in seqs.values()) == 251
    for name, steps in seqs.items():
        acc_w = steps[0]["acc_w"]
        y, sf, st, cnt = orc.acc_run(acc_w, [s["x"] for s in steps])
        for i, s in enumerate(steps):
            assert s["step"] == i
            assert (y[i], sf[i], st[i], cnt[i]) == (s["y"], s["step_flags"], s["sticky"], s["count"]), (name, i)


def test_cmult_vectors_and_mult3(orc):
    recs = vecio.load_cmult()
    assert len(recs) == 864
    for r in recs:
        o = orc.cmult(r["a"], r["b"])
        assert (o["p_re"], o["p_im"]) == (r["p_re"], r["p_im"]), r["id"]
        assert (o["p3_re"], o["p3_im"]) == (r["p_re"], r["p_im"]), r["id"]  # MULT3 == MULT4
        assert (o["y"], o["f_re"], o["f_im"]) == (r["y"], r["f_re"], r["f_im"]), r["id"]


def test_cmult_corner_identity(orc):
    """(-1-j)^2 has imaginary part +2^31 (exact, 33-bit) and saturates to 0x7FFF."""
    m = vecio.pack(-32768, -32768)
    o = orc.cmult(m, m)
    assert o["p_im"] == 2**31 and o["p_re"] == 0
    assert vecio.unpack(o["y"]) == (0, 32767) and o["f_im"] == 2


def test_rne_bias_zero(orc):
    """RNE has zero bias over a full residue sweep; half-up does not (NUMERICS.md:197-202)."""
    s = 3
    tot_rne = tot_hu = 0
    for v in range(-65536, 65536):
        y, _ = orc.fxp_op("round_even", v, 0, s, 64)
        h, _ = orc.fxp_op("round_half_up", v, 0, s, 64)
        tot_rne += y * (1 << s) - v
        tot_hu += h * (1 << s) - v
    assert tot_rne == 0
    assert tot_hu == 65536


def test_saturation_idempotent(orc):
    rng = np.random.default_rng(7)
    for w in range(2, 49):
        for v in rng.integers(-(2**50), 2**50, 20):
            a, _ = orc.fxp_op("sat", int(v), 0, 0, w)
            b, fl = orc.fxp_op("sat", a, 0, 0, w)
            assert a == b and fl == 0


def _digest(words):
    h = hashlib.sha256()
    for w in words:
        h.update(int(w).to_bytes(4, "little"))
    return h.hexdigest()[:16]


def test_twiddle_table_decimates_to_reference_digest(orc):
    t = orc.twiddle_table(4096)
    assert _digest(t[::4]) == vecio.TWIDDLE_DIGEST_1024
    assert _digest(orc.twiddle_table(1024)) == vecio.TWIDDLE_DIGEST_1024


def test_twiddle_table_matches_reference_header(ref, orc):
    assert np.array_equal(ref.twiddle_table(1024), orc.twiddle_table(4096)[::4])


def test_twiddle_axis_points(orc):
    t = orc.twiddle_table(4096)
    assert vecio.unpack(t[0]) == (32767, 0)
    assert vecio.unpack(t[1024]) == (0, -32768)
    assert vecio.unpack(t[2048]) == (-32768, 0)
    assert vecio.unpack(t[3072]) == (0, 32767)


def _cos_sin_hp(num, den):
    """cos/sin(2*pi*num/den) to ~40 digits by Taylor series in Decimal."""
    getcontext().prec = 50
    pi = Decimal("3.14159265358979323846264338327950288419716939937510")
    x = 2 * pi * num / den
    c = s = Decimal(0)
    term = Decimal(1)
    for k in range(0, 60):
        if k % 2 == 0:
            c += term * (1 if (k // 2) % 2 == 0 else -1)
        else:
            s += term * (1 if (k // 2) % 2 == 0 else -1)
        term = term * x / (k + 1)
    return c, s


def test_twiddle_4096_has_no_near_ties_and_matches_high_precision(orc):
    """Every first-quadrant entry of the 4096 table is >1e-9 LSB away from a .5 tie,
    so any faithful libm quantises it to the same integer; and the integers equal
    the high-precision rounding. This pins the 3072 entries the 1024 table leaves open."""
    t = orc.twiddle_table(4096)
    scale = Decimal(32768)
    for e in range(1, 1024):
        c, s = _cos_sin_hp(e, 4096)
        for v in (c * scale, -s * scale):
            frac = abs(v - v.to_integral_value(rounding="ROUND_FLOOR") - Decimal("0.5"))
            assert frac > Decimal("1e-9"), e
        re = min(32767, int(c * scale + Decimal("0.5")) if c >= 0 else -int(-c * scale + Decimal("0.5")))
        im_v = -s * scale
        im = -int(-im_v + Decimal("0.5")) if im_v < 0 else int(im_v + Decimal("0.5"))
        assert vecio.unpack(t[e]) == (re, im), e


@pytest.mark.parametrize("fault", [1])
def test_planted_rounding_fault_is_caught(orc, fault):
    """With RNE replaced by half-up in the oracle, the committed vectors must fail."""
    orc.lib.rdo_set_fault(fault)
    try:
        bad = 0
        for r in vecio.load_ops():
            if r["op"] in ("round_even", "round", "round_sat", "mul_q15_rs", "cmul_re", "cmul_im"):
                y, fl = orc.fxp_op(r["op"], r["a"], r["b"], r["sh"], r["w"])
                bad += (y, fl) != (r["y"], r["flags"])
        assert bad > 0
    finally:
        orc.lib.rdo_set_fault(0)
