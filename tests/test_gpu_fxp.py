"""Device gate of the fixed-point primitives: every record of the reference's
committed vectors (fxp_ops.vec, fxp_accum.vec, cmult.vec, the gates of
model/cpp/test/test_fxp_vectors.cpp) evaluated on the GPU by the device
functions the stage kernels inline (rd_fxp.h), through rd_fxp_eval /
rd_fxp_accum / rd_fxp_cmult. The specialised forms the kernels use (16-bit
saturation, shift-15 rounding in 32 and 64 bits, the one-bit butterfly
rounding) are checked against the general form on every record whose
operands they cover; a random sweep over widths and shifts adds operands the
files do not hold, checked against the oracle."""
import numpy as np
import pytest

import rd_amd
import vecio

pytestmark = pytest.mark.gpu


def _records(rows):
    ops = np.zeros(len(rows), rd_amd.FXP_OP_DTYPE)
    for i, r in enumerate(rows):
        ops[i] = (r["a"], r["b"], rd_amd.FXP_OPS[r["op"]], r["sh"], r["w"], 0)
    return ops


def test_fxp_ops_vectors_on_device(gpu):
    rows = vecio.load_ops()
    y, fl, forms = gpu.fxp_eval(_records(rows))
    bad = [r["id"] for i, r in enumerate(rows) if y[i] != r["y"] or fl[i] != r["flags"]]
    assert not bad, f"{len(bad)} of {len(rows)} records differ: {bad[:10]}"
    assert not np.any(forms & 2), [rows[i]["id"] for i in np.flatnonzero(forms & 2)[:10]]
    assert int(np.count_nonzero(forms & 1)) >= 300  # the kernels' forms were exercised


def test_fxp_accum_vectors_on_device(gpu):
    for sid, steps in vecio.load_accum().items():
        acc_w = steps[0]["acc_w"]
        y, sf, st, cnt = gpu.fxp_accum(acc_w, [s["x"] for s in steps])
        for i, s in enumerate(steps):
            assert (y[i], sf[i], st[i], cnt[i]) == (s["y"], s["step_flags"], s["sticky"], s["count"]), (sid, i)


def test_cmult_vectors_on_device(gpu):
    recs = vecio.load_cmult()
    a = np.array([r["a"] for r in recs], np.uint32)
    b = np.array([r["b"] for r in recs], np.uint32)
    p4, p3, y, fl = gpu.fxp_cmult(a, b)
    assert np.array_equal(p4[:, 0], [r["p_re"] for r in recs]) and np.array_equal(p4[:, 1], [r["p_im"] for r in recs])
    assert np.array_equal(p3, p4)  # MULT3 == MULT4 (cmult.hpp:139-168)
    assert np.array_equal(y, [r["y"] for r in recs])
    assert np.array_equal(fl & 3, [r["f_re"] for r in recs]) and np.array_equal(fl >> 2, [r["f_im"] for r in recs])


def test_random_sweep_against_oracle(gpu, orc):
    rng = np.random.default_rng(2026)
    names = ["sat", "trunc", "round_even", "round_half_up", "round_sat", "add_sat", "sub_sat", "neg_sat", "mul_q15",
             "mul_q15_rs", "cmul_re", "cmul_im", "growth_bits", "acc_w"]
    rows = []
    for _ in range(6000):
        op = names[int(rng.integers(len(names)))]
        mag = int(rng.choice([16, 31, 32, 40, 47, 62, 63]))
        a = int(rng.integers(-(1 << mag), 1 << mag))
        b = int(rng.integers(-(1 << mag), 1 << mag))
        if op in ("cmul_re", "cmul_im", "mul_q15", "mul_q15_rs", "growth_bits", "acc_w"):
            a, b = int(rng.integers(0, 1 << 32)), int(rng.integers(0, 1 << 32))
        rows.append(dict(op=op, a=a, b=b, sh=int(rng.integers(0, 66)), w=int(rng.integers(1, 66))))
    y, fl, forms = gpu.fxp_eval(_records(rows))
    for i, r in enumerate(rows):
        oy, ofl = orc.fxp_op(r["op"], r["a"], r["b"], r["sh"], r["w"])
        assert (int(y[i]), int(fl[i])) == (oy, ofl), r
    assert not np.any(forms & 2)
