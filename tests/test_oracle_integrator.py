"""The oracle's restatements of the reference's stateful / per-element model
calls, pinned against the reference headers (oracle/_ref) and closed forms:

* covar::Integrator (covar_model.hpp:186-262) over random control streams —
  windows of every length (0 counts as 1), both modes, shifts past 63,
  enable toggles, flush and sat_clear at random steps, configuration changes
  that latch only at window boundaries, every accumulator width from 1 to 64
  and data that saturates and wraps;
* the same lanes fed by a covariance's cross powers;
* pfb::LaneModel::accumulate (pfb_model.hpp:135-150) with and without history;
* cfar::decide (cfar_model.hpp:243-285) with the configuration unclamped,
  including windows that leave the row (partial sums), GO ties and huge
  reference counts.
"""
import numpy as np
import pytest

import stim

I64_MIN, I64_MAX = -(1 << 63), (1 << 63) - 1


def random_steps(rng, n, p_flush=0.05, p_clear=0.05, p_cfg=0.1, max_len=6):
    cfg = [int(rng.integers(0, max_len + 1)), int(rng.integers(0, 2)), int(rng.integers(0, 70)), 1]
    out = []
    for _ in range(n):
        if rng.random() < p_cfg:
            cfg = [int(rng.integers(0, max_len + 1)), int(rng.integers(0, 2)), int(rng.integers(0, 70)),
                   int(rng.random() < 0.9)]
        out.append(cfg + [int(rng.random() < 0.8), int(rng.random() < p_flush), int(rng.random() < p_clear)])
    return np.array(out, np.uint32)


def random_data(rng, n, lanes, kind):
    if kind == "small":
        return rng.integers(-1000, 1000, (n, lanes))
    if kind == "q31":
        return rng.integers(-(1 << 31), (1 << 31) + 1, (n, lanes))
    return rng.integers(I64_MIN, I64_MAX, (n, lanes), dtype=np.int64, endpoint=True)


def _same(a, b):
    for k in ("acc", "flags", "count", "results", "windows", "ctl", "n_windows"):
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k


@pytest.mark.parametrize("acc_w", [1, 2, 8, 20, 33, 40, 63, 64])
@pytest.mark.parametrize("kind", ["small", "q31", "wide"])
def test_integ_run_vs_reference(orc, ref, acc_w, kind):
    rng = np.random.default_rng(acc_w * 17 + len(kind))
    for trial in range(4):
        n = int(rng.integers(1, 120))
        steps = random_steps(rng, n)
        cfg0 = steps[0, :4] if trial % 2 else np.array([3, 0, 0, 1], np.uint32)
        data = random_data(rng, n, 5, kind)
        _same(orc.integ_run(acc_w, cfg0, steps, data, cap=200), ref.integ_run(acc_w, cfg0, steps, data, cap=200))


def test_integ_run_capacity_and_window_fields(orc, ref):
    rng = np.random.default_rng(9)
    steps = random_steps(rng, 300, p_cfg=0.0, max_len=2)
    data = random_data(rng, 300, 2, "small")
    a, b = orc.integ_run(40, [2, 0, 0, 1], steps, data, cap=7), ref.integ_run(40, [2, 0, 0, 1], steps, data, cap=7)
    _same(a, b)
    assert a["n_windows"] > 7 and a["results"].shape[0] == 7


def test_integ_block_windows_are_plain_sums(orc):
    """Block mode at acc_w = 64, windows of L accepted samples, no flush: each
    window result is the sum of its L samples, window ids count up."""
    rng = np.random.default_rng(3)
    L, n = 5, 53
    steps = np.array([[L, 0, 0, 1, 1, 0, 0]] * n, np.uint32)
    data = rng.integers(-(1 << 31), 1 << 31, (n, 3))
    o = orc.integ_run(64, [L, 0, 0, 1], steps, data, cap=64)
    assert o["n_windows"] == n // L
    assert np.array_equal(o["results"], data[: (n // L) * L].reshape(n // L, L, 3).sum(1))
    assert np.array_equal(o["windows"][:, 1], np.arange(n // L))
    assert np.all(o["windows"][:, 2] == L) and not o["windows"][:, 3:].any()


def test_integ_exp_k0_passes_the_sample_through(orc):
    """Exponential mode with k = 0: acc + (x - acc) = x, so every window of one
    sample emits that sample (saturated to acc_w)."""
    rng = np.random.default_rng(4)
    n = 40
    steps = np.array([[1, 1, 0, 1, 1, 0, 0]] * n, np.uint32)
    data = rng.integers(-(1 << 40), 1 << 40, (n, 2))
    o = orc.integ_run(36, [1, 1, 0, 1], steps, data, cap=64)
    assert np.array_equal(o["results"], np.clip(data, -(1 << 35), (1 << 35) - 1))


@pytest.mark.parametrize("CH,F,N", [(4, 37, 5), (7, 20, 3), (16, 12, 2)])
def test_cov_integrate_windows_vs_reference(orc, ref, CH, F, N):
    rng = np.random.default_rng(CH + F + N)
    for Xa, acc_w in ((stim.rand_cx(rng, F * N * CH), 40), (stim.rand_extreme(rng, F * N * CH), 33)):
        steps = random_steps(rng, F)
        a = orc.cov_integrate_windows(Xa, CH, F, N, acc_w, steps[0, :4], steps, cap=64, threads=4)
        b = ref.cov_integrate_windows(Xa, CH, F, N, acc_w, steps[0, :4], steps, cap=64, threads=4)
        _same(a, b)


def test_cov_integrate_windows_one_window_is_the_covariance(orc):
    rng = np.random.default_rng(8)
    CH, F, N = 5, 30, 4
    Xa = stim.rand_cx(rng, F * N * CH)
    steps = np.array([[F, 0, 0, 1, 1, 0, 0]] * F, np.uint32)
    o = orc.cov_integrate_windows(Xa, CH, F, N, 64, [F, 0, 0, 1], steps, cap=4)
    assert o["n_windows"] == 1
    assert np.array_equal(o["results"][0], orc.cov(Xa, CH, F, N)[0])


@pytest.mark.parametrize("P,T,nb,ch", [(4, 8, 11, 2), (16, 16, 5, 1), (8, 1, 4, 2), (32, 5, 3, 3)])
def test_pfb_acc_vs_reference(orc, ref, P, T, nb, ch):
    rng = np.random.default_rng(P + T + nb)
    for x, h in ((stim.rand_cx(rng, ch * nb * P), stim.rand_cx(rng, P * T)),
                 (stim.rand_extreme(rng, ch * nb * P), stim.rand_extreme(rng, P * T))):
        hist = stim.rand_cx(rng, ch * max(T - 1, 0) * P) if T > 1 else None
        for hh in (None, hist):
            assert np.array_equal(orc.pfb_acc(x, h, P, T, ch, nb, hh), ref.pfb_acc(x, h, P, T, ch, nb, hh))


def test_pfb_acc_rounds_to_the_pfb_output(orc):
    """round_sat(acc, 15, 16) of the exact accumulators is the PFB output."""
    rng = np.random.default_rng(11)
    P, T, nb, ch = 8, 16, 9, 2
    x, h = stim.rand_extreme(rng, ch * nb * P), stim.rand_extreme(rng, P * T)
    acc = orc.pfb_acc(x, h, P, T, ch, nb).reshape(-1, 2)
    y, _ = orc.pfb(x, h, P, T, ch, nb)
    exp = stim.pack(np.array([stim.rne_int(int(v), 15) for v in acc[:, 0]]),
                    np.array([stim.rne_int(int(v), 15) for v in acc[:, 1]]))
    assert np.array_equal(exp, y)


@pytest.mark.parametrize("cfg", [[1, 0, 0, 2, 3, 8, 5, 0x300], [1, 1, 0, 1, 2, 4, 6, 0x200], [1, 1, 0, 0, 0, 3, 3, 0x100],
                                 [0, 0, 0, 2, 2, 8, 8, 0x800], [1, 1, 0, 2, 2, 0, 8, 0x800],
                                 [1, 0, 0, 4, 1, 1000, 2, 0x400], [1, 0, 0, 0xFFFFFFFF, 2, 3, 3, 0x100],
                                 [1, 0, 0, 0, 0, 0, 0, 0x100]])
def test_cfar_decide_vs_reference(orc, ref, cfg):
    rng = np.random.default_rng(sum(cfg) % 1000)
    n, rows = 97, 3
    for bins in (rng.integers(0, 1 << 20, n * rows).astype(np.uint64),
                 rng.integers(0, 1 << 63, n * rows, dtype=np.uint64),
                 np.full(n * rows, 5, np.uint64)):
        assert np.array_equal(orc.cfar_decide(bins, n, cfg), ref.cfar_decide(bins, n, cfg))


def test_cfar_decide_matches_run_frame_detections(orc):
    """The detections of decide() over a row are the DETECT events of
    run_frame with the same (in-range) configuration."""
    rng = np.random.default_rng(12)
    n = 300
    bins = rng.integers(0, 1 << 16, n).astype(np.uint64)
    bins[[50, 120, 121, 200]] = 1 << 22
    cfg = [1, 0, 0, 2, 2, 8, 8, 0x400]
    det = orc.cfar_decide(bins, n, cfg)
    ev, _ = orc.cfar(bins, n, cfg, 31, 63)
    kinds = ev[:, 0] & 3
    bins_ev = (ev[:, 0] >> 2) & 0xFFFF
    assert sorted(bins_ev[kinds == 0].tolist()) == np.flatnonzero(det[:, 6] & 2).tolist()
