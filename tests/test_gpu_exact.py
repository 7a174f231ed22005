"""GPU parity of the exact per-element entry points behind the C++ mirror's
scalar APIs, against the oracle (itself pinned to the reference headers in
tests/test_oracle_integrator.py):

* rd_integ_run — covar::Integrator lanes over random control streams
  (covar_model.hpp:186-262), one call and the same stream split over two
  calls (the control and lane state carried between them);
* rd_cov_integrate_windows — the covariance's cross powers as the lanes;
* rd_pfb_acc_run — pfb::LaneModel::accumulate (pfb_model.hpp:135-150);
* rd_cfar_decide — cfar::decide with the configuration unclamped
  (cfar_model.hpp:243-285).
"""
import numpy as np
import pytest

import rd_amd
import stim
from test_oracle_integrator import random_data, random_steps

pytestmark = pytest.mark.gpu


def _lanes(gpu, n):
    return gpu.put(np.zeros(n, rd_amd.INTEG_LANE_DTYPE))


def _ctl(cfg4):
    c = rd_amd.IntegCtl()
    assert rd_amd.load().lib.rd_integ_reset(rd_amd.C.byref(rd_amd.IntegCfg(*[int(v) for v in cfg4])),
                                            rd_amd.C.byref(c)) == 0
    return c


def _check_lanes(lanes, o, n):
    got = lanes.download(np.uint8, 16 * n).view(rd_amd.INTEG_LANE_DTYPE)
    assert np.array_equal(got["acc"], o["acc"])
    assert np.array_equal(got["flags"], o["flags"])
    assert np.array_equal(got["sat_count"], o["count"])


def _ctl_list(c):
    return [c.active.window_len, c.active.mode, c.active.exp_k, c.active.enable, c.count, c.window_id]


@pytest.mark.parametrize("acc_w", [1, 2, 20, 33, 40, 63, 64])
@pytest.mark.parametrize("kind", ["small", "q31", "wide"])
def test_integ_run_matches_oracle(gpu, orc, acc_w, kind):
    rng = np.random.default_rng(acc_w * 5 + len(kind))
    n, L = 150, 300
    steps = random_steps(rng, n)
    data = random_data(rng, n, L, kind)
    cfg0 = steps[0, :4]
    o = orc.integ_run(acc_w, cfg0, steps, data, cap=256)
    ctl, lanes = _ctl(cfg0), _lanes(gpu, L)
    res, win, nw = gpu.integ_run(acc_w, ctl, steps, data, lanes, cap=256)
    assert nw == o["n_windows"]
    assert np.array_equal(win, o["windows"])
    assert np.array_equal(res, o["results"])
    assert _ctl_list(ctl) == o["ctl"].tolist()
    _check_lanes(lanes, o, L)
    # the same stream in two calls: the state carries over
    cut = int(rng.integers(1, n))
    ctl2, lanes2 = _ctl(cfg0), _lanes(gpu, L)
    r1, w1, n1 = gpu.integ_run(acc_w, ctl2, steps[:cut], data[:cut], lanes2, cap=256)
    r2, w2, n2 = gpu.integ_run(acc_w, ctl2, steps[cut:], data[cut:], lanes2, cap=256)
    assert n1 + n2 == o["n_windows"]
    w2 = w2.copy()
    w2[:, 0] += cut
    assert np.array_equal(np.concatenate([w1, w2]), o["windows"])
    assert np.array_equal(np.concatenate([r1, r2]), o["results"])
    _check_lanes(lanes2, o, L)


def test_integ_run_capacity(gpu, orc):
    rng = np.random.default_rng(2)
    steps = random_steps(rng, 200, p_cfg=0.0, max_len=1)
    data = random_data(rng, 200, 7, "small")
    o = orc.integ_run(40, [1, 0, 0, 1], steps, data, cap=10)
    res, win, nw = gpu.integ_run(40, _ctl([1, 0, 0, 1]), steps, data, _lanes(gpu, 7), cap=10)
    assert nw == o["n_windows"] > 10
    assert np.array_equal(res, o["results"]) and np.array_equal(win, o["windows"])


@pytest.mark.parametrize("CH,F,N,acc_w", [(4, 37, 5, 40), (16, 64, 9, 36), (33, 20, 3, 64), (64, 12, 2, 33)])
def test_cov_integrate_windows_matches_oracle(gpu, orc, CH, F, N, acc_w):
    rng = np.random.default_rng(CH * 3 + F)
    for Xa in (stim.rand_cx(rng, F * N * CH), stim.rand_extreme(rng, F * N * CH)):
        steps = random_steps(rng, F)
        o = orc.cov_integrate_windows(Xa, CH, F, N, acc_w, steps[0, :4], steps, cap=64, threads=16)
        n_l = N * CH * (CH + 1)
        ctl, lanes = _ctl(steps[0, :4]), _lanes(gpu, n_l)
        res, win, nw = gpu.cov_integrate_windows(Xa, CH, F, N, acc_w, ctl, steps, lanes, cap=64)
        assert nw == o["n_windows"]
        assert np.array_equal(win, o["windows"]) and np.array_equal(res, o["results"])
        assert _ctl_list(ctl) == o["ctl"].tolist()
        _check_lanes(lanes, o, n_l)


def test_cov_integrate_windows_one_window_equals_cov_run(gpu):
    rng = np.random.default_rng(6)
    CH, F, N = 16, 48, 7
    Xa = stim.rand_extreme(rng, F * N * CH)
    steps = np.array([[F, 0, 0, 1, 1, 0, 0]] * F, np.uint32)
    res, win, nw = gpu.cov_integrate_windows(Xa, CH, F, N, 64, _ctl([F, 0, 0, 1]), steps,
                                             _lanes(gpu, N * CH * (CH + 1)), cap=4)
    assert nw == 1 and win[0].tolist() == [F - 1, 0, F, 0, 0]
    assert np.array_equal(res[0], gpu.cov(Xa, CH, F, N))


@pytest.mark.parametrize("P,T,nb,ch", [(4, 8, 11, 2), (1024, 16, 5, 1), (8, 1, 4, 2), (32, 5, 3, 3), (64, 64, 2, 1)])
def test_pfb_acc_matches_oracle_and_rounds_to_pfb(gpu, orc, P, T, nb, ch):
    rng = np.random.default_rng(P + 3 * T + nb)
    for x, h in ((stim.rand_cx(rng, ch * nb * P), stim.rand_cx(rng, P * T)),
                 (stim.rand_extreme(rng, ch * nb * P), stim.rand_extreme(rng, P * T))):
        hist = stim.rand_extreme(rng, ch * (T - 1) * P) if T > 1 else None
        for hh in (None, hist):
            acc = gpu.pfb_acc(x, h, P, T, ch, nb, hh)
            assert np.array_equal(acc, orc.pfb_acc(x, h, P, T, ch, nb, hh))
        # round_sat(acc, 15, 16) is the PFB output (both with the zero history)
        y, _ = gpu.pfb(x, h, P, T, ch, nb)
        a = gpu.pfb_acc(x, h, P, T, ch, nb).reshape(-1, 2)
        assert np.array_equal(stim.pack(np.array([stim.rne_int(int(v), 15) for v in a[:, 0]]),
                                        np.array([stim.rne_int(int(v), 15) for v in a[:, 1]])), y)


CFGS = [[1, 0, 0, 2, 3, 8, 5, 0x300], [1, 1, 0, 1, 2, 4, 6, 0x200], [1, 1, 0, 0, 0, 3, 3, 0x100],
        [0, 0, 0, 2, 2, 8, 8, 0x800], [1, 1, 0, 2, 2, 0, 8, 0x800], [1, 0, 0, 4, 1, 1000, 2, 0x400],
        [1, 0, 0, 0xFFFFFFFF, 2, 3, 3, 0x100], [1, 0, 0, 0, 0, 0, 0, 0x100]]


@pytest.mark.parametrize("cfg", CFGS)
@pytest.mark.parametrize("n", [1, 97, 1024])
def test_cfar_decide_matches_oracle(gpu, orc, cfg, n):
    rng = np.random.default_rng(n + sum(cfg) % 997)
    rows = 3
    for bins in (rng.integers(0, 1 << 20, n * rows).astype(np.uint64),
                 rng.integers(0, 1 << 63, n * rows, dtype=np.uint64)):
        d = gpu.cfar_decide(bins, n, cfg)
        o = orc.cfar_decide(bins, n, cfg)
        for i, f in enumerate(("cut_power", "sum_lead", "sum_lag", "sel_sum", "sel_n", "window_ok")):
            assert np.array_equal(d[f].astype(np.uint64), o[:, i]), f
        assert np.array_equal(d["evaluable"] | (d["detected"] << 1) | (d["suppressed"] << 2), o[:, 6])
        assert np.array_equal(d["take_lead"].astype(np.uint64), o[:, 7])
