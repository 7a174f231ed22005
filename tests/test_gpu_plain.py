"""GPU parity of the plain (no matrix cores) contractions and chain that
bench.py measures as the "plain GPU" comparison point: bit-exact against the
oracle, like the MFMA path they are compared with.

Semantics: bf::dot + covar::power (model/cpp/beamformer/beamformer_model.hpp:
162-179, model/cpp/covariance/covar_model.hpp:99-102) and covar::cross_power
summed by an acc_w = 64 block-mode Integrator (covar_model.hpp:137-151,
186-262).
"""
import numpy as np
import pytest

import rd_amd
import rd_scenario
import stim
import vecio

pytestmark = pytest.mark.gpu

FULL_NEG = vecio.pack(-32768, -32768)


def _eq(a, b, what):
    bad = np.flatnonzero(np.asarray(a) != np.asarray(b))
    assert bad.size == 0, f"{what}: {bad.size} mismatches, first at {bad[:5]}"


@pytest.mark.parametrize("CH,B", [(4, 4), (16, 16), (64, 64), (7, 3), (33, 17), (64, 128), (1, 1)])
def test_bf_plain_matches_oracle(gpu, orc, CH, B):
    rng = np.random.default_rng(CH * 7 + B)
    cols = 333
    for Xa, W in [(stim.rand_cx(rng, cols * CH), stim.rand_cx(rng, B * CH)),
                  (stim.rand_extreme(rng, cols * CH), stim.rand_extreme(rng, B * CH)),
                  (np.full(cols * CH, FULL_NEG, np.uint32), np.full(B * CH, FULL_NEG, np.uint32))]:
        y, fl, pw = gpu.bf_plain(Xa, W, CH, B)
        oy, ofl, opw, _ = orc.bf(Xa, W, CH, B, want_acc=False)
        _eq(y, oy, f"y {CH}x{B}")
        _eq(fl, ofl, f"flags {CH}x{B}")
        _eq(pw, opw, f"power {CH}x{B}")


@pytest.mark.parametrize("CH,F,N", [(4, 16, 64), (16, 40, 33), (64, 20, 5), (1, 7, 9), (33, 35, 6), (64, 512, 2)])
def test_cov_plain_matches_oracle(gpu, orc, CH, F, N):
    rng = np.random.default_rng(CH + F * 3 + N)
    for Xa in (stim.rand_cx(rng, F * N * CH), stim.rand_extreme(rng, F * N * CH),
               np.full(F * N * CH, FULL_NEG, np.uint32)):
        _eq(gpu.cov_plain(Xa, CH, F, N), orc.cov(Xa, CH, F, N)[0], f"R {CH}x{F}x{N}")


@pytest.mark.parametrize("name,frames", [("tiny", None), ("medium", 12), ("large", 2)])
def test_plain_chain_matches_oracle(gpu, orc, name, frames):
    sc = rd_scenario.config(name)
    if frames:
        sc.F = frames
    stages = ("u", "X", "Xa", "y", "yflags", "power", "R")
    ch = rd_amd.Chain(gpu, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg(), rd_scenario.MAX_GUARD, rd_scenario.MAX_REF)
    iq, taps, W = sc.iq(), sc.taps(), sc.weights()
    ch.load(iq, taps, W)
    ch.enqueue(plain=True)
    gpu.sync()
    tot = int(ch.total.download(np.uint64, 1)[0])
    got = ch.fetch(stages, n_events=tot)
    exp = orc.chain(iq, taps, W, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg(), rd_scenario.MAX_GUARD,
                    rd_scenario.MAX_REF, keep=stages + ("events",), threads=16)
    for k in stages + ("row_count",):
        _eq(got[k], exp[k], f"{name} {k}")
    assert tot == exp["n_events"]
    _eq(got["events"], exp["events"], f"{name} events")
