"""GPU parity of the batched path (rd_batch_*): many CPIs of uneven geometry in
one graph launch, every CPI's every stage output and packed event list equal to
the oracle's composed chain for that CPI, bit for bit — with the heuristic
shapes, with the autotuned shapes, and across repeated launches.
"""
import numpy as np
import pytest

import rd_amd
import rd_scenario

pytestmark = pytest.mark.gpu

STAGES = ("u", "pfb_flags", "X", "fft_flags", "Xa", "y", "yflags", "power", "R")
ALL = rd_scenario.SCHED_ALL


def _scene(i, CH, N, T, F, B, spc=2, sched=ALL, cfg=None):
    rng = np.random.default_rng(100 + i)
    sc = rd_scenario.Scene(f"b{i}", CH=CH, N=N, T=T, F=F, B=B, sigma=256.0, guard=2, ref=8, spc=spc, sched=sched,
                           targets=rd_scenario._random_targets(rng, int(rng.integers(0, 6)), N),
                           beam_angles=list(np.linspace(-50.0, 50.0, B)), seed=300 + i)
    cfg = sc.cfar_cfg() if cfg is None else np.array(cfg, np.uint32)
    return sc, cfg


# Shape classes of every stage are mixed: PFB taps 16/8/4 and generic 5;
# FFT lengths 64..4096, spc 1 and 2, partial schedules; beamformer steps
# 1..4 with beams above and below the channel count; covariance chp 16..64;
# CFAR CA/GO/dense, rows of 64..4096 bins.
SPECS = [
    dict(CH=8, N=256, T=16, F=6, B=8),
    dict(CH=16, N=1024, T=16, F=4, B=16),
    dict(CH=64, N=4096, T=16, F=2, B=64),
    dict(CH=32, N=256, T=16, F=5, B=32),
    dict(CH=7, N=64, T=5, F=9, B=3, spc=1, sched=0x15, cfg=[1, 1, 0, 1, 2, 4, 3, 0x300]),
    dict(CH=20, N=128, T=8, F=17, B=40, cfg=[1, 0, 1, 1, 1, 4, 4, 0x200]),
    dict(CH=33, N=512, T=4, F=3, B=17, spc=1),
    dict(CH=16, N=1024, T=16, F=40, B=16),
    dict(CH=48, N=256, T=16, F=3, B=5, sched=0x0F0),
    dict(CH=8, N=256, T=16, F=33, B=8),
]


def _build(gpu, specs, lean=False):
    chains, exp = [], []
    for i, s in enumerate(specs):
        s = dict(s)
        cfg = s.pop("cfg", None)
        sc, cfg = _scene(i, cfg=cfg, **s)
        ch = rd_amd.Chain(gpu, sc.CH, sc.N, sc.T, sc.F, sc.B, cfg, spc=sc.spc, sched=sc.sched)
        iq, taps, W = sc.iq(), sc.taps(), sc.weights()
        ch.load(iq, taps, W)
        chains.append(ch)
        exp.append((sc, cfg, iq, taps, W))
    return chains, exp


def _check(orc, chains, exp, totals, checked=STAGES):
    for i, (ch, (sc, cfg, iq, taps, W)) in enumerate(zip(chains, exp)):
        o = orc.chain(iq, taps, W, sc.CH, sc.N, sc.T, sc.F, sc.B, cfg, 31, 63, spc=sc.spc, sched=sc.sched,
                      keep=STAGES + ("events",), threads=16)
        assert int(totals[i]) == o["n_events"], (i, int(totals[i]), o["n_events"])
        got = ch.fetch(checked, n_events=int(totals[i]))
        for k in checked + ("row_count",):
            bad = np.flatnonzero(got[k] != o[k])
            assert bad.size == 0, f"CPI {i} {sc.CH}x{sc.N}x{sc.F} stage {k}: {bad.size} mismatches (first {bad[:4]})"
        assert np.array_equal(got["events"], o["events"]), f"CPI {i}: events differ"


@pytest.mark.parametrize("autotune", [False, True])
def test_batch_matches_oracle_per_cpi(gpu, orc, autotune):
    chains, exp = _build(gpu, SPECS)
    b = rd_amd.Batch.of_chains(gpu, chains, autotune=autotune)
    for ch in chains:  # the batch launch alone must produce every output
        for k in ("u", "X", "Xa", "y", "power", "R", "events", "row_count"):
            ch.b[k].zero()
    b.launch()
    totals = b.events()
    plan = b.plan()
    assert "pfb" in plan and "cfar_gather" in plan
    if autotune:
        assert "tuned" in plan
    _check(orc, chains, exp, totals)
    b.close()


def test_batch_replays_with_eager_work_between(gpu, orc):
    """Lean batch graph replayed three times; between replays every output is
    clobbered and an eager single-CPI chain runs on the same context."""
    chains, exp = _build(gpu, SPECS[:6])
    b = rd_amd.Batch.of_chains(gpu, chains, lean=True)
    other = rd_amd.Chain(gpu, 4, 64, 8, 16, 4, rd_scenario.config("tiny").cfar_cfg())
    sc = rd_scenario.config("tiny")
    other.load(sc.iq(seed=5), sc.taps(), sc.weights())
    for rep in range(3):
        for ch in chains:
            for k in ("u", "X", "Xa", "power", "R", "events", "row_count"):
                ch.b[k].zero()
        other.run()
        b.launch()
        totals = b.events()
        _check(orc, chains, exp, totals, checked=("u", "X", "Xa", "power", "R"))
    b.close()


def test_batch_equals_single_cpi_launches(gpu):
    """Two batches over the same CPIs (heuristic and tuned shapes) and the
    one-CPI-per-launch chain give identical outputs."""
    specs = [dict(CH=c, N=n, T=16, F=f, B=c) for c, n, f in
             [(8, 256, 40), (16, 1024, 12), (32, 4096, 3), (64, 256, 20), (8, 4096, 5), (32, 1024, 7)]]
    chains, _ = _build(gpu, specs)
    ref = []
    for ch in chains:
        n = ch.run()
        ref.append((n, ch.fetch(("u", "X", "Xa", "power", "R"), n_events=n)))
    for tune in (False, True):
        b = rd_amd.Batch.of_chains(gpu, chains, autotune=tune, lean=True)
        for ch in chains:
            for k in ("u", "X", "Xa", "power", "R", "events"):
                ch.b[k].zero()
        b.launch()
        totals = b.events()
        for i, ch in enumerate(chains):
            n, r = ref[i]
            assert int(totals[i]) == n
            got = ch.fetch(("u", "X", "Xa", "power", "R"), n_events=n)
            for k in ("u", "X", "Xa", "power", "R", "row_count", "events"):
                assert np.array_equal(got[k], r[k]), (tune, i, k)
        b.close()


def test_batch_loader_fed_classes_cross_cpis(gpu, orc):
    """Loader-fed batch classes where the persistent beamformer's blocks move
    from one CPI's tiles to the next (two 48-channel CPIs of 512 tiles of 128
    columns each, more tiles than resident blocks) and the 32-channel
    covariance class: every stage of every CPI against the oracle."""
    specs = [dict(CH=48, N=1024, T=16, F=64, B=16), dict(CH=48, N=1024, T=16, F=64, B=32),
             dict(CH=32, N=256, T=16, F=20, B=32)]
    chains, exp = _build(gpu, specs)
    b = rd_amd.Batch.of_chains(gpu, chains)
    assert b.plan().count("loader-fed") == 2, b.plan()
    b.launch()
    _check(orc, chains, exp, b.events())
    b.close()


def test_batch_capacity_reported_per_cpi(gpu, orc):
    sc = rd_scenario.config("tiny")
    chains = []
    for cap in (None, 5):
        ch = rd_amd.Chain(gpu, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg(), events_cap=cap)
        ch.load(sc.iq(), sc.taps(), sc.weights())
        chains.append(ch)
    b = rd_amd.Batch.of_chains(gpu, chains)
    b.launch()
    out = np.zeros(2, np.uint64)
    rc = gpu.L.rd_batch_events(b.h, out.ctypes.data)
    assert rc == rd_amd.RD_E_CAPACITY
    o = orc.chain(sc.iq(), sc.taps(), sc.weights(), sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg(), 31, 63,
                  keep=("events",))
    assert int(out[0]) == int(out[1]) == o["n_events"]
    got = chains[1].fetch((), n_events=5)["events"]
    assert np.array_equal(got, o["events"][:5])
    b.close()


def test_batch_rejects_bad_cpi(gpu):
    sc = rd_scenario.config("tiny")
    ch = rd_amd.Chain(gpu, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg())
    bad = rd_amd.Chain(gpu, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg())
    bad.desc.n_fft = 48
    with pytest.raises(rd_amd.RdError) as e:
        rd_amd.Batch.of_chains(gpu, [ch, bad])
    assert e.value.code == rd_amd.RD_E_GEOM
