"""GPU parity of the published workloads at their stated sizes.

* large (64 ch, 4096-pt PFB/FFT, 16 taps, 512 frames, 64 beams, CA-CFAR 32/4):
  one whole CPI through rd_chain_run against the oracle's composed chain on
  the host's cores — every stage output, every flag, the covariance, the row
  counts and the packed events, bit for bit;
* the uneven batch (256 CPIs, channels {8,16,32,64}, FFT {256,1024,4096},
  frames U[32,512], beams = channels): one autotuned graph launch, every CPI's
  outputs equal to that CPI run alone (rd_chain_run), the smallest CPIs also
  against the oracle.

Inputs are generated on the device (rd_workloads.device_iq: the construction
of rd_scenario.Scene.iq in torch), so the large CPI's 134 M samples take a
second instead of half a minute on the host.
"""
import os

import numpy as np
import pytest

import rd_amd
import rd_scenario

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

STAGES = ("u", "pfb_flags", "X", "fft_flags", "Xa", "y", "yflags", "power", "R")
THREADS = max(1, min(16, os.cpu_count() or 1))


@pytest.fixture(scope="module")
def wl():
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    import rd_workloads

    return rd_workloads


def test_large_cpi_full_size(gpu, orc, wl):
    sc = rd_scenario.config("large")
    assert (sc.CH, sc.N, sc.T, sc.F, sc.B) == (64, 4096, 16, 512, 64)
    print(f"large CPI: {sc.samples} samples", flush=True)
    iq = wl.device_iq(torch, sc, 7001).cpu().numpy().view(np.uint32)
    torch.cuda.empty_cache()
    taps, W = sc.taps(), sc.weights()
    ch = rd_amd.Chain(gpu, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg(), rd_scenario.MAX_GUARD, rd_scenario.MAX_REF)
    ch.load(iq, taps, W)
    tot = ch.run()
    got = ch.fetch(STAGES, n_events=tot)
    del ch
    print(f"large CPI: {tot} events on the GPU; oracle on {THREADS} threads", flush=True)
    exp = orc.chain(iq, taps, W, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg(), rd_scenario.MAX_GUARD,
                    rd_scenario.MAX_REF, keep=STAGES + ("events",), threads=THREADS)
    for k in STAGES + ("row_count",):
        bad = np.flatnonzero(got[k] != exp[k])
        assert bad.size == 0, f"large: stage {k}: {bad.size} mismatches (first {bad[:4]})"
    assert tot == exp["n_events"]
    assert np.array_equal(got["events"], exp["events"]), "large: events differ"
    assert tot > sc.B * sc.F
    # the covariance-trace invariant in FP64 on the GPU: tr R[k] is the sum of
    # every channel's power in bin k over the frames (covar_model.hpp:99-102)
    tr = wl.covariance_trace_check(torch, got["Xa"], got["R"], sc.CH, sc.F, sc.N)
    assert tr["ok"], tr


def test_uneven_batch_256_matches_per_cpi(gpu, orc, wl):
    scenes = rd_scenario.uneven_batch(256)
    print(f"uneven batch: {len(scenes)} CPIs, {sum(s.samples for s in scenes)} samples", flush=True)
    bs = wl.BatchSet(torch, scenes, seed0=77_000)
    # one pass per CPI to size the event buffers (the count is exact even when short)
    totals_dev = torch.zeros(len(scenes), dtype=torch.int64, device="cuda")
    for i, (d, b) in enumerate(zip(bs.descs, bs.bufs)):
        gpu.lib.check("rd_chain_enqueue", gpu.L.rd_chain_enqueue(gpu.h, rd_amd.C.byref(d), rd_amd.C.byref(b),
                                                                 totals_dev.data_ptr() + 8 * i))
    gpu.sync()
    bs.fit_event_caps(torch, totals_dev.cpu().numpy())
    batch = rd_amd.Batch(gpu, bs.descs, bs.bufs, autotune=True, keep=bs)
    for t in bs.t:  # the batch launch alone must produce every output
        for k in ("u", "X", "Xa", "power", "R", "row_count"):
            t[k].zero_()
    torch.cuda.synchronize()
    batch.launch()
    totals = batch.events()
    print("uneven batch: launched; checking every CPI against its own launch", flush=True)
    assert np.all(totals > 0)
    bad = wl.verify_batch_vs_plain(torch, gpu, bs, batch, totals)
    assert bad == 0, f"{bad} CPIs differ from their one-CPI launch"
    orc_bad, n = wl.verify_batch_vs_oracle(torch, gpu, bs, totals, 3)
    assert n == 3 and orc_bad == 0
    batch.close()
    del bs
    torch.cuda.empty_cache()
