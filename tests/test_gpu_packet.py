"""GPU parity of the detection packets (rd_event_flits) against the oracle,
whose restatement is itself checked against packet_model::hdr_pack / make_flit:
random events at every packet width, a wrapping sequence number, and the
chain's own events with the count read from the device total."""
import numpy as np
import pytest

import rd_amd
import rd_scenario

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("packet_w", [36, 64, 100, 128, 176, 177, 256, 511, 512])
def test_event_flits_random(gpu, orc, packet_w):
    rng = np.random.default_rng(packet_w + 1)
    ev = rng.integers(0, 2**63, size=(300, 3), dtype=np.uint64) | (
        rng.integers(0, 2, size=(300, 3), dtype=np.uint64) << np.uint64(63))
    for dest, src, vc, seq0 in [(0, 0, 0, 0), (15, 31, 3, 65400), (6, 11, 1, 77)]:
        g = gpu.event_flits(ev, packet_w, dest, src, vc, seq0)
        o = orc.event_flits(ev, packet_w, dest, src, vc, seq0)
        assert np.array_equal(g, o), (packet_w, dest, src, vc, seq0)


def test_event_flits_after_the_chain_with_device_count(gpu, orc):
    sc = rd_scenario.config("medium")
    sc.F = 8
    ch = rd_amd.Chain(gpu, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg())
    ch.load(sc.iq(), sc.taps(), sc.weights())
    ch.enqueue()  # the event total lands in ch.total (device)
    gpu.sync()
    n = int(ch.total.download(np.uint64, 1)[0])
    cap = ch.events_cap
    flits = gpu.event_flits(ch.b["events"], 128, dest=3, src=7, vc=1, seq0=9, n_dev=ch.total, cap=cap)
    ev = ch.fetch((), n_events=n)["events"]
    want = orc.event_flits(ev, 128, dest=3, src=7, vc=1, seq0=9)
    assert np.array_equal(flits[:n], want)


def test_event_flits_rejects_bad_geometry(gpu):
    ev = np.zeros((2, 3), np.uint64)
    for pw, dest, src, vc in [(35, 0, 0, 0), (513, 0, 0, 0), (64, 16, 0, 0), (64, 0, 32, 0), (64, 0, 0, 4)]:
        with pytest.raises(rd_amd.RdError) as e:
            gpu.event_flits(ev, pw, dest, src, vc)
        assert e.value.code == rd_amd.RD_E_ARG
