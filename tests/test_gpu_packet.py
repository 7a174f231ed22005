"""GPU parity of the detection packets (rd_event_flits) and their writeback
(rd_event_writeback) against the oracle, whose restatements are checked
against packet_model::hdr_pack / make_flit and packet_model::Scoreboard:
random events at every packet width, a wrapping sequence number, planted and
random flit faults, a wrapping ring, and the chain's own events with the
counts read from the device total."""
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


@pytest.mark.parametrize("packet_w", [36, 64, 100, 177, 256, 512])
def test_event_writeback_vs_oracle(gpu, orc, packet_w):
    """Writeback on the GPU against the oracle: a clean stream, one stream
    with every planted fault kind on its own packet plus random bit flips,
    and a ring shorter than the stream (wrapping from addr0)."""
    import flitfault

    rng = np.random.default_rng(packet_w + 21)
    n = 700
    ev = rng.integers(0, 2**63, size=(n, 3), dtype=np.uint64)
    fl = orc.event_flits(ev, packet_w, dest=9, src=3, vc=1, seq0=65000)
    g_ring, g_st = gpu.event_writeback(fl, packet_w, port=9, seq0=65000)
    o_ring, o_st = orc.event_writeback(fl, packet_w, port=9, seq0=65000)
    assert np.array_equal(g_ring, o_ring) and np.array_equal(g_st, o_st)
    assert g_st.tolist() == [n, 0, 0, 0, 0]
    bad = fl
    for i, kind in enumerate(flitfault.ALL):
        bad = flitfault.plant(bad, 17 + 41 * i, kind, packet_w)
    flat = bad.reshape(-1)
    nbits = packet_w + 5
    for _ in range(60):
        f = int(rng.integers(0, n * fl.shape[1]))
        b = int(rng.integers(0, nbits))
        flat[f * fl.shape[2] + b // 32] ^= np.uint32(1 << (b % 32))
    for ring_words, addr0 in [(None, 0), (257, 200)]:
        g_ring, g_st = gpu.event_writeback(bad, packet_w, port=9, seq0=65000, ring_words=ring_words, addr0=addr0)
        o_ring, o_st = orc.event_writeback(bad, packet_w, port=9, seq0=65000, ring_words=ring_words, addr0=addr0)
        assert np.array_equal(g_ring, o_ring), (ring_words, addr0)
        assert np.array_equal(g_st, o_st), (g_st, o_st)
        assert g_st[1] > 0 and g_st[2] > 0 and g_st[3] > 0 and g_st[4] > 0


def test_event_writeback_after_the_chain_with_device_counts(gpu, orc):
    """The chain's events -> flits -> writeback, both counts read from the
    chain's device total: the ring holds every event's low 176 bits in order,
    and a second call continuing at the next address adds to the same
    counters."""
    import flitfault

    sc = rd_scenario.config("medium")
    sc.F = 8
    ch = rd_amd.Chain(gpu, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg())
    ch.load(sc.iq(), sc.taps(), sc.weights())
    ch.enqueue()
    gpu.sync()
    n = int(ch.total.download(np.uint64, 1)[0])
    cap = ch.events_cap
    ev = ch.fetch((), n_events=n)["events"]
    L, fw = int(gpu.L.rd_event_flit_count(512)), int(gpu.L.rd_flit_words(512))
    flits = gpu.buf(4 * cap * L * fw)
    rc = gpu.L.rd_event_flits(gpu.h, ch.b["events"].ptr, cap, ch.total.ptr, 512, 0, 0, 0, 0, flits.ptr)
    gpu.lib.check("rd_event_flits", rc)
    R = 2 * n + 3
    ring = gpu.buf(64 * R).zero()
    stats = gpu.buf(8 * rd_amd.RD_WB_NSTATS).zero()
    g_ring, g_st = gpu.event_writeback(flits, 512, ring=ring, stats=stats, ring_words=R, n_dev=ch.total, cap=cap)
    assert g_st.tolist() == [n, 0, 0, 0, 0]
    assert np.array_equal(g_ring, flitfault.ring_of_events(ev, R, 0))
    # a second delivery of the same packets at the next addresses: the seq
    # numbers now restart, which the counters report as n breaks
    g_ring, g_st = gpu.event_writeback(flits, 512, seq0=n, ring=ring, stats=stats, ring_words=R, addr0=n,
                                       n_dev=ch.total, cap=cap)
    assert g_st.tolist() == [2 * n, 0, 0, 0, n]
    want = flitfault.ring_of_events(ev, R, 0)
    want[n:2 * n] = want[:n]
    assert np.array_equal(g_ring, want)


def test_event_writeback_rejects_bad_arguments(gpu):
    d = gpu.buf(256)
    st = gpu.buf(8 * rd_amd.RD_WB_NSTATS)
    for pw, port, rw, off in [(35, 0, 4, 0), (513, 0, 4, 0), (512, 16, 4, 0), (512, 0, 0, 0), (512, 0, 2**48 + 1, 0),
                              (512, 0, 2, 4)]:
        rc = gpu.L.rd_event_writeback(gpu.h, d.ptr, 1, None, pw, port, 0, d.ptr + off, rw, 0, st.ptr)
        assert rc == rd_amd.RD_E_ARG, (pw, port, rw, off)
    assert gpu.L.rd_event_writeback(gpu.h, d.ptr, 1, None, 512, 0, 0, d.ptr, 2, 0, None) == rd_amd.RD_E_ARG
    # Counts past one dispatch's work-item range are refused before any launch.
    assert gpu.L.rd_event_writeback(gpu.h, d.ptr, 2**31 + 1, None, 512, 0, 0, d.ptr, 2, 0, st.ptr) == rd_amd.RD_E_ARG
    assert gpu.L.rd_event_flits(gpu.h, d.ptr, 2**30 + 1, None, 512, 0, 0, 0, 0, d.ptr) == rd_amd.RD_E_ARG
    assert gpu.L.rd_event_flits(gpu.h, d.ptr, 2**62, None, 512, 0, 0, 0, 0, d.ptr) == rd_amd.RD_E_ARG
