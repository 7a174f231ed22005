"""The allocation-free chain entry point (rd_chain_enqueue_static) against the
allocating one (rd_chain_enqueue): every stage output, the covariance, the row
counts, the packed events and the event total identical, bit for bit, on two
calls in a row (the second CPI smaller than the first, so the static array is
reused with stale contents from the first), and its size limit reported as
RD_E_SCRATCH before anything is launched.
"""
import ctypes as C

import numpy as np
import pytest

import rd_amd
import rd_scenario

pytestmark = pytest.mark.gpu

OUTS = ("u", "pfb_flags", "X", "fft_flags", "Xa", "y", "yflags", "power", "R")


def _scene(CH, N, T, F, B, seed, cfg=None):
    rng = np.random.default_rng(seed)
    sc = rd_scenario.Scene(f"s{seed}", CH=CH, N=N, T=T, F=F, B=B, sigma=256.0, guard=2, ref=8,
                           targets=rd_scenario._random_targets(rng, 5, N),
                           beam_angles=list(np.linspace(-40.0, 40.0, B)), seed=seed)
    return sc, (sc.cfar_cfg() if cfg is None else np.array(cfg, np.uint32))


def _run(gpu, sc, cfg, static):
    ch = rd_amd.Chain(gpu, sc.CH, sc.N, sc.T, sc.F, sc.B, cfg)
    ch.load(sc.iq(), sc.taps(), sc.weights())
    ch.enqueue(static=static)
    gpu.sync()
    n = int(ch.total.download(np.uint64, 1)[0])
    out = ch.fetch(OUTS, n_events=n)
    out["total"] = n
    return out


# the first CPI larger than the second; the dense CFAR mode of the first makes
# rows emit more events than the static array stages per row at that size
CASES = [
    [dict(CH=16, N=1024, T=16, F=64, B=16, seed=11, cfg=[1, 0, 1, 4, 4, 16, 16, 0x800]),
     dict(CH=8, N=256, T=16, F=12, B=8, seed=12)],
    [dict(CH=64, N=4096, T=16, F=8, B=64, seed=21), dict(CH=4, N=64, T=8, F=16, B=4, seed=22)],
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_static_entry_point_matches_allocating_one(gpu, case):
    for spec in CASES[case]:
        spec = dict(spec)
        seed, cfg = spec.pop("seed"), spec.pop("cfg", None)
        sc, cfg = _scene(seed=seed, cfg=cfg, **spec)
        print(f"static vs allocating: {spec} seed {seed}", flush=True)
        want = _run(gpu, sc, cfg, static=False)
        got = _run(gpu, sc, cfg, static=True)
        assert got["total"] == want["total"], (got["total"], want["total"])
        for k in OUTS + ("row_count", "events"):
            bad = np.flatnonzero(np.asarray(got[k]) != np.asarray(want[k]))
            assert bad.size == 0, f"{k}: {bad.size} mismatches (first {bad[:4]})"


def test_static_entry_point_rejects_rows_beyond_its_scratch(gpu):
    B, F = 128, 65535  # B * F rows > RD_STATIC_MAX_ROWS
    assert B * F > rd_amd.RD_STATIC_MAX_ROWS
    ch = rd_amd.Chain(gpu, 4, 64, 8, 2, 4, np.array([1, 0, 0, 2, 2, 8, 8, 0x800], np.uint32))
    desc = rd_amd.CpiDesc.from_buffer_copy(ch.desc)
    desc.n_beams, desc.n_frames = B, F
    rc = gpu.L.rd_chain_enqueue_static(gpu.h, C.byref(desc), C.byref(ch.bufs), ch.total.ptr)
    assert rc == rd_amd.RD_E_SCRATCH
    assert b"static scratch" in gpu.lib.lib.rd_status_str(rc)
    gpu.sync()
