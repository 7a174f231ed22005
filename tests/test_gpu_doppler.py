"""GPU parity of the slow-time (Doppler) transform (rd_doppler_run) against
the oracle's composition (ring window gather, the FFT restatement along the
frame axis, optional fftshift), bit for bit, values and stage flags."""
import numpy as np
import pytest

import stim

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rows,depth,n_bins,newest,H,spc,sched,shift", [
    (4, 16, 64, 15, 16, 2, 0xFFFFFFFF, False),
    (3, 40, 100, 5, 32, 1, 0xFFFFFFFF, True),
    (16, 256, 1024, 255, 256, 2, 0xFFFFFFFF, True),
    (2, 70, 33, 69, 64, 2, 0x15, False),
    (1, 4096, 8, 100, 4096, 2, 0xFFFFFFFF, True),
    (5, 9, 130, 0, 8, 2, 0, False),
])
def test_doppler_matches_oracle(gpu, orc, rows, depth, n_bins, newest, H, spc, sched, shift):
    rng = np.random.default_rng(rows * 1000 + depth + n_bins + H)
    for ring in (stim.rand_cx(rng, rows * depth * n_bins), stim.rand_extreme(rng, rows * depth * n_bins)):
        g, gf = gpu.doppler(ring, rows, depth, n_bins, newest, H, spc, sched, shift)
        o, of = orc.doppler(ring, rows, depth, n_bins, newest, H, spc, sched, shift)
        assert np.array_equal(g, o)
        assert np.array_equal(gf, of)


def test_doppler_rejects_bad_geometry(gpu):
    import rd_amd

    ring = np.zeros(2 * 16 * 8, np.uint32)
    for depth, newest, H, code in [(16, 16, 16, rd_amd.RD_E_ARG), (16, 3, 32, rd_amd.RD_E_ARG),
                                   (16, 3, 12, rd_amd.RD_E_GEOM), (16, 3, 4, rd_amd.RD_E_GEOM)]:
        with pytest.raises(rd_amd.RdError) as e:
            gpu.doppler(ring, 2, depth, 8, newest, H)
        assert e.value.code == code
