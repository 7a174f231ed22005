"""The reference file formats read and written by the library's host side
(rd_read_coeff, rd_read_pfb_vec, rd_read_iq_ascii, rd_write_iq_ascii): every
committed .coeff/.vec fixture of the reference against the test-side reader,
the reference loaders' rejections, and the iq_ascii_v1 layout."""
import numpy as np
import pytest

import rd_amd
import stim
import vecio


@pytest.fixture(scope="module")
def lib():
    return rd_amd.load()


@pytest.mark.parametrize("set_name,geom", vecio.PFB_SETS)
def test_coeff_and_vec_match_the_fixtures(lib, set_name, geom):
    P, T, coeff, v = vecio.load_pfb(set_name, geom)
    c, p, t = lib.read_coeff(vecio.VEC_DIR / f"pfb_{set_name}_{geom}.coeff")
    assert (p, t) == (P, T) and np.array_equal(c, coeff)
    r = lib.read_pfb_vec(vecio.VEC_DIR / f"pfb_{set_name}_{geom}.vec")
    assert (r["phases"], r["taps"], r["beats"]) == (P, T, v["beats"])
    for k in ("x", "y", "f_re", "f_im"):
        assert np.array_equal(r[k], v[k]), k


def _mutate(src, dst, fn):
    lines = src.read_text().splitlines()
    dst.write_text("\n".join(fn(lines)) + "\n")
    return dst


@pytest.mark.parametrize("what,fn", [
    ("index order", lambda L: [l for l in L if not l.startswith("1 ")]),
    ("kind", lambda L: [l.replace("pfb_coeff", "pfb_io") for l in L]),
    ("count", lambda L: [l.replace("# count: 32", "# count: 31") for l in L]),
    ("phase-major", lambda L: [("1 1 0 " + l.split(None, 3)[3]) if l.startswith("1 0 1 ") else l for l in L]),
    ("fields", lambda L: [l + " 7" if l.startswith("5 ") else l for l in L]),
])
def test_coeff_rejections(lib, tmp_path, what, fn):
    bad = _mutate(vecio.VEC_DIR / "pfb_proto_p4t8.coeff", tmp_path / "bad.coeff", fn)
    with pytest.raises(rd_amd.RdError) as e:
        lib.read_coeff(bad)
    assert e.value.code == rd_amd.RD_E_ARG, what
    assert "bad.coeff" in str(e.value)


def test_missing_file(lib, tmp_path):
    with pytest.raises(rd_amd.RdError) as e:
        lib.read_coeff(tmp_path / "none.coeff")
    assert "cannot open" in str(e.value)


@pytest.mark.parametrize("A,N,H,P", [(4, 64, 16, 2), (3, 32, 5, 1), (2, 256, 3, 4)])
def test_iq_ascii_round_trip_and_layout(lib, tmp_path, A, N, H, P):
    rng = np.random.default_rng(A + N + H)
    iq = stim.rand_extreme(rng, A * H * N)
    path = tmp_path / "scene.iq"
    lib.write_iq_ascii(path, iq, A, N, H, P, "round_trip")
    back, a, n, h, p = lib.read_iq_ascii(path)
    assert (a, n, h, p) == (A, N, H, P) and np.array_equal(back, iq)
    # layout as documented by the writer of the format: header keys, then rows
    # in (frame, antenna, beat) order with sof/eof on the first/last beat
    text = path.read_text().splitlines()
    hdr = {l[1:].split(":", 1)[0].strip(): l.split(":", 1)[1].strip() for l in text if l.startswith("#") and ":" in l}
    assert hdr["format"] == "iq_ascii_v1" and int(hdr["N_ANTENNAS"]) == A and int(hdr["FFT_SIZE"]) == N
    assert int(hdr["SAMPLES_PER_CYCLE"]) == P and int(hdr["total_frames"]) == H
    rows = [list(map(int, l.split())) for l in text if l and not l.startswith("#")]
    bpf = N // P
    assert len(rows) == H * A * bpf
    for i, r in enumerate(rows):
        f, rest = divmod(i, A * bpf)
        a_, b = divmod(rest, bpf)
        assert r[:5] == [f, a_, b, int(b == 0), int(b == bpf - 1)]
        for k in range(P):
            w = int(iq[a_ * H * N + f * N + b * P + k])
            re, im = w & 0xFFFF, w >> 16
            assert r[5 + 2 * k] == re - (re >> 15) * 65536 and r[6 + 2 * k] == im - (im >> 15) * 65536


def test_iq_ascii_rejects_broken_framing(lib, tmp_path):
    iq = np.arange(2 * 16 * 2, dtype=np.uint32)
    path = tmp_path / "s.iq"
    lib.write_iq_ascii(path, iq, 2, 16, 2, 2)
    for what, fn in [("sof", lambda L: [l.replace("0 0 0 1 0", "0 0 0 0 0", 1) if l.startswith("0 0 0 1 0") else l
                                        for l in L]),
                     ("duplicate", lambda L: L + [next(l for l in L if l.startswith("1 1 3 "))]),
                     ("missing", lambda L: [l for l in L if not l.startswith("1 1 7 ")])]:
        bad = _mutate(path, tmp_path / f"{what}.iq", fn)
        with pytest.raises(rd_amd.RdError) as e:
            lib.read_iq_ascii(bad)
        assert e.value.code == rd_amd.RD_E_ARG, what
