"""Readers for the reference's committed golden-vector formats.

The files under ``tests/golden/ref_vectors/`` are the reference's own data
(``model/vectors/``).  Their line formats are documented in their headers and in
``model/vectors/README.md``; the loaders here check the declared schema, kind,
rounding mode, count and (for the FFT) twiddle digest, so a stale or truncated
file fails loudly instead of passing short (the same policy as
``model/cpp/fft/fft_vectors.hpp:204-244`` and ``model/cpp/pfb/pfb_model.hpp:409-435``).
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

VEC_DIR = Path(__file__).resolve().parent / "golden" / "ref_vectors"
TWIDDLE_DIGEST_1024 = "c756048fbc00d64f"


def pack(re: int, im: int) -> int:
    return ((im & 0xFFFF) << 16) | (re & 0xFFFF)


def unpack(p: int) -> tuple[int, int]:
    p = int(p)
    re = p & 0xFFFF
    im = (p >> 16) & 0xFFFF
    return (re - 0x10000 if re & 0x8000 else re, im - 0x10000 if im & 0x8000 else im)


def _read(path: Path):
    header: dict[str, str] = {}
    rows: list[list[str]] = []
    for raw in path.read_text().splitlines():
        line = raw.rstrip("\r")
        if line.startswith("#"):
            body = line[1:]
            if ":" in body:
                k, v = body.split(":", 1)
                header[k.strip()] = v.strip()
            continue
        if line.strip():
            rows.append(line.split())
    return header, rows


def _check(header, kind: str, path: Path):
    assert header.get("schema") == "1", f"{path}: schema {header.get('schema')}"
    assert header.get("kind") == kind, f"{path}: kind {header.get('kind')}"
    assert header.get("rounding_mode") == "nearest_even", f"{path}: rounding_mode"


def load_ops(path: Path = VEC_DIR / "fxp_ops.vec"):
    h, rows = _read(path)
    _check(h, "ops", path)
    assert int(h["count"]) == len(rows), "declared count"
    out = []
    for r in rows:
        assert len(r) == 8, r
        out.append(dict(id=r[0], op=r[1], a=int(r[2]), b=int(r[3]), sh=int(r[4]), w=int(r[5]), y=int(r[6]),
                        flags=int(r[7])))
    return out


def load_accum(path: Path = VEC_DIR / "fxp_accum.vec"):
    h, rows = _read(path)
    _check(h, "accum", path)
    assert int(h["count"]) == len(rows), "declared count"
    seqs: dict[str, list] = {}
    for r in rows:
        assert len(r) == 8, r
        seqs.setdefault(r[0], []).append(dict(step=int(r[1]), acc_w=int(r[2]), x=int(r[3]), y=int(r[4]),
                                              step_flags=int(r[5]), sticky=int(r[6]), count=int(r[7])))
    return seqs


def load_cmult(path: Path = VEC_DIR / "cmult.vec"):
    h, rows = _read(path)
    _check(h, "cmult", path)
    assert int(h["count"]) == len(rows), "declared count"
    out = []
    for r in rows:
        assert len(r) == 11, r
        v = [int(t) for t in r[1:]]
        out.append(dict(id=r[0], a=pack(v[0], v[1]), b=pack(v[2], v[3]), p_re=v[4], p_im=v[5],
                        y=pack(v[6], v[7]), f_re=v[8], f_im=v[9]))
    return out


def load_fft(path: Path = VEC_DIR / "fft64.vec"):
    """Returns (header, records); each record has in/out packed arrays in beat order."""
    h, rows = _read(path)
    _check(h, "fft", path)
    assert h.get("twiddle_digest") == TWIDDLE_DIGEST_1024, "twiddle digest"
    n = int(h["fft_size"])
    recs = []
    cur = None
    for r in rows:
        if r[0] == "vec":
            cur = dict(id=r[1], sched=int(r[2], 16), reorder=int(r[3]), flags=int(r[4], 16),
                       inp=np.zeros(n, np.uint32), out=np.zeros(n, np.uint32), seen=0)
            recs.append(cur)
        elif r[0] == "s":
            i = int(r[1])
            cur["inp"][i] = pack(int(r[2]), int(r[3]))
            cur["out"][i] = pack(int(r[4]), int(r[5]))
            cur["seen"] += 1
        else:
            raise AssertionError(f"unknown record tag {r[0]}")
    assert int(h["count"]) == len(recs), "declared count"
    for rec in recs:
        assert rec["seen"] == n, f"record {rec['id']} holds {rec['seen']} samples"
    return dict(n=n, spc=int(h["spc"]), stages=int(h["stages"])), recs


def load_pfb(set_name: str, geom: str):
    """Loads pfb_<set>_<geom>.coeff/.vec; returns (P, T, coeff[P*T], beats list)."""
    ch, crow = _read(VEC_DIR / f"pfb_{set_name}_{geom}.coeff")
    _check(ch, "pfb_coeff", VEC_DIR)
    P, T = int(ch["phases"]), int(ch["taps"])
    assert int(ch["count"]) == len(crow) == P * T
    coeff = np.zeros(P * T, np.uint32)
    for i, r in enumerate(crow):
        idx, ph, tap, re, im = (int(t) for t in r)
        assert idx == i == ph * T + tap, "phase-major, tap-minor order"
        coeff[i] = pack(re, im)
    vh, vrow = _read(VEC_DIR / f"pfb_{set_name}_{geom}.vec")
    _check(vh, "pfb_io", VEC_DIR)
    assert int(vh["phases"]) == P and int(vh["taps"]) == T
    assert int(vh["count"]) == len(vrow)
    nb = len(vrow)
    x = np.zeros(nb * P, np.uint32)
    y = np.zeros(nb * P, np.uint32)
    fre = np.zeros(nb * P, np.uint8)
    fim = np.zeros(nb * P, np.uint8)
    for b, r in enumerate(vrow):
        v = [int(t) for t in r]
        assert len(v) == 1 + 6 * P and v[0] == b
        for p in range(P):
            x[b * P + p] = pack(v[1 + 2 * p], v[2 + 2 * p])
            y[b * P + p] = pack(v[1 + 2 * P + 2 * p], v[2 + 2 * P + 2 * p])
            fre[b * P + p] = v[1 + 4 * P + 2 * p]
            fim[b * P + p] = v[2 + 4 * P + 2 * p]
    return P, T, coeff, dict(x=x, y=y, f_re=fre, f_im=fim, acc_w=int(vh.get("acc_w", 0)), beats=nb)


PFB_SETS = [(s, g) for s in ("ident", "max", "mixed", "proto", "random") for g in ("p4t8", "p8t16")]
