"""CPU oracle for the bit-exact radar chain — TEST INFRASTRUCTURE ONLY.

Loads ``oracle/librd_oracle.so`` (the C++ restatement of the reference's
``model/cpp`` semantics, ``rd_oracle.cpp``) and, when it was built in a tree
that had the reference sources, ``oracle/_ref/librd_ref.so`` (the reference's
own header models behind a C shim).  Both expose the same entry points with
prefixes ``rdo_`` and ``rdr_``; :func:`load` returns a numpy-friendly wrapper
for either.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may use this module, and only as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "librd_oracle.so"
REF_SO = HERE / "_ref" / "librd_ref.so"

_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Oracle:
    """Thin wrapper; ``prefix`` is ``rdo`` (restatement) or ``rdr`` (reference)."""

    def __init__(self, path: Path, prefix: str):
        self.lib = C.CDLL(str(path))
        self.prefix = prefix
        self.path = path

    def _f(self, name):
        return getattr(self.lib, f"{self.prefix}_{name}")

    # -- numerics --------------------------------------------------------------
    def fxp_op(self, op: str, a: int, b: int, sh: int, w: int):
        y = C.c_int64()
        fl = C.c_uint()
        ok = self._f("fxp_op")(op.encode(), C.c_int64(a), C.c_int64(b), C.c_uint(sh), C.c_uint(w),
                               C.byref(y), C.byref(fl))
        if not ok:
            raise ValueError(f"unknown op {op}")
        return y.value, fl.value

    def cmult(self, a: int, b: int):
        vals = [C.c_int64() for _ in range(4)]
        y = C.c_uint32()
        fr = C.c_uint()
        fi = C.c_uint()
        self._f("cmult")(C.c_uint32(a), C.c_uint32(b), *[C.byref(v) for v in vals], C.byref(y), C.byref(fr),
                         C.byref(fi))
        return dict(p_re=vals[0].value, p_im=vals[1].value, p3_re=vals[2].value, p3_im=vals[3].value,
                    y=y.value, f_re=fr.value, f_im=fi.value)

    def acc_run(self, acc_w: int, xs):
        x = np.ascontiguousarray(xs, dtype=np.int64)
        n = x.size
        y = np.zeros(n, np.int64)
        sf = np.zeros(n, np.uint32)
        st = np.zeros(n, np.uint32)
        cnt = np.zeros(n, np.uint32)
        self.lib.rdo_acc_run(C.c_uint(acc_w), _ptr(x), C.c_uint(n), _ptr(y), _ptr(sf), _ptr(st), _ptr(cnt))
        return y, sf, st, cnt

    def twiddle_table(self, M: int = 4096):
        out = np.zeros(M, np.uint32)
        if self.prefix == "rdr":
            self.lib.rdr_twiddle_table(_ptr(out))
        else:
            self.lib.rdo_twiddle_table(C.c_uint(M), _ptr(out))
        return out

    def power(self, s: int) -> int:
        f = self._f("power")
        f.restype = C.c_int64
        return f(C.c_uint32(s))

    def cross_power(self, x: int, y: int):
        v = [C.c_int64() for _ in range(4)]
        self._f("cross_power")(C.c_uint32(x), C.c_uint32(y), *[C.byref(t) for t in v])
        return tuple(t.value for t in v)

    # -- stages ----------------------------------------------------------------
    def pfb(self, x, coeff, P, T, n_chan, n_beats, threads=1, hist=None):
        x = np.ascontiguousarray(x, np.uint32)
        coeff = np.ascontiguousarray(coeff, np.uint32)
        y = np.zeros(n_chan * n_beats * P, np.uint32)
        fl = np.zeros(n_chan * n_beats * P, np.uint8)
        if self.prefix == "rdr":
            if hist is not None:
                raise ValueError("reference shim has no history import")
            self.lib.rdr_pfb_run(_ptr(x), _ptr(coeff), C.c_uint(P), C.c_uint(T), C.c_uint(n_chan),
                                 C.c_uint(n_beats), _ptr(y), _ptr(fl), C.c_uint(threads))
        else:
            self.lib.rdo_pfb_run(_ptr(x), _ptr(coeff), C.c_uint(P), C.c_uint(T), C.c_uint(n_chan),
                                 C.c_uint(n_beats), _ptr(hist), _ptr(y), _ptr(fl), C.c_uint(threads))
        return y, fl

    def fft(self, x, n, spc=2, sched=0xFFFFFFFF, reorder=1, threads=1):
        x = np.ascontiguousarray(x, np.uint32)
        frames = x.size // n
        out = np.zeros(frames * n, np.uint32)
        bins = np.zeros(frames * n, np.uint32)
        sf = np.zeros(frames, np.uint32)
        rc = self._f("fft_run")(_ptr(x), C.c_uint(n), C.c_uint(spc), C.c_uint32(sched), C.c_int(reorder),
                                C.c_uint(frames), _ptr(out), _ptr(bins), _ptr(sf), C.c_uint(threads))
        if rc != 0:
            raise ValueError(f"invalid FFT geometry n={n} spc={spc}")
        return out, bins, sf

    def align(self, X, n_chan, n_frames, n_bins):
        X = np.ascontiguousarray(X, np.uint32)
        Xa = np.zeros_like(X)
        self.lib.rdo_align_run(_ptr(X), C.c_uint(n_chan), C.c_uint(n_frames), C.c_uint(n_bins), _ptr(Xa))
        return Xa

    def bf(self, Xa, W, CH, B, threads=1, want_acc=True):
        Xa = np.ascontiguousarray(Xa, np.uint32)
        W = np.ascontiguousarray(W, np.uint32)
        cols = Xa.size // CH
        y = np.zeros(B * cols, np.uint32)
        fl = np.zeros(B * cols, np.uint8)
        pw = np.zeros(B * cols, np.uint32)
        acc = np.zeros(B * cols * 2, np.int64) if want_acc else None
        self._f("bf_run")(_ptr(Xa), _ptr(W), C.c_uint(CH), C.c_uint(B), C.c_size_t(cols), _ptr(y), _ptr(fl),
                          _ptr(pw), _ptr(acc), C.c_uint(threads))
        return y, fl, pw, acc

    def cov(self, Xa, CH, F, N, acc_w=64, threads=1):
        Xa = np.ascontiguousarray(Xa, np.uint32)
        npairs = CH * (CH + 1) // 2
        R = np.zeros(N * npairs * 2, np.int64)
        sf = np.zeros(N * npairs, np.uint8)
        self._f("cov_run")(_ptr(Xa), C.c_uint(CH), C.c_uint(F), C.c_uint(N), C.c_uint(acc_w), _ptr(R), _ptr(sf),
                           C.c_uint(threads))
        return R, sf

    def cfar(self, bins, n, cfg8, max_guard, max_ref, frame_ids=None):
        bins = np.ascontiguousarray(bins, np.uint64)
        rows = bins.size // n
        cfg = np.ascontiguousarray(cfg8, np.uint32)
        fid = None if frame_ids is None else np.ascontiguousarray(frame_ids, np.uint32)
        cap = rows * (n + 1)
        ev = np.zeros(cap * 3, np.uint64)
        rc = np.zeros(rows, np.uint32)
        f = self._f("cfar_run")
        f.restype = C.c_int64
        tot = f(_ptr(bins), C.c_uint(n), C.c_uint(rows), _ptr(cfg), C.c_uint(max_guard), C.c_uint(max_ref),
                _ptr(fid), _ptr(ev), C.c_int64(cap), _ptr(rc))
        if tot < 0:
            raise RuntimeError("event overflow")
        return ev[: 3 * tot].reshape(tot, 3), rc

    def event_pack(self, fields11):
        f = np.ascontiguousarray(fields11, np.uint64)
        w = np.zeros(3, np.uint64)
        self.lib.rdo_event_pack(_ptr(f), _ptr(w))
        return w

    def doppler(self, ring, rows, depth, n_bins, newest, n_dop, spc=2, sched=0xFFFFFFFF, shift=False):
        """Slow-time transform: for every (row, bin) the FFT restatement across
        the last n_dop frames of the ring (frame window oldest first, newest in
        slot `newest`; fft_ref.hpp:409-412 semantics, natural-order bins), as
        [rows][n_dop][n_bins], optionally fftshifted on the Doppler axis
        (model/python/range_doppler.py:254-265). Also the stage flags [rows][n_bins]."""
        r = np.asarray(ring, np.uint32).reshape(rows, depth, n_bins)
        slots = (newest - (n_dop - 1) + np.arange(n_dop)) % depth
        cols = np.ascontiguousarray(r[:, slots, :].transpose(0, 2, 1)).reshape(-1)  # [rows][bins][n_dop]
        _, bins, sf = self.fft(cols, n_dop, spc, sched, 1)
        spec = bins.reshape(rows, n_bins, n_dop).transpose(0, 2, 1)
        if shift:
            spec = np.roll(spec, n_dop // 2, axis=1)
        return np.ascontiguousarray(spec).reshape(-1), sf.reshape(-1)

    def cov_integrate(self, Xa, CH, F, N, mode, exp_k, acc_w, threads=1):
        """Covariance through the Integrator, block (0) or exponential (1)
        mode; returns (R [N][pairs][2] int64, sticky sat flags [N][pairs])."""
        Xa = np.ascontiguousarray(Xa, np.uint32)
        npairs = CH * (CH + 1) // 2
        R = np.zeros(N * npairs * 2, np.int64)
        fl = np.zeros(N * npairs, np.uint8)
        self._f("cov_integrate")(_ptr(Xa), C.c_uint(CH), C.c_uint(F), C.c_uint(N), C.c_uint(mode), C.c_uint(exp_k),
                                 C.c_uint(acc_w), _ptr(R), _ptr(fl), C.c_uint(threads))
        return R, fl

    def _integ_out(self, n_lanes, cap):
        return dict(acc=np.zeros(n_lanes, np.int64), flags=np.zeros(n_lanes, np.uint32),
                    count=np.zeros(n_lanes, np.uint32), results=np.zeros(max(cap, 1) * n_lanes, np.int64),
                    windows=np.zeros((max(cap, 1), 5), np.uint32), ctl=np.zeros(6, np.uint32))

    def integ_run(self, acc_w, cfg4, steps7, data, cap=64):
        """covar::Integrator lanes from reset(cfg4) over a control stream
        steps7 [n_steps][7] = (window_len, mode, exp_k, enable, valid, flush,
        sat_clear) with data [n_steps][n_lanes] int64. Returns dict(acc, flags,
        count per lane, results [n_windows][n_lanes], windows [n_windows][5] =
        (step, window_id, sample_count, flushed, truncated), ctl (active cfg
        (4), count, window_id), n_windows)."""
        steps7 = np.ascontiguousarray(steps7, np.uint32).reshape(-1, 7)
        data = np.ascontiguousarray(data, np.int64).reshape(steps7.shape[0], -1)
        n_lanes = data.shape[1]
        o = self._integ_out(n_lanes, cap)
        f = self._f("integ_run")
        f.restype = C.c_uint32
        nw = f(C.c_uint(acc_w), _ptr(np.ascontiguousarray(cfg4, np.uint32)), _ptr(steps7),
               C.c_uint32(steps7.shape[0]), _ptr(data), C.c_size_t(n_lanes), _ptr(o["acc"]), _ptr(o["flags"]),
               _ptr(o["count"]), _ptr(o["results"]), _ptr(o["windows"]), C.c_uint32(cap), _ptr(o["ctl"]))
        k = min(nw, cap)
        o.update(n_windows=nw, results=o["results"][: k * n_lanes].reshape(k, n_lanes), windows=o["windows"][:k])
        return o

    def cov_integrate_windows(self, Xa, CH, F, N, acc_w, cfg4, steps7, cap=64, threads=1):
        """The Integrator lanes of a covariance: lane ((k*pairs + p)*2 + c) is
        component c of cross_power(X_i, X_j) of bin k, frame m = step m. Same
        return layout as integ_run (results [n_windows][N*pairs*2])."""
        Xa = np.ascontiguousarray(Xa, np.uint32)
        steps7 = np.ascontiguousarray(steps7, np.uint32).reshape(-1, 7)
        assert steps7.shape[0] == F
        n_lanes = N * CH * (CH + 1)
        o = self._integ_out(n_lanes, cap)
        f = self._f("cov_integrate_windows")
        f.restype = C.c_uint32
        nw = f(_ptr(Xa), C.c_uint(CH), C.c_uint(F), C.c_uint(N), C.c_uint(acc_w),
               _ptr(np.ascontiguousarray(cfg4, np.uint32)), _ptr(steps7), _ptr(o["acc"]), _ptr(o["flags"]),
               _ptr(o["count"]), _ptr(o["results"]), _ptr(o["windows"]), C.c_uint32(cap), _ptr(o["ctl"]),
               C.c_uint(threads))
        k = min(nw, cap)
        o.update(n_windows=nw, results=o["results"][: k * n_lanes].reshape(k, n_lanes), windows=o["windows"][:k])
        return o

    def pfb_acc(self, x, coeff, P, T, n_chan, n_beats, hist=None):
        """pfb::LaneModel::accumulate for every output: acc [n_chan][n_beats][P][2] int64."""
        acc = np.zeros(n_chan * n_beats * P * 2, np.int64)
        h = None if hist is None else np.ascontiguousarray(hist, np.uint32)
        self._f("pfb_acc")(_ptr(np.ascontiguousarray(x, np.uint32)), _ptr(np.ascontiguousarray(coeff, np.uint32)),
                           C.c_uint(P), C.c_uint(T), C.c_uint(n_chan), C.c_uint(n_beats), _ptr(h), _ptr(acc))
        return acc

    def cfar_decide(self, bins, n, cfg8):
        """cfar::decide for every bin of every row (configuration as given, no
        clamp): [rows*n][8] u64 = (cut_power, sum_lead, sum_lag, sel_sum, sel_n,
        window_ok, evaluable | detected<<1 | suppressed<<2, take_lead)."""
        bins = np.ascontiguousarray(bins, np.uint64)
        rows = bins.size // n
        dec = np.zeros(rows * n * 8, np.uint64)
        self._f("cfar_decide")(_ptr(bins), C.c_uint(n), C.c_uint(rows), _ptr(np.ascontiguousarray(cfg8, np.uint32)),
                               _ptr(dec))
        return dec.reshape(rows * n, 8)

    def event_flits(self, events, packet_w, dest=0, src=0, vc=0, seq0=0):
        """Detection packets of packed events [n][3] -> flits [n][L][fw] u32."""
        ev = np.ascontiguousarray(events, np.uint64).reshape(-1, 3)
        n = ev.shape[0]
        L = 1 + (176 + packet_w - 1) // packet_w
        fw = (packet_w + 5 + 31) // 32
        out = np.zeros(max(n, 1) * L * fw, np.uint32)
        f = self._f("event_flits")
        f.restype = C.c_int
        rc = f(_ptr(ev), C.c_uint64(n), C.c_uint(packet_w), C.c_uint(dest), C.c_uint(src), C.c_uint(vc),
               C.c_uint(seq0), _ptr(out))
        if rc < 0:
            raise ValueError(f"event_flits failed ({rc})")
        return out[: n * L * fw].reshape(n, L, fw)

    def event_writeback(self, flits, packet_w, port=0, seq0=0, ring_words=None, addr0=0, ring=None):
        """Detection writeback of flits [n][L][fw] -> (ring [ring_words][16] u32,
        stats [5] u64). ring (optional) is updated in place."""
        fl = np.ascontiguousarray(flits, np.uint32)
        n = fl.shape[0]
        if ring is None:
            ring = np.zeros((max(n, 1) if ring_words is None else ring_words, 16), np.uint32)
        stats = np.zeros(5, np.uint64)
        f = self._f("event_writeback")
        f.restype = C.c_int
        rc = f(_ptr(fl), C.c_uint64(n), C.c_uint(packet_w), C.c_uint(port), C.c_uint(seq0), _ptr(ring),
               C.c_uint64(ring.shape[0]), C.c_uint64(addr0), _ptr(stats))
        if rc < 0:
            raise ValueError(f"event_writeback failed ({rc})")
        return ring, stats

    def scoreboard(self, flits, packet_w, port, events, dest=0, src=0, vc=0, seq0=0):
        """packet_model::Scoreboard over flits [n][L][fw] against packets built
        from events [m][3] (reference build only) -> dict of its counts."""
        fl = np.ascontiguousarray(flits, np.uint32)
        ev = np.ascontiguousarray(events, np.uint64).reshape(-1, 3)
        out = np.zeros(10, np.uint64)
        f = self.lib.rdr_scoreboard
        f.restype = C.c_int
        rc = f(_ptr(fl), C.c_uint64(fl.shape[0]), C.c_uint(packet_w), C.c_uint(port), _ptr(ev),
               C.c_uint64(ev.shape[0]), C.c_uint(dest), C.c_uint(src), C.c_uint(vc), C.c_uint(seq0), _ptr(out))
        if rc < 0:
            raise ValueError(f"scoreboard failed ({rc})")
        names = ("delivered", "lost", "duplicated", "misrouted", "corrupted", "reordered", "framing", "parity",
                 "unexpected", "injected")
        return {k: int(v) for k, v in zip(names, out)}

    def chain(self, iq, taps, W, CH, N, T, F, B, cfar_cfg8, max_guard, max_ref, spc=2, sched=0xFFFFFFFF,
              keep=("u", "X", "y", "power", "R", "events"), threads=1):
        """Run the composed chain for one CPI; returns a dict of the kept outputs."""
        iq = np.ascontiguousarray(iq, np.uint32)
        taps = np.ascontiguousarray(taps, np.uint32)
        W = np.ascontiguousarray(W, np.uint32)
        cfg = np.ascontiguousarray(cfar_cfg8, np.uint32)
        S = CH * F * N
        out = {}

        def mk(name, size, dt):
            if name in keep:
                out[name] = np.zeros(size, dt)
                return out[name]
            return None

        u = mk("u", S, np.uint32)
        pfl = mk("pfb_flags", S, np.uint8)
        X = mk("X", S, np.uint32)
        ffl = mk("fft_flags", CH * F, np.uint32)
        Xa = mk("Xa", S, np.uint32)
        y = mk("y", B * F * N, np.uint32)
        yfl = mk("yflags", B * F * N, np.uint8)
        pw = mk("power", B * F * N, np.uint32)
        R = mk("R", N * CH * (CH + 1), np.int64)
        rc = np.zeros(B * F, np.uint32)
        want_ev = "events" in keep
        cap = B * F * (N + 1) if want_ev else 0
        ev = np.zeros(max(cap, 1) * 3, np.uint64) if want_ev else None
        f = self._f("chain_run")
        f.restype = C.c_int64
        tot = f(_ptr(iq), _ptr(taps), _ptr(W), C.c_uint(CH), C.c_uint(N), C.c_uint(T), C.c_uint(F), C.c_uint(B),
                C.c_uint(spc), C.c_uint32(sched), _ptr(cfg), C.c_uint(max_guard), C.c_uint(max_ref), _ptr(u),
                _ptr(pfl), _ptr(X), _ptr(ffl), _ptr(Xa), _ptr(y), _ptr(yfl), _ptr(pw), _ptr(R), _ptr(ev),
                C.c_int64(cap), _ptr(rc), C.c_uint(threads))
        if tot < 0:
            raise RuntimeError(f"chain failed ({tot})")
        out["row_count"] = rc
        out["n_events"] = int(tot)
        if want_ev:
            out["events"] = ev[: 3 * tot].reshape(tot, 3)
        return out


_cache: dict = {}


def load(kind: str = "oracle") -> Oracle:
    """``kind`` = "oracle" (the restatement) or "ref" (the reference models)."""
    if kind not in _cache:
        if kind == "oracle":
            if not ORACLE_SO.exists():
                raise FileNotFoundError(f"{ORACLE_SO} not built (make -C oracle)")
            _cache[kind] = Oracle(ORACLE_SO, "rdo")
        elif kind == "ref":
            if not REF_SO.exists():
                raise FileNotFoundError(f"{REF_SO} not built (needs the reference tree)")
            _cache[kind] = Oracle(REF_SO, "rdr")
        else:
            raise ValueError(kind)
    return _cache[kind]


def ref_available() -> bool:
    return REF_SO.exists()


def build(quiet: bool = True) -> None:
    """Compile the oracle (and the reference shim when the reference is present)."""
    import subprocess

    cmd = ["make", "-B", "-C", str(HERE)]
    if os.environ.get("RD_REFERENCE"):
        cmd.append(f"REFERENCE={os.environ['RD_REFERENCE']}")
    subprocess.run(cmd, check=True, capture_output=quiet)
