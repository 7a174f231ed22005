"""ctypes binding of librd_amd.so (the C ABI in include/rd_amd.h).

Plumbing for tests, the smoke entry point and bench.py: device buffers are
allocated and copied through the library's own rd_malloc / rd_h2d / rd_d2h, and
every compute call goes to the HIP kernels. There is no CPU fallback: if the
library is missing, :func:`load` raises.
"""
from __future__ import annotations

import ctypes as C
import weakref
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "librd_amd.so"

RD_OK, RD_E_ARG, RD_E_GEOM, RD_E_HIP, RD_E_CAPACITY, RD_E_NODEV, RD_E_SCRATCH = 0, -1, -2, -3, -4, -5, -6
RD_STATIC_SCRATCH_BYTES = 32 << 20
RD_STATIC_MAX_ROWS = RD_STATIC_SCRATCH_BYTES // 8 - 1

EXPORTS = [
    "rd_device_count", "rd_create", "rd_destroy", "rd_status_str", "rd_sync", "rd_malloc", "rd_free", "rd_h2d",
    "rd_d2h", "rd_memset", "rd_stream", "rd_twiddle_table", "rd_pfb_run", "rd_fft_run", "rd_align_run",
    "rd_bf_run", "rd_bf_stage", "rd_bf_cov_stage", "rd_fft_stage", "rd_cov_run", "rd_cfar_run", "rd_chain_run", "rd_chain_enqueue", "rd_chain_graph",
    "rd_graph_launch", "rd_graph_destroy", "rd_batch_create", "rd_batch_launch", "rd_batch_events", "rd_batch_plan",
    "rd_batch_destroy", "rd_event_flit_count", "rd_flit_words", "rd_event_flits", "rd_event_writeback", "rd_doppler_run", "rd_read_coeff", "rd_read_pfb_vec", "rd_read_iq_ascii", "rd_write_iq_ascii", "rd_cov_integrate",
    "rd_chain_enqueue_static", "rd_bf_plain_run", "rd_cov_plain_run", "rd_chain_enqueue_plain", "rd_cfar_enqueue",
    "rd_integ_reset", "rd_integ_run", "rd_cov_integrate_windows", "rd_pfb_acc_run", "rd_cfar_decide",
    "rd_fxp_eval", "rd_fxp_accum", "rd_fxp_cmult",
]
FXP_OPS = {"sat": 0, "trunc": 1, "round_even": 2, "round": 2, "round_half_up": 3, "round_sat": 4, "add_sat": 5,
           "sub_sat": 6, "neg_sat": 7, "mul_q15": 8, "mul_q15_rs": 9, "cmul_re": 10, "cmul_im": 11, "acc_w": 12,
           "growth_bits": 13}
FXP_OP_DTYPE = np.dtype([("a", np.int64), ("b", np.int64), ("op", np.uint32), ("sh", np.uint32), ("w", np.uint32),
                         ("pad", np.uint32)])
RD_BATCH_AUTOTUNE = 1


class CfarCfg(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in
                ("enable", "mode", "out_mode", "guard_lead", "guard_lag", "ref_lead", "ref_lag", "alpha")]

    @classmethod
    def of(cls, words):
        return cls(*[int(w) for w in words])


def blocked_x(X, CH, F, N):
    """X [CH][F][N] (natural) -> [CH][N/4][F][4] (bin-quad blocked), flat."""
    return np.ascontiguousarray(np.asarray(X, np.uint32).reshape(CH, F, N // 4, 4).transpose(0, 2, 1, 3)).reshape(-1)


def natural_x(Xb, CH, F, N):
    """The inverse of blocked_x."""
    return np.ascontiguousarray(np.asarray(Xb, np.uint32).reshape(CH, N // 4, F, 4).transpose(0, 2, 1, 3)).reshape(-1)


class CpiDesc(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in
                ("n_chan", "n_fft", "taps", "n_frames", "n_beams", "spc", "scale_sched", "max_guard", "max_ref")] + [
        ("cfar", CfarCfg), ("x_layout", C.c_uint32)]


class CpiBufs(C.Structure):
    _fields_ = [("iq", C.c_void_p), ("taps", C.c_void_p), ("weights", C.c_void_p), ("u", C.c_void_p),
                ("pfb_flags", C.c_void_p), ("X", C.c_void_p), ("fft_flags", C.c_void_p), ("Xa", C.c_void_p),
                ("y", C.c_void_p), ("yflags", C.c_void_p), ("power", C.c_void_p), ("R", C.c_void_p),
                ("events", C.c_void_p), ("events_cap", C.c_uint64), ("row_count", C.c_void_p)]


class IntegCfg(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("window_len", "mode", "exp_k", "enable")]


class IntegStep(C.Structure):
    _fields_ = [("cfg", IntegCfg), ("valid", C.c_uint32), ("flush", C.c_uint32), ("sat_clear", C.c_uint32)]


class IntegWindow(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("step", "window_id", "sample_count", "flushed", "truncated")]


class IntegCtl(C.Structure):
    _fields_ = [("active", IntegCfg), ("count", C.c_uint32), ("window_id", C.c_uint32)]


INTEG_LANE_DTYPE = np.dtype([("acc", np.int64), ("flags", np.uint32), ("sat_count", np.uint32)])
CFAR_DECISION_DTYPE = np.dtype([("cut_power", np.uint64), ("sum_lead", np.uint64), ("sum_lag", np.uint64),
                                ("sel_sum", np.uint64), ("sel_n", np.uint32), ("window_ok", np.uint8),
                                ("evaluable", np.uint8), ("detected", np.uint8), ("suppressed", np.uint8),
                                ("take_lead", np.uint8), ("pad", np.uint8, 7)])


def integ_steps(steps7):
    """A ctypes array of IntegStep from rows (window_len, mode, exp_k, enable,
    valid, flush, sat_clear)."""
    rows = np.asarray(steps7, np.uint32).reshape(-1, 7)
    arr = (IntegStep * max(1, rows.shape[0]))()
    for i, r in enumerate(rows.tolist()):
        arr[i] = IntegStep(IntegCfg(*r[:4]), r[4], r[5], r[6])
    return arr


class RdError(RuntimeError):
    def __init__(self, fn, code, msg):
        super().__init__(f"{fn} failed: {code} ({msg})")
        self.code = code


class Lib:
    def __init__(self, path: Path = LIB_PATH):
        if not path.exists():
            raise FileNotFoundError(f"{path} is not built; run __graft_entry__.build() or make -C {HERE}")
        self.lib = C.CDLL(str(path))
        L = self.lib
        vp, u32, u64, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
        L.rd_device_count.restype = i32
        L.rd_status_str.restype = C.c_char_p
        L.rd_status_str.argtypes = [i32]
        L.rd_create.argtypes = [i32, C.POINTER(vp)]
        L.rd_destroy.argtypes = [vp]
        L.rd_sync.argtypes = [vp]
        L.rd_malloc.argtypes = [vp, C.c_size_t, C.POINTER(vp)]
        L.rd_free.argtypes = [vp, vp]
        L.rd_h2d.argtypes = [vp, vp, vp, C.c_size_t]
        L.rd_d2h.argtypes = [vp, vp, vp, C.c_size_t]
        L.rd_memset.argtypes = [vp, vp, i32, C.c_size_t]
        L.rd_stream.argtypes = [vp]
        L.rd_stream.restype = vp
        L.rd_twiddle_table.argtypes = [vp]
        L.rd_pfb_run.argtypes = [vp, vp, vp, u32, u32, u32, u32, vp, vp, vp]
        L.rd_fft_run.argtypes = [vp, vp, u32, u32, u32, i32, u32, vp, vp, vp]
        L.rd_align_run.argtypes = [vp, vp, u32, u32, u32, vp]
        L.rd_bf_run.argtypes = [vp, vp, vp, u32, u32, u64, vp, vp, vp, vp]
        L.rd_bf_stage.argtypes = [vp, vp, vp, u32, u32, u32, u32, vp, vp, vp, vp]
        L.rd_bf_cov_stage.argtypes = [vp, vp, vp, u32, u32, u32, u32, vp, vp, vp, vp, vp, u32]
        L.rd_fft_stage.argtypes = [vp, vp, u32, u32, u32, u32, u32, u32, vp, vp]
        L.rd_cov_run.argtypes = [vp, vp, u32, u32, u32, vp]
        L.rd_cfar_run.argtypes = [vp, vp, u32, u32, u32, C.POINTER(CfarCfg), u32, u32, vp, u32, vp, u64, vp,
                                  C.POINTER(u64)]
        L.rd_fxp_eval.argtypes = [vp, vp, u64, vp, vp, vp]
        L.rd_fxp_accum.argtypes = [vp, u32, vp, u32, vp, vp, vp, vp]
        L.rd_fxp_cmult.argtypes = [vp, vp, vp, u64, vp, vp, vp, vp]
        L.rd_integ_reset.argtypes = [C.POINTER(IntegCfg), C.POINTER(IntegCtl)]
        L.rd_integ_run.argtypes = [vp, u32, C.POINTER(IntegCtl), vp, u32, vp, u64, vp, vp, vp, u32, C.POINTER(u32)]
        L.rd_cov_integrate_windows.argtypes = [vp, vp, u32, u32, u32, u32, C.POINTER(IntegCtl), vp, vp, vp, vp, u32,
                                               C.POINTER(u32)]
        L.rd_pfb_acc_run.argtypes = [vp, vp, vp, u32, u32, u32, u32, vp, vp]
        L.rd_cfar_decide.argtypes = [vp, vp, u32, u32, C.POINTER(CfarCfg), vp]
        L.rd_cfar_enqueue.argtypes = [vp, vp, u32, u32, u32, C.POINTER(CfarCfg), u32, u32, vp, u32, vp, u64, vp, vp]
        L.rd_chain_run.argtypes = [vp, C.POINTER(CpiDesc), C.POINTER(CpiBufs), C.POINTER(u64)]
        L.rd_chain_enqueue.argtypes = [vp, C.POINTER(CpiDesc), C.POINTER(CpiBufs), vp]
        L.rd_chain_enqueue_static.argtypes = [vp, C.POINTER(CpiDesc), C.POINTER(CpiBufs), vp]
        L.rd_chain_enqueue_plain.argtypes = [vp, C.POINTER(CpiDesc), C.POINTER(CpiBufs), vp]
        L.rd_bf_plain_run.argtypes = [vp, vp, vp, u32, u32, u64, vp, vp, vp]
        L.rd_cov_plain_run.argtypes = [vp, vp, u32, u32, u32, vp]
        L.rd_chain_graph.argtypes = [vp, C.POINTER(CpiDesc), C.POINTER(CpiBufs), vp, C.POINTER(vp)]
        L.rd_graph_launch.argtypes = [vp]
        L.rd_graph_destroy.argtypes = [vp]
        L.rd_batch_create.argtypes = [vp, u32, C.POINTER(CpiDesc), C.POINTER(CpiBufs), u32, C.POINTER(vp)]
        L.rd_batch_launch.argtypes = [vp]
        L.rd_batch_events.argtypes = [vp, vp]
        L.rd_batch_plan.argtypes = [vp, C.c_char_p, C.c_size_t]
        L.rd_batch_destroy.argtypes = [vp]
        L.rd_event_flit_count.argtypes = [u32]
        L.rd_event_flit_count.restype = u32
        L.rd_flit_words.argtypes = [u32]
        L.rd_flit_words.restype = u32
        L.rd_event_flits.argtypes = [vp, vp, u64, vp, u32, u32, u32, u32, u32, vp]
        L.rd_event_writeback.argtypes = [vp, vp, u64, vp, u32, u32, u32, vp, u64, u64, vp]
        L.rd_doppler_run.argtypes = [vp, vp, u32, u32, u32, u32, u32, u32, u32, i32, vp, vp]
        L.rd_cov_integrate.argtypes = [vp, vp, u32, u32, u32, u32, u32, u32, vp, vp]
        sz, pu = C.c_size_t, C.POINTER(u32)
        L.rd_read_coeff.argtypes = [C.c_char_p, vp, sz, pu, pu, C.c_char_p, sz]
        L.rd_read_pfb_vec.argtypes = [C.c_char_p, vp, vp, vp, vp, sz, pu, pu, pu, C.c_char_p, sz]
        L.rd_read_iq_ascii.argtypes = [C.c_char_p, vp, sz, pu, pu, pu, pu, C.c_char_p, sz]
        L.rd_write_iq_ascii.argtypes = [C.c_char_p, vp, u32, u32, u32, u32, C.c_char_p]
        for name in EXPORTS:
            fn = getattr(L, name)
            if name not in ("rd_device_count", "rd_status_str", "rd_stream", "rd_event_flit_count", "rd_flit_words"):
                fn.restype = i32

    def device_count(self) -> int:
        return int(self.lib.rd_device_count())

    def check(self, fn: str, code: int, ok=(RD_OK,)):
        if code not in ok:
            raise RdError(fn, code, self.lib.rd_status_str(code).decode())
        return code

    def twiddle_table(self) -> np.ndarray:
        out = np.zeros(4096, np.uint32)
        self.check("rd_twiddle_table", self.lib.rd_twiddle_table(out.ctypes.data_as(C.c_void_p)))
        return out

    def context(self, device: int = 0) -> "Ctx":
        return Ctx(self, device)

    # ---- reference file formats (host side)
    def _sized(self, fn, name, path, make_bufs, dims):
        """Two calls: sizes first (RD_E_CAPACITY), then the data."""
        err = C.create_string_buffer(512)
        d = [C.c_uint32() for _ in range(dims)]
        rc = fn(str(path).encode(), *make_bufs(None), *[C.byref(x) for x in d], err, len(err))
        if rc == RD_E_CAPACITY:
            args, out = make_bufs([x.value for x in d])
            rc = fn(str(path).encode(), *args, *[C.byref(x) for x in d], err, len(err))
            if rc == RD_OK:
                return out, [x.value for x in d]
        raise RdError(name, rc, err.value.decode() or self.lib.rd_status_str(rc).decode())

    def read_coeff(self, path):
        """.coeff -> (packed phase-major taps, phases, taps)."""
        def bufs(dims):
            if dims is None:
                return (None, 0)
            a = np.zeros(dims[0] * dims[1], np.uint32)
            return (a.ctypes.data_as(C.c_void_p), a.size), a
        a, (P, T) = self._sized(self.lib.rd_read_coeff, "rd_read_coeff", path, bufs, 2)
        return a, P, T

    def read_pfb_vec(self, path):
        """.vec -> dict(x, y, f_re, f_im [beats*phases], phases, taps, beats)."""
        def bufs(dims):
            if dims is None:
                return (None, None, None, None, 0)
            n = dims[0] * dims[2]
            out = dict(x=np.zeros(n, np.uint32), y=np.zeros(n, np.uint32), f_re=np.zeros(n, np.uint8),
                       f_im=np.zeros(n, np.uint8))
            ptrs = tuple(out[k].ctypes.data_as(C.c_void_p) for k in ("x", "y", "f_re", "f_im")) + (dims[2],)
            return ptrs, out
        out, (P, T, nb) = self._sized(self.lib.rd_read_pfb_vec, "rd_read_pfb_vec", path, bufs, 3)
        out.update(phases=P, taps=T, beats=nb)
        return out

    def read_iq_ascii(self, path):
        """iq_ascii_v1 -> (iq [n_ant][frames*fft] packed, n_ant, fft_size, frames, spc)."""
        def bufs(dims):
            if dims is None:
                return (None, 0)
            a = np.zeros(dims[0] * dims[1] * dims[2], np.uint32)
            return (a.ctypes.data_as(C.c_void_p), a.size), a
        a, (A, N, H, P) = self._sized(self.lib.rd_read_iq_ascii, "rd_read_iq_ascii", path, bufs, 4)
        return a, A, N, H, P

    def write_iq_ascii(self, path, iq, n_ant, fft_size, frames, spc, scenario="scene"):
        iq = np.ascontiguousarray(iq, np.uint32)
        self.check("rd_write_iq_ascii", self.lib.rd_write_iq_ascii(str(path).encode(), iq.ctypes.data_as(C.c_void_p),
                                                                   n_ant, fft_size, frames, spc, scenario.encode()))


class DevBuf:
    """A device allocation owned by a context."""

    def __init__(self, ctx: "Ctx", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        ctx.lib.check("rd_malloc", ctx.L.rd_malloc(ctx.h, max(self.nbytes, 4), C.byref(p)))
        self.ptr = p.value

    def upload(self, arr: np.ndarray) -> "DevBuf":
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        self.ctx.lib.check("rd_h2d", self.ctx.L.rd_h2d(self.ctx.h, self.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes))
        return self

    def download(self, dtype, count=None) -> np.ndarray:
        dt = np.dtype(dtype)
        n = self.nbytes // dt.itemsize if count is None else int(count)
        out = np.empty(n, dt)
        if n:
            self.ctx.lib.check("rd_d2h", self.ctx.L.rd_d2h(self.ctx.h, out.ctypes.data_as(C.c_void_p), self.ptr,
                                                             out.nbytes))
        return out

    def zero(self) -> "DevBuf":
        self.ctx.lib.check("rd_memset", self.ctx.L.rd_memset(self.ctx.h, self.ptr, 0, self.nbytes))
        return self

    def free(self):
        if self.ptr and self.ctx.h:
            self.ctx.L.rd_free(self.ctx.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Ctx:
    def __init__(self, lib: Lib, device: int = 0):
        self.lib = lib
        self.L = lib.lib
        h = C.c_void_p()
        lib.check("rd_create", self.L.rd_create(device, C.byref(h)))
        self.h = h.value
        self.graphs = weakref.WeakSet()

    def close(self):
        """Destroys the context's graphs, then the context (buffers still held by
        DevBuf objects become inert)."""
        if self.h:
            for g in list(self.graphs):
                g.close()
            self.L.rd_sync(self.h)
            self.lib.check("rd_destroy", self.L.rd_destroy(self.h))
            self.h = None

    @property
    def stream(self) -> int:
        return int(self.L.rd_stream(self.h) or 0)

    def sync(self):
        self.lib.check("rd_sync", self.L.rd_sync(self.h))

    def buf(self, nbytes: int) -> DevBuf:
        return DevBuf(self, nbytes)

    def put(self, arr: np.ndarray) -> DevBuf:
        a = np.ascontiguousarray(arr)
        return DevBuf(self, a.nbytes).upload(a)

    # ---- per-stage host-array conveniences (upload, run, download) ----------
    def pfb(self, x, coeff, P, T, n_chan, n_beats, hist=None, flags=True):
        """PFB through rd_pfb_run -> (y, flag bytes or None when flags=False)."""
        S = n_chan * n_beats * P
        dx, dc = self.put(np.asarray(x, np.uint32)), self.put(np.asarray(coeff, np.uint32))
        dy, df = self.buf(4 * S), (self.buf(S) if flags else None)
        dh = self.put(np.asarray(hist, np.uint32)) if hist is not None else None
        self.lib.check("rd_pfb_run", self.L.rd_pfb_run(self.h, dx.ptr, dc.ptr, P, T, n_chan, n_beats,
                                                       dh.ptr if dh else None, dy.ptr, df.ptr if df else None))
        y, f = dy.download(np.uint32, S), (df.download(np.uint8, S) if df else None)
        if hist is not None:
            hist[...] = dh.download(np.uint32, hist.size)
        return y, f

    def fft(self, x, n, spc=2, sched=0xFFFFFFFF, reorder=1, flags=True):
        """rd_fft_run -> (out, bins, stage flags or None when flags=False)."""
        x = np.asarray(x, np.uint32)
        frames = x.size // n
        dx = self.put(x)
        do, db = self.buf(4 * x.size), self.buf(4 * x.size)
        dfl = self.buf(4 * max(frames, 1)) if flags else None
        self.lib.check("rd_fft_run", self.L.rd_fft_run(self.h, dx.ptr, n, spc, sched, reorder, frames, do.ptr, db.ptr,
                                                       dfl.ptr if dfl else None))
        return (do.download(np.uint32, x.size), db.download(np.uint32, x.size),
                dfl.download(np.uint32, frames) if dfl else None)

    def fft_stage(self, u, n, n_chan, n_frames, x_layout=0, spc=2, sched=0xFFFFFFFF, flags=True):
        """rd_fft_stage -> (X in natural order, stage flags or None); with
        x_layout 1 the device writes it bin-quad blocked and it is returned
        un-blocked."""
        u = np.asarray(u, np.uint32)
        du, dX = self.put(u), self.buf(4 * u.size)
        dfl = self.buf(4 * max(n_chan * n_frames, 1)) if flags else None
        self.lib.check("rd_fft_stage", self.L.rd_fft_stage(self.h, du.ptr, n, spc, sched, n_chan, n_frames, x_layout,
                                                           dX.ptr, dfl.ptr if dfl else None))
        X = dX.download(np.uint32, u.size)
        if x_layout == 1:
            X = natural_x(X, n_chan, n_frames, n)
        return X, (dfl.download(np.uint32, n_chan * n_frames) if dfl else None)

    def align(self, X, n_chan, n_frames, n_bins):
        dX = self.put(np.asarray(X, np.uint32))
        dA = self.buf(dX.nbytes)
        self.lib.check("rd_align_run", self.L.rd_align_run(self.h, dX.ptr, n_chan, n_frames, n_bins, dA.ptr))
        return dA.download(np.uint32, np.asarray(X).size)

    def bf(self, Xa, W, CH, B, want_acc=True):
        Xa = np.asarray(Xa, np.uint32)
        cols = Xa.size // CH
        dX, dW = self.put(Xa), self.put(np.asarray(W, np.uint32))
        dy, df, dp = self.buf(4 * B * cols), self.buf(B * cols), self.buf(4 * B * cols)
        da = self.buf(16 * B * cols) if want_acc else None
        self.lib.check("rd_bf_run", self.L.rd_bf_run(self.h, dX.ptr, dW.ptr, CH, B, cols, dy.ptr, df.ptr, dp.ptr,
                                                     da.ptr if da else None))
        return (dy.download(np.uint32, B * cols), df.download(np.uint8, B * cols), dp.download(np.uint32, B * cols),
                da.download(np.int64, 2 * B * cols) if da else None)

    def bf_stage(self, X, W, CH, B, F, N):
        """The chain's beamformer on the FFT layout; returns y, flags, power, Xa."""
        X = np.asarray(X, np.uint32)
        cols = F * N
        dX, dW = self.put(X), self.put(np.asarray(W, np.uint32))
        dy, df, dp, da = self.buf(4 * B * cols), self.buf(B * cols), self.buf(4 * B * cols), self.buf(4 * CH * cols)
        self.lib.check("rd_bf_stage", self.L.rd_bf_stage(self.h, dX.ptr, dW.ptr, CH, B, F, N, dy.ptr, df.ptr, dp.ptr,
                                                         da.ptr))
        return (dy.download(np.uint32, B * cols), df.download(np.uint8, B * cols), dp.download(np.uint32, B * cols),
                da.download(np.uint32, CH * cols))

    def bf_cov_stage(self, X, W, CH, B, F, N, outputs=("y", "flags", "power", "Xa"), x_layout=0):
        """rd_bf_cov_stage on the FFT layout (X given in natural order; with
        x_layout 1 it is passed bin-quad blocked); returns (y, flags, power,
        Xa, R), None for an output not in `outputs` (passed as NULL)."""
        X = np.asarray(X, np.uint32)
        if x_layout == 1:
            X = blocked_x(X, CH, F, N)
        cols = F * N
        dX, dW = self.put(X), self.put(np.asarray(W, np.uint32))
        dy = self.buf(4 * B * cols) if "y" in outputs else None
        df = self.buf(B * cols) if "flags" in outputs else None
        dp = self.buf(4 * B * cols) if "power" in outputs else None
        da = self.buf(4 * CH * cols) if "Xa" in outputs else None
        dR = self.buf(16 * N * CH * (CH + 1) // 2)
        ptr = lambda d: d.ptr if d is not None else None  # noqa: E731
        self.lib.check("rd_bf_cov_stage", self.L.rd_bf_cov_stage(self.h, dX.ptr, dW.ptr, CH, B, F, N, ptr(dy),
                                                                 ptr(df), ptr(dp), ptr(da), dR.ptr, x_layout))
        return (dy.download(np.uint32, B * cols) if dy is not None else None,
                df.download(np.uint8, B * cols) if df is not None else None,
                dp.download(np.uint32, B * cols) if dp is not None else None,
                da.download(np.uint32, CH * cols) if da is not None else None,
                dR.download(np.int64, N * CH * (CH + 1)))

    def bf_plain(self, Xa, W, CH, B):
        """rd_bf_plain_run -> (y, flags, power)."""
        Xa = np.asarray(Xa, np.uint32)
        cols = Xa.size // CH
        dX, dW = self.put(Xa), self.put(np.asarray(W, np.uint32))
        dy, df, dp = self.buf(4 * B * cols), self.buf(B * cols), self.buf(4 * B * cols)
        self.lib.check("rd_bf_plain_run", self.L.rd_bf_plain_run(self.h, dX.ptr, dW.ptr, CH, B, cols, dy.ptr, df.ptr,
                                                                 dp.ptr))
        return dy.download(np.uint32, B * cols), df.download(np.uint8, B * cols), dp.download(np.uint32, B * cols)

    def cov_plain(self, Xa, CH, F, N):
        dX = self.put(np.asarray(Xa, np.uint32))
        npairs = CH * (CH + 1) // 2
        dR = self.buf(16 * N * npairs)
        self.lib.check("rd_cov_plain_run", self.L.rd_cov_plain_run(self.h, dX.ptr, CH, F, N, dR.ptr))
        return dR.download(np.int64, 2 * N * npairs)

    def cov(self, Xa, CH, F, N):
        dX = self.put(np.asarray(Xa, np.uint32))
        npairs = CH * (CH + 1) // 2
        dR = self.buf(16 * N * npairs)
        self.lib.check("rd_cov_run", self.L.rd_cov_run(self.h, dX.ptr, CH, F, N, dR.ptr))
        return dR.download(np.int64, 2 * N * npairs)

    def cfar(self, cells, n, cfg8, max_guard, max_ref, frame_ids=None, frame_mod=0, cap=None):
        cells = np.ascontiguousarray(cells)
        assert cells.dtype in (np.uint32, np.uint64)
        rows = cells.size // n
        cap = rows * (n + 1) if cap is None else cap
        dc = self.put(cells)
        dfid = self.put(np.asarray(frame_ids, np.uint32)) if frame_ids is not None else None
        dev, drc = self.buf(24 * max(cap, 1)), self.buf(4 * max(rows, 1))
        tot = C.c_uint64()
        cfg = CfarCfg.of(cfg8)
        rc = self.L.rd_cfar_run(self.h, dc.ptr, cells.dtype.itemsize, n, rows, C.byref(cfg), max_guard, max_ref,
                                dfid.ptr if dfid else None, frame_mod, dev.ptr, cap, drc.ptr, C.byref(tot))
        self.lib.check("rd_cfar_run", rc, ok=(RD_OK, RD_E_CAPACITY))
        t = min(tot.value, cap)
        return dev.download(np.uint64, 3 * t).reshape(t, 3), drc.download(np.uint32, rows), tot.value, rc


def _ctx_event_flits(self, events, packet_w, dest=0, src=0, vc=0, seq0=0, n_dev=None, cap=None):
    """Detection packets of packed events [n][3] (host array or DevBuf with
    cap events) -> flits [n][L][fw] u32."""
    L = int(self.L.rd_event_flit_count(packet_w))
    fw = int(self.L.rd_flit_words(packet_w))
    if isinstance(events, DevBuf):
        dev, n = events, int(cap)
    else:
        ev = np.ascontiguousarray(events, np.uint64).reshape(-1, 3)
        n = ev.shape[0]
        dev = self.put(ev)
    out = self.buf(4 * max(n, 1) * L * fw)
    rc = self.L.rd_event_flits(self.h, dev.ptr, n, n_dev.ptr if n_dev is not None else None, packet_w, dest, src,
                               vc, seq0, out.ptr)
    self.lib.check("rd_event_flits", rc)
    return out.download(np.uint32, n * L * fw).reshape(n, L, fw)


Ctx.event_flits = _ctx_event_flits

RD_WB_NSTATS = 5


def _ctx_event_writeback(self, flits, packet_w, port=0, seq0=0, ring_words=None, addr0=0, ring=None, stats=None,
                         n_dev=None, cap=None):
    """Detection writeback of flits [n][L][fw] (host array, or DevBuf with cap
    packets) into a ring of 64-byte words. ring / stats: DevBufs to keep across
    calls (created, zeroed, when None). Returns (ring [ring_words][16] u32,
    stats [5] u64) as host arrays."""
    if isinstance(flits, DevBuf):
        dev, n = flits, int(cap)
    else:
        fl = np.ascontiguousarray(flits, np.uint32)
        n = fl.shape[0]
        dev = self.put(fl)
    if ring_words is None:
        ring_words = max(n, 1)
    if ring is None:
        ring = self.buf(64 * ring_words).zero()
    if stats is None:
        stats = self.buf(8 * RD_WB_NSTATS).zero()
    rc = self.L.rd_event_writeback(self.h, dev.ptr, n, n_dev.ptr if n_dev is not None else None, packet_w, port,
                                   seq0, ring.ptr, ring_words, addr0, stats.ptr)
    self.lib.check("rd_event_writeback", rc)
    return ring.download(np.uint32, 16 * ring_words).reshape(ring_words, 16), stats.download(np.uint64, RD_WB_NSTATS)


Ctx.event_writeback = _ctx_event_writeback


def _ctx_doppler(self, ring, rows, depth, n_bins, newest, n_dop, spc=2, sched=0xFFFFFFFF, shift=False):
    """Slow-time transform of a ring [rows][depth][n_bins] -> (out
    [rows][n_dop][n_bins], stage_flags [rows][n_bins])."""
    dev = self.put(np.ascontiguousarray(ring, np.uint32))
    out = self.buf(4 * rows * n_dop * n_bins)
    sf = self.buf(4 * rows * n_bins)
    rc = self.L.rd_doppler_run(self.h, dev.ptr, rows, depth, n_bins, newest, n_dop, spc, sched, 1 if shift else 0,
                               out.ptr, sf.ptr)
    self.lib.check("rd_doppler_run", rc)
    return out.download(np.uint32, rows * n_dop * n_bins), sf.download(np.uint32, rows * n_bins)


Ctx.doppler = _ctx_doppler


def _ctx_cov_integrate(self, Xa, CH, F, N, mode, exp_k, acc_w):
    """Covariance through the Integrator -> (R int64 [N][pairs][2], sat flags [N][pairs])."""
    dX = self.put(np.ascontiguousarray(Xa, np.uint32))
    npairs = CH * (CH + 1) // 2
    dR = self.buf(16 * N * npairs)
    df = self.buf(N * npairs)
    self.lib.check("rd_cov_integrate", self.L.rd_cov_integrate(self.h, dX.ptr, CH, F, N, mode, exp_k, acc_w, dR.ptr,
                                                               df.ptr))
    return dR.download(np.int64, 2 * N * npairs), df.download(np.uint8, N * npairs)


Ctx.cov_integrate = _ctx_cov_integrate


def _ctx_integ_run(self, acc_w, ctl, steps7, data, lanes, cap=64):
    """rd_integ_run: data [n_steps][n_lanes] int64 (host), lanes a DevBuf of
    INTEG_LANE_DTYPE records carried across calls, ctl an IntegCtl (advanced in
    place). Returns (results [n_windows][n_lanes], windows [n_windows][5],
    n_windows)."""
    data = np.ascontiguousarray(data, np.int64)
    n_steps = data.shape[0]
    n_lanes = data.shape[1] if data.ndim == 2 else 0
    steps = integ_steps(steps7)
    dd = self.put(data) if data.size else None
    res = self.buf(8 * max(cap, 1) * max(n_lanes, 1))
    wins = (IntegWindow * max(cap, 1))()
    nw = C.c_uint32()
    rc = self.L.rd_integ_run(self.h, acc_w, C.byref(ctl), steps, n_steps, dd.ptr if dd else None, n_lanes,
                             lanes.ptr, res.ptr, wins, cap, C.byref(nw))
    self.lib.check("rd_integ_run", rc, ok=(RD_OK, RD_E_CAPACITY))
    k = min(nw.value, cap)
    w = np.array([[x.step, x.window_id, x.sample_count, x.flushed, x.truncated] for x in wins[:k]],
                 np.uint32).reshape(k, 5)
    return res.download(np.int64, k * n_lanes).reshape(k, n_lanes), w, nw.value


Ctx.integ_run = _ctx_integ_run


def _ctx_cov_integrate_windows(self, Xa, CH, F, N, acc_w, ctl, steps7, lanes, cap=64):
    """rd_cov_integrate_windows on host Xa; lanes a DevBuf of N*pairs*2
    INTEG_LANE_DTYPE records. Returns (results [n_windows][N*pairs*2],
    windows [n_windows][5], n_windows)."""
    dX = self.put(np.ascontiguousarray(Xa, np.uint32))
    n_lanes = N * CH * (CH + 1)
    steps = integ_steps(steps7)
    res = self.buf(8 * max(cap, 1) * n_lanes)
    wins = (IntegWindow * max(cap, 1))()
    nw = C.c_uint32()
    rc = self.L.rd_cov_integrate_windows(self.h, dX.ptr, CH, F, N, acc_w, C.byref(ctl), steps, lanes.ptr, res.ptr,
                                         wins, cap, C.byref(nw))
    self.lib.check("rd_cov_integrate_windows", rc, ok=(RD_OK, RD_E_CAPACITY))
    k = min(nw.value, cap)
    w = np.array([[x.step, x.window_id, x.sample_count, x.flushed, x.truncated] for x in wins[:k]],
                 np.uint32).reshape(k, 5)
    return res.download(np.int64, k * n_lanes).reshape(k, n_lanes), w, nw.value


Ctx.cov_integrate_windows = _ctx_cov_integrate_windows


def _ctx_pfb_acc(self, x, coeff, P, T, n_chan, n_beats, hist=None):
    """rd_pfb_acc_run -> acc [n_chan][n_beats][P][2] int64."""
    dx, dc = self.put(np.asarray(x, np.uint32)), self.put(np.asarray(coeff, np.uint32))
    dh = self.put(np.asarray(hist, np.uint32)) if hist is not None else None
    n = n_chan * n_beats * P
    da = self.buf(16 * max(n, 1))
    self.lib.check("rd_pfb_acc_run", self.L.rd_pfb_acc_run(self.h, dx.ptr, dc.ptr, P, T, n_chan, n_beats,
                                                           dh.ptr if dh else None, da.ptr))
    return da.download(np.int64, 2 * n)


Ctx.pfb_acc = _ctx_pfb_acc


def _ctx_cfar_decide(self, cells, n, cfg8):
    """rd_cfar_decide on uint64 cells [rows][n] -> structured array of
    CFAR_DECISION_DTYPE records [rows*n]."""
    cells = np.ascontiguousarray(cells, np.uint64)
    rows = cells.size // n
    dc = self.put(cells)
    do = self.buf(48 * max(rows * n, 1))
    cfg = CfarCfg.of(cfg8)
    self.lib.check("rd_cfar_decide", self.L.rd_cfar_decide(self.h, dc.ptr, n, rows, C.byref(cfg), do.ptr))
    raw = do.download(np.uint8, 48 * rows * n)
    return raw.view(CFAR_DECISION_DTYPE)


Ctx.cfar_decide = _ctx_cfar_decide


def _ctx_fxp_eval(self, ops):
    """rd_fxp_eval on records (FXP_OP_DTYPE) -> (y int64, flags u32, forms u32)."""
    ops = np.ascontiguousarray(ops, FXP_OP_DTYPE)
    n = ops.size
    d = self.put(ops.view(np.uint8))
    dy, df, dm = self.buf(8 * max(n, 1)), self.buf(4 * max(n, 1)), self.buf(4 * max(n, 1))
    self.lib.check("rd_fxp_eval", self.L.rd_fxp_eval(self.h, d.ptr, n, dy.ptr, df.ptr, dm.ptr))
    return dy.download(np.int64, n), df.download(np.uint32, n), dm.download(np.uint32, n)


def _ctx_fxp_accum(self, acc_w, xs):
    """rd_fxp_accum from a clear -> (y, step_flags, sticky, count) per step."""
    x = np.ascontiguousarray(xs, np.int64)
    n = x.size
    dx = self.put(x)
    b = [self.buf(8 * max(n, 1))] + [self.buf(4 * max(n, 1)) for _ in range(3)]
    self.lib.check("rd_fxp_accum", self.L.rd_fxp_accum(self.h, acc_w, dx.ptr, n, *[t.ptr for t in b]))
    return (b[0].download(np.int64, n),) + tuple(t.download(np.uint32, n) for t in b[1:])


def _ctx_fxp_cmult(self, a, b):
    """rd_fxp_cmult -> (p4 [n][2], p3 [n][2], y packed, flags f_re | f_im << 2)."""
    a, b = np.ascontiguousarray(a, np.uint32), np.ascontiguousarray(b, np.uint32)
    n = a.size
    da, db = self.put(a), self.put(b)
    d4, d3, dy, df = self.buf(16 * max(n, 1)), self.buf(16 * max(n, 1)), self.buf(4 * max(n, 1)), self.buf(max(n, 1))
    self.lib.check("rd_fxp_cmult", self.L.rd_fxp_cmult(self.h, da.ptr, db.ptr, n, d4.ptr, d3.ptr, dy.ptr, df.ptr))
    return (d4.download(np.int64, 2 * n).reshape(n, 2), d3.download(np.int64, 2 * n).reshape(n, 2),
            dy.download(np.uint32, n), df.download(np.uint8, n))


Ctx.fxp_eval = _ctx_fxp_eval
Ctx.fxp_accum = _ctx_fxp_accum
Ctx.fxp_cmult = _ctx_fxp_cmult


class Chain:
    """Device buffers and descriptor for one CPI geometry (reused across runs)."""

    def __init__(self, ctx: Ctx, CH, N, T, F, B, cfar_cfg8, max_guard=31, max_ref=63, spc=2, sched=0xFFFFFFFF,
                 keep_y=True, keep_flags=True, keep_R=True, events_cap=None, x_layout=0):
        """x_layout 1: the chain keeps X bin-quad blocked [CH][N/4][F][4] (the
        fused path's fast layout); fetch() returns it in natural order."""
        self.ctx = ctx
        S = CH * N * F
        self.shape = dict(CH=CH, N=N, T=T, F=F, B=B)
        self.desc = CpiDesc(CH, N, T, F, B, spc, sched, max_guard, max_ref, CfarCfg.of(cfar_cfg8), x_layout)
        self.events_cap = B * F * (N + 1) if events_cap is None else events_cap
        b = {}
        b["iq"] = ctx.buf(4 * S)
        b["taps"] = ctx.buf(4 * N * T)
        b["weights"] = ctx.buf(4 * B * CH)
        b["u"] = ctx.buf(4 * S)
        b["pfb_flags"] = ctx.buf(S) if keep_flags else None
        b["X"] = ctx.buf(4 * S)
        b["fft_flags"] = ctx.buf(4 * CH * F) if keep_flags else None
        b["Xa"] = ctx.buf(4 * S)
        b["y"] = ctx.buf(4 * B * N * F) if keep_y else None
        b["yflags"] = ctx.buf(B * N * F) if keep_flags else None
        b["power"] = ctx.buf(4 * B * N * F)
        b["R"] = ctx.buf(16 * N * CH * (CH + 1) // 2) if keep_R else None
        b["events"] = ctx.buf(24 * max(self.events_cap, 1))
        b["row_count"] = ctx.buf(4 * B * F)
        self.total = ctx.buf(8)
        self.b = b
        self.bufs = CpiBufs(**{k: (v.ptr if v is not None else None) for k, v in b.items()},
                            events_cap=self.events_cap)

    def load(self, iq, taps, weights):
        self.b["iq"].upload(np.asarray(iq, np.uint32))
        self.b["taps"].upload(np.asarray(taps, np.uint32))
        self.b["weights"].upload(np.asarray(weights, np.uint32))

    def run(self) -> int:
        tot = C.c_uint64()
        rc = self.ctx.L.rd_chain_run(self.ctx.h, C.byref(self.desc), C.byref(self.bufs), C.byref(tot))
        self.ctx.lib.check("rd_chain_run", rc)
        return tot.value

    def enqueue(self, iq_ptr=None, total_ptr=None, static=False, plain=False, lean=False):
        """rd_chain_enqueue; with static=True the allocation-free
        rd_chain_enqueue_static (the CFAR scratch in the library's static
        array); with plain=True rd_chain_enqueue_plain (no matrix cores).
        lean: the optional per-stage outputs are not written (as graph())."""
        bufs = self.bufs
        if iq_ptr is not None or lean:
            bufs = CpiBufs.from_buffer_copy(self.bufs)
        if iq_ptr is not None:
            bufs.iq = iq_ptr
        if lean:
            bufs.pfb_flags = bufs.fft_flags = bufs.y = bufs.yflags = bufs.Xa = None
        fn = "rd_chain_enqueue_static" if static else ("rd_chain_enqueue_plain" if plain else "rd_chain_enqueue")
        rc = getattr(self.ctx.L, fn)(self.ctx.h, C.byref(self.desc), C.byref(bufs),
                                     total_ptr if total_ptr is not None else self.total.ptr)
        self.ctx.lib.check(fn, rc)

    def graph(self, iq_ptr=None, total_ptr=None, lean=False) -> "Graph":
        """The whole chain for this CPI geometry captured as one HIP graph.
        lean: the optional per-stage outputs (PFB/FFT/beamformer flags, the
        beam samples y and the aligned copy Xa) are not written; every stage
        still runs and the intermediates the next stage reads, the covariance
        and the events are produced."""
        bufs = CpiBufs.from_buffer_copy(self.bufs)
        if iq_ptr is not None:
            bufs.iq = iq_ptr
        if lean:
            bufs.pfb_flags = bufs.fft_flags = bufs.y = bufs.yflags = bufs.Xa = None
        g = C.c_void_p()
        rc = self.ctx.L.rd_chain_graph(self.ctx.h, C.byref(self.desc), C.byref(bufs),
                                       total_ptr if total_ptr is not None else self.total.ptr, C.byref(g))
        self.ctx.lib.check("rd_chain_graph", rc)
        return Graph(self.ctx, g.value, keep=self)

    def fetch(self, names=("u", "pfb_flags", "X", "fft_flags", "Xa", "y", "yflags", "power", "R"), n_events=None):
        CH, N, F, B = (self.shape[k] for k in ("CH", "N", "F", "B"))
        S = CH * N * F
        spec = dict(u=(np.uint32, S), pfb_flags=(np.uint8, S), X=(np.uint32, S), fft_flags=(np.uint32, CH * F),
                    Xa=(np.uint32, S), y=(np.uint32, B * N * F), yflags=(np.uint8, B * N * F),
                    power=(np.uint32, B * N * F), R=(np.int64, N * CH * (CH + 1)), row_count=(np.uint32, B * F))
        out = {}
        for k in names:
            if self.b.get(k) is not None:
                dt, n = spec[k]
                out[k] = self.b[k].download(dt, n)
        if "X" in out and self.desc.x_layout == 1:
            out["X"] = natural_x(out["X"], CH, F, N)
        out["row_count"] = self.b["row_count"].download(np.uint32, B * F)
        if n_events is not None:
            t = min(n_events, self.events_cap)
            out["events"] = self.b["events"].download(np.uint64, 3 * t).reshape(t, 3)
        return out


class Graph:
    """A captured chain. `keep` (the Chain) stays referenced while the graph
    lives: the graph replays into its device buffers."""

    def __init__(self, ctx: Ctx, h, keep=None):
        self.ctx, self.h, self._keep = ctx, h, keep
        ctx.graphs.add(self)

    def launch(self):
        self.ctx.lib.check("rd_graph_launch", self.ctx.L.rd_graph_launch(self.h))

    def close(self):
        if self.h and self.ctx.h:
            self.ctx.L.rd_graph_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many CPIs through the chain with one graph launch (rd_batch_*):
    `descs` / `bufs` are lists of :class:`CpiDesc` / :class:`CpiBufs`."""

    def __init__(self, ctx: Ctx, descs, bufs, autotune=False, keep=None):
        self.ctx = ctx
        self.keep = keep  # objects owning the buffers
        n = len(descs)
        da = (CpiDesc * n)(*descs)
        ba = (CpiBufs * n)(*bufs)
        h = C.c_void_p()
        rc = ctx.L.rd_batch_create(ctx.h, n, da, ba, RD_BATCH_AUTOTUNE if autotune else 0, C.byref(h))
        ctx.lib.check("rd_batch_create", rc)
        self.h = h.value
        self.n = n
        ctx.graphs.add(self)

    @classmethod
    def of_chains(cls, ctx: Ctx, chains, autotune=False, lean=False, iq_ptrs=None) -> "Batch":
        """A batch over :class:`Chain` objects (each with its own buffers). lean:
        the optional per-stage outputs are not written (as :meth:`Chain.graph`)."""
        chains = list(chains)
        bufs = []
        for i, c in enumerate(chains):
            b = CpiBufs.from_buffer_copy(c.bufs)
            if iq_ptrs is not None:
                b.iq = iq_ptrs[i]
            if lean:
                b.pfb_flags = b.fft_flags = b.y = b.yflags = None
            bufs.append(b)
        return cls(ctx, [c.desc for c in chains], bufs, autotune=autotune, keep=chains)

    def launch(self):
        self.ctx.lib.check("rd_batch_launch", self.ctx.L.rd_batch_launch(self.h))

    def events(self) -> np.ndarray:
        out = np.zeros(self.n, np.uint64)
        rc = self.ctx.L.rd_batch_events(self.h, out.ctypes.data_as(C.c_void_p))
        self.ctx.lib.check("rd_batch_events", rc, ok=(RD_OK, RD_E_CAPACITY))
        return out

    def plan(self) -> str:
        buf = C.create_string_buffer(1 << 16)
        self.ctx.lib.check("rd_batch_plan", self.ctx.L.rd_batch_plan(self.h, buf, len(buf)))
        return buf.value.decode()

    def close(self):
        if self.h and self.ctx.h:
            self.ctx.L.rd_batch_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_lib = None


def load() -> Lib:
    global _lib
    if _lib is None:
        _lib = Lib()
    return _lib
