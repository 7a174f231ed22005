"""Device-side workloads shared by the GPU tests and bench.py: seeded scenes
generated on the GPU (the construction of rd_scenario.Scene.iq in torch), the
uneven CPI batch with its buffers, and the per-CPI checks of a batch against
one-CPI launches and against the oracle."""
from __future__ import annotations

import os

import numpy as np

import rd_amd
import rd_scenario


def note(msg: str) -> None:
    import sys

    print(f"[workload] {msg}", file=sys.stderr, flush=True)


def device_iq(torch, sc, seed, chan_err=None):
    """A scene's I/Q generated on the device (same construction as
    rd_scenario.Scene.iq: complex Gaussian noise plus plane-wave tones on the
    array, rounded half-to-even and saturated to int16, packed {im, re}).
    chan_err (complex [CH], optional) multiplies each channel before
    quantisation: receiver gain/phase errors."""
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    CH, L = sc.CH, sc.N * sc.F
    re = torch.randn((CH, L), generator=g, device="cuda") * sc.sigma
    im = torch.randn((CH, L), generator=g, device="cuda") * sc.sigma
    if sc.targets:
        rng = np.random.default_rng(seed)
        a = np.arange(CH)
        steer = np.empty((CH, len(sc.targets)), np.complex64)
        freqs = np.empty(len(sc.targets), np.float64)
        for i, (k, ang, snr) in enumerate(sc.targets):
            amp = sc.sigma * 10 ** (snr / 20.0)
            steer[:, i] = amp * np.exp(1j * (np.pi * a * np.sin(np.deg2rad(ang)) + rng.uniform(0, 2 * np.pi)))
            freqs[i] = k / sc.N
        t = torch.arange(L, device="cuda", dtype=torch.float64)
        ph = torch.outer(torch.from_numpy(2 * np.pi * freqs).cuda(), t)  # exact phase in f64, then f32 tones
        tones = torch.polar(torch.ones_like(ph, dtype=torch.float32), ph.float())
        sig = torch.from_numpy(steer).cuda() @ tones
        re += sig.real
        im += sig.imag
        del ph, tones, sig
    if chan_err is not None:
        e = torch.from_numpy(np.asarray(chan_err, np.complex64)).cuda()[:, None]
        re, im = re * e.real - im * e.imag, re * e.imag + im * e.real
    rq = torch.clamp(torch.round(re), -32768, 32767).to(torch.int32)
    iq_ = torch.clamp(torch.round(im), -32768, 32767).to(torch.int32)
    return ((iq_ << 16) | (rq & 0xFFFF)).reshape(-1)


class BatchSet:
    """Device buffers (torch allocations) and descriptors of an uneven batch."""

    def __init__(self, torch, scenes, seed0, events_per_row=64):
        self.scenes = scenes
        self.t = []
        self.descs, self.bufs, self.caps = [], [], []
        taps_of, w_of = {}, {}
        dev = "cuda"
        for i, sc in enumerate(scenes):
            S, P = sc.samples, sc.B * sc.F * sc.N
            if sc.N not in taps_of:
                taps_of[sc.N] = torch.from_numpy(sc.taps().view(np.int32)).to(dev)
            key = (sc.CH, sc.B)
            if key not in w_of:
                w_of[key] = torch.from_numpy(sc.weights().view(np.int32)).to(dev)
            cap = sc.B * sc.F * events_per_row
            d = dict(iq=device_iq(torch, sc, seed0 + i), taps=taps_of[sc.N], weights=w_of[key],
                     u=torch.empty(S, dtype=torch.int32, device=dev), X=torch.empty(S, dtype=torch.int32, device=dev),
                     Xa=torch.empty(S, dtype=torch.int32, device=dev),
                     power=torch.empty(P, dtype=torch.int32, device=dev),
                     R=torch.empty(sc.N * sc.CH * (sc.CH + 1), dtype=torch.int64, device=dev),
                     events=torch.empty(3 * cap, dtype=torch.int64, device=dev),
                     row_count=torch.empty(sc.B * sc.F, dtype=torch.int32, device=dev))
            self.t.append(d)
            self.caps.append(cap)
            self.descs.append(rd_amd.CpiDesc(sc.CH, sc.N, sc.T, sc.F, sc.B, sc.spc, sc.sched, rd_scenario.MAX_GUARD,
                                             rd_scenario.MAX_REF, rd_amd.CfarCfg.of(sc.cfar_cfg())))
            self.bufs.append(rd_amd.CpiBufs(**{k: v.data_ptr() for k, v in d.items()}, events_cap=cap))
            if (i + 1) % 64 == 0:
                note(f"  batch: {i + 1}/{len(scenes)} CPIs generated")
        torch.cuda.synchronize()

    @property
    def samples(self) -> int:
        return sum(sc.samples for sc in self.scenes)

    def fit_event_caps(self, torch, counts):
        """Grow the event buffers of CPIs whose first pass counted more events
        than fit (the count is exact even when the buffer is short)."""
        for i, n in enumerate(counts.tolist()):
            if n > self.caps[i]:
                cap = int(n * 1.25) + 64
                self.t[i]["events"] = torch.empty(3 * cap, dtype=torch.int64, device="cuda")
                self.caps[i] = cap
                self.bufs[i].events = self.t[i]["events"].data_ptr()
                self.bufs[i].events_cap = cap
        torch.cuda.synchronize()


def verify_batch_vs_plain(torch, ctx, bs, batch, totals):
    """Each CPI's batched outputs (already in its buffers) against a fresh
    one-CPI launch: outputs are copied aside, the CPI is rerun alone, compared."""
    bad = 0
    tot = np.zeros(1, np.uint64)
    names = ("u", "X", "Xa", "power", "R", "row_count")
    for i, (d, b) in enumerate(zip(bs.descs, bs.bufs)):
        t = bs.t[i]
        n = int(totals[i])
        saved = {k: t[k].clone() for k in names}
        saved["events"] = t["events"][: 3 * n].clone()
        torch.cuda.synchronize()
        for k in names + ("events",):
            t[k].zero_()
        torch.cuda.synchronize()
        ctx.lib.check("rd_chain_run", ctx.L.rd_chain_run(ctx.h, rd_amd.C.byref(d), rd_amd.C.byref(b),
                                                         tot.ctypes.data_as(rd_amd.C.POINTER(rd_amd.C.c_uint64))))
        ok = int(tot[0]) == n and all(torch.equal(saved[k], t[k]) for k in names)
        ok = ok and torch.equal(saved["events"], t["events"][: 3 * n])
        bad += 0 if ok else 1
    return bad


def verify_batch_vs_oracle(torch, ctx, bs, totals, k):
    """The k smallest CPIs of the batch against the CPU chain (power, covariance, events)."""
    if k <= 0:
        return None, 0
    import oracle

    orc = oracle.load("oracle")
    order = sorted(range(len(bs.scenes)), key=lambda i: bs.scenes[i].samples)[:k]
    bad = 0
    for i in order:
        sc, t = bs.scenes[i], bs.t[i]
        iq = t["iq"].cpu().numpy().view(np.uint32)
        o = orc.chain(iq, sc.taps(), sc.weights(), sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg(), rd_scenario.MAX_GUARD,
                      rd_scenario.MAX_REF, keep=("power", "R", "events"), threads=min(16, os.cpu_count() or 1))
        n = int(totals[i])
        bad += int(np.count_nonzero(t["power"].cpu().numpy().view(np.uint32) != o["power"]))
        bad += int(np.count_nonzero(t["R"].cpu().numpy() != o["R"]))
        if n != o["n_events"]:
            bad += abs(n - o["n_events"]) + 1
        else:
            bad += int(np.count_nonzero(t["events"][: 3 * n].cpu().numpy().view(np.uint64).reshape(n, 3) != o["events"]))
    return bad, len(order)


def covariance_trace_check(torch, Xa, R, CH, F, N):
    """The covariance-trace invariant checked in FP64 on the GPU: for every bin
    k, tr R[k] = sum_c sum_m power(X_c[m][k]) and every diagonal entry is real
    (covar::cross_power(x, x) = power(x) + 0j, covar_model.hpp:99-102, 137-151).
    Every term is an integer below 2^31 and every sum below 2^53, so the float64
    sums are exact and the check is equality. Xa [F][N][CH] packed, R as the
    chain writes it ([N][pairs][2] int64)."""
    x = torch.from_numpy(np.ascontiguousarray(Xa).view(np.int32)).cuda().reshape(F, N, CH)
    re = ((x << 16) >> 16).double()
    im = (x >> 16).double()
    p = (re * re + im * im).sum(dim=(0, 2))
    del x, re, im
    npairs = CH * (CH + 1) // 2
    r = torch.from_numpy(np.ascontiguousarray(R)).cuda().reshape(N, npairs, 2)
    diag = torch.tensor([i * CH - i * (i - 1) // 2 for i in range(CH)], device="cuda")
    tr = r[:, diag, 0].double().sum(1)
    err = float((tr - p).abs().max())
    im_nz = int((r[:, diag, 1] != 0).sum())
    return dict(bins=N, frames=F, channels=CH, max_abs_err=err, diag_imag_nonzero=im_nz,
                ok=bool(err == 0.0 and im_nz == 0))
