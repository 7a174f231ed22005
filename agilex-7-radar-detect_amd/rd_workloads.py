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


def hard_cfar_hits(torch, power, N, guard, ref, alpha):
    """The bit-exact CA-CFAR decision (cfar_model.hpp:296-301) on int power
    rows [..., N] in exact int64 torch ops: hit iff (C * n) << 8 > alpha * S,
    for the bins whose symmetric window fits (the others are suppressed)."""
    P = power.to(torch.int64)
    lo, hi = guard + ref, N - guard - ref
    S = torch.zeros_like(P[..., lo:hi])
    for d in range(guard + 1, guard + ref + 1):
        S += P[..., lo + d: hi + d] + P[..., lo - d: hi - d]
    return ((P[..., lo:hi] * (2 * ref)) << 8) > alpha * S


class ShadowProblem:
    """The differentiable FP32 shadow's workload (32 ch, 1024-pt, 16 taps, 128
    frames, 32 beams): n_cpi device-generated CPIs whose channels carry seeded
    gain/phase errors (within +-1 dB and +-10 degrees), 1-8 targets each
    (bins away from the row ends, angles in +-55 degrees, per-channel SNR -30 to
    -10 dB), the steering weights of 32 beams as the initial weights, and the
    soft CA-CFAR labels over the bins whose window fits:
      label 1  the target's bin in the beam nearest to it, every frame;
      don't care  the target's bins k-2..k+2 in the beams of its main lobe
               (|sin(beam) - sin(target)| < 2/CH, the first null of the array)
               other than the labelled cell: a target between two beams and
               the PFB channel overlap light them legitimately;
      label 0  every other cell, the target's sidelobes included.
    The first n_train CPIs train the weights, the rest are held out."""

    C, N, T, F, B = 32, 1024, 16, 128, 32
    guard, ref, alpha_q, tau = 4, 16, 0x800, 0.1

    def __init__(self, torch, n_cpi=64, n_train=48, seed=55, pos_weight=512.0):
        import rd_shadow as sh

        self.torch, self.sh = torch, sh
        C, N, T, F, B = self.C, self.N, self.T, self.F, self.B
        self.n_cpi, self.n_train = n_cpi, min(n_train, n_cpi)
        self.alpha = self.alpha_q / 256.0
        rng = np.random.default_rng(seed)
        self.err = 10 ** (rng.uniform(-1, 1, C) / 20) * np.exp(1j * np.deg2rad(rng.uniform(-10, 10, C)))
        angles = list(np.rad2deg(np.arcsin(np.linspace(-0.9, 0.9, B))))
        sin_b = np.sin(np.deg2rad(angles))
        lo = self.guard + self.ref
        self.lo = lo
        label = torch.zeros((n_cpi, B, F, N), dtype=torch.float32, device="cuda")
        care = torch.ones((n_cpi, B, F, N), dtype=torch.float32, device="cuda")
        near = torch.zeros((n_cpi, B, F, N), dtype=torch.bool, device="cuda")  # a target's bins k-2..k+2
        self.scenes, self.iqs = [], []
        for q in range(n_cpi):
            nt = int(rng.integers(1, 9))
            tg = [(int(rng.integers(lo + 3, N - lo - 3)), float(rng.uniform(-55, 55)), float(rng.uniform(-30, -10)))
                  for _ in range(nt)]
            sc = rd_scenario.Scene(f"c5_{q}", CH=C, N=N, T=T, F=F, B=B, sigma=256.0, guard=self.guard, ref=self.ref,
                                   targets=tg, beam_angles=angles, seed=9000 + q)
            self.scenes.append(sc)
            self.iqs.append(device_iq(torch, sc, 9000 + q, chan_err=self.err))
            for k, ang, _ in tg:
                st = np.sin(np.deg2rad(ang))
                near[q, :, :, k - 2:k + 3] = True
                for b in np.flatnonzero(np.abs(sin_b - st) < 2.0 / C):
                    care[q, b, :, k - 2:k + 3] = 0.0
            for k, ang, _ in tg:
                bb = int(np.argmin(np.abs(sin_b - np.sin(np.deg2rad(ang)))))
                label[q, bb, :, k] = 1.0
                care[q, bb, :, k] = 1.0
        self.label = label[..., lo:N - lo].contiguous()
        self.care = care[..., lo:N - lo].contiguous()
        self.near = near[..., lo:N - lo].contiguous()
        del label, care, near
        self.pos_weight = torch.tensor(pos_weight, device="cuda")
        self.taps_w, self.W0_w = self.scenes[0].taps(), self.scenes[0].weights()
        self.x = torch.stack([self.to_cx(t).reshape(C, F * N) for t in self.iqs])
        self.h = self.to_cx(torch.from_numpy(self.taps_w.view(np.int32)).cuda()).reshape(N, T).contiguous()
        self.W0 = self.to_cx(torch.from_numpy(self.W0_w.view(np.int32)).cuda()).reshape(B, C).contiguous()

    @property
    def samples(self):
        return self.n_cpi * self.C * self.N * self.F

    def to_cx(self, words_i32):
        """packed int32 words -> complex64 / 2^15"""
        torch = self.torch
        w = words_i32.to(torch.int32)
        return torch.complex(((w << 16) >> 16).float(), (w >> 16).float()) / 32768.0

    def words_to_weights(self, words):
        return self.to_cx(self.torch.from_numpy(np.asarray(words, np.uint32).view(np.int32)).cuda()).reshape(
            self.B, self.C).contiguous()

    def _loss(self, p, sl, dtype=None):
        z = self.sh.soft_cfar_logits(p, self.N, self.guard, self.ref, self.alpha, self.tau)
        y, c = self.label[sl], self.care[sl]
        if dtype is not None:
            y, c = y.to(dtype), c.to(dtype)
        return self.torch.nn.functional.binary_cross_entropy_with_logits(
            z, y, weight=c, pos_weight=self.pos_weight.to(z.dtype))

    def loss32(self, W, gain, sl=slice(None)):
        """The FP32 shadow (hand-written HIP forward/backward) on CPIs `sl`."""
        return self._loss(self.sh.FrontEnd.apply(self.x[sl], W, gain, self.h), sl)

    def loss64(self, W, gain, x64, h64, sl=slice(None)):
        """The same model in float64 plain torch ops (the gradient oracle)."""
        return self._loss(self.sh.reference(x64[sl], W, gain, h64), sl, dtype=self.torch.float64)

    def gradient_check(self, W, gain, fd_dirs=2, eps=1e-5):
        """FP32 gradients of the loss over every CPI against float64 autograd
        and float64 central finite differences along random directions.
        Returns (relative error vs autograd, [relative errors vs FD])."""
        torch = self.torch
        W = W.detach().clone().requires_grad_()
        gain = gain.detach().clone().requires_grad_()
        self.loss32(W, gain).backward()
        gW32, gg32 = W.grad.detach().to(torch.complex128), gain.grad.detach().to(torch.complex128)
        x64, h64 = self.x.to(torch.complex128), self.h.to(torch.complex128)
        W64 = W.detach().to(torch.complex128).requires_grad_()
        g64 = gain.detach().to(torch.complex128).requires_grad_()
        gW64, gg64 = torch.autograd.grad(self.loss64(W64, g64, x64, h64), (W64, g64))
        g32 = torch.cat([gW32.reshape(-1), gg32])
        g64v = torch.cat([gW64.reshape(-1), gg64])
        rel = float((g32 - g64v).abs().norm() / g64v.abs().norm())
        gen = torch.Generator(device="cpu").manual_seed(5)
        fd_rel = []
        with torch.no_grad():
            for _ in range(fd_dirs):
                vW = torch.complex(torch.randn(self.B, self.C, generator=gen, dtype=torch.float64),
                                   torch.randn(self.B, self.C, generator=gen, dtype=torch.float64)).cuda()
                vg = torch.complex(torch.randn(self.C, generator=gen, dtype=torch.float64),
                                   torch.randn(self.C, generator=gen, dtype=torch.float64)).cuda()
                lp = self.loss64(W64 + eps * vW, g64 + eps * vg, x64, h64)
                lm = self.loss64(W64 - eps * vW, g64 - eps * vg, x64, h64)
                fd = float((lp - lm) / (2 * eps))
                # directional derivative from the gradient: Re <g, v> over re/im parts
                an = float((gW32.conj() * vW).real.sum() + (gg32.conj() * vg).real.sum())
                fd_rel.append(abs(an - fd) / abs(fd))
        return rel, fd_rel

    def train(self, steps, lr=3e-3):
        """Adam on the weights and the per-channel gains over the training
        CPIs; returns (W, gain, losses)."""
        torch = self.torch
        W = self.W0.clone().requires_grad_()
        gain = torch.ones(self.C, dtype=torch.complex64, device="cuda").requires_grad_()
        opt = torch.optim.Adam([W, gain], lr=lr)
        losses = []
        for _ in range(steps):
            opt.zero_grad()
            L = self.loss32(W, gain, slice(0, self.n_train))
            L.backward()
            opt.step()
            losses.append(float(L.detach()))
        return W, gain, losses

    def bit_exact_hits(self, ctx, words):
        """The bit-exact batched chain with Q1.15 weights `words` on every CPI;
        detections scored per region, on the training and the held-out CPIs.
        Also returns the number of (beam, frame) rows whose detection count
        differs between the CFAR rule above and the chain's CFAR."""
        torch = self.torch
        C, N, T, F, B = self.C, self.N, self.T, self.F, self.B
        wt = torch.from_numpy(np.asarray(words, np.uint32).view(np.int32)).cuda()
        tp = torch.from_numpy(self.taps_w.view(np.int32)).cuda()
        cfg = np.array([1, 0, 0, self.guard, self.guard, self.ref, self.ref, self.alpha_q], np.uint32)
        descs, bufs, keep = [], [], []
        S = C * N * F
        for q in range(self.n_cpi):
            d = dict(iq=self.iqs[q], taps=tp, weights=wt, u=torch.empty(S, dtype=torch.int32, device="cuda"),
                     X=torch.empty(S, dtype=torch.int32, device="cuda"),
                     Xa=torch.empty(S, dtype=torch.int32, device="cuda"),
                     power=torch.empty(B * F * N, dtype=torch.int32, device="cuda"),
                     row_count=torch.empty(B * F, dtype=torch.int32, device="cuda"))
            keep.append(d)
            descs.append(rd_amd.CpiDesc(C, N, T, F, B, 2, rd_scenario.SCHED_ALL, rd_scenario.MAX_GUARD,
                                        rd_scenario.MAX_REF, rd_amd.CfarCfg.of(cfg)))
            bufs.append(rd_amd.CpiBufs(**{k: v.data_ptr() for k, v in d.items()}, events_cap=0))
        torch.cuda.synchronize()
        bt = rd_amd.Batch(ctx, descs, bufs, keep=keep)
        bt.launch()
        bt.events()
        bt.close()
        pw = torch.stack([d["power"].view(torch.int32).reshape(B, F, N) for d in keep]).to(torch.int64) & 0xFFFFFFFF
        hit = hard_cfar_hits(torch, pw, N, self.guard, self.ref, self.alpha_q)
        rc = torch.stack([d["row_count"].reshape(B, F) for d in keep]).to(torch.int64) - 1
        consistency = int((hit.sum(-1) != rc).sum())
        out = {}
        for name, sl in (("train", slice(0, self.n_train)), ("held_out", slice(self.n_train, self.n_cpi))):
            h, y, c, nr = hit[sl], self.label[sl] > 0, self.care[sl] > 0, self.near[sl]
            neg = c & ~y
            cells_noise = int((neg & ~nr).sum())
            out[name] = dict(cpis=int(h.shape[0]), pd=round(float(h[y].float().mean()), 4) if bool(y.any()) else None,
                             noise_false_alarms=int(h[neg & ~nr].sum()),
                             noise_fa_rate=round(int(h[neg & ~nr].sum()) / max(1, cells_noise), 7),
                             sidelobe_hits=int(h[neg & nr].sum()), mainlobe_neighbour_hits=int(h[~c].sum()))
        del pw, hit, keep
        return out, consistency
