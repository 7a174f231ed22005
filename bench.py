#!/usr/bin/env python3
"""Throughput of the bit-exact PFB -> FFT -> align -> beamformer/power ->
covariance -> CFAR -> event-packing chain on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload medium]

A step is one CPI of the workload through the whole chain (one HIP graph launch),
inputs already resident in HBM. Steps cycle through a ring of distinct input
CPIs (default 32, 537 MB for the medium workload) so the input stream is not
served from the 256 MB last-level cache. One JSON line is printed:

  value      whole-job Msamples/s (one sample = one complex int16 input sample
             of one channel), wall clock over exactly K steps
  roofline   the dominant kernel: algorithmic bytes per launch / its mean
             launch time (HIP events on the library's stream) vs 8 TB/s HBM
  chain      the fused-chain algorithmic bytes (input + taps + weights +
             covariance + events) per CPI over the chain's time per CPI
  cpu_baseline  the reference's C++ models (oracle/_ref) or the CPU
             restatement, on this host's cores, on a bounded sample of CPIs
  mismatches bit-exact comparison of one CPI's power, covariance and events
             against the CPU run (must be 0)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path[:0] = [str(ROOT), str(ROOT / "agilex-7-radar-detect_amd")]

import rd_amd  # noqa: E402
import rd_scenario  # noqa: E402
import rd_workloads  # noqa: E402

HBM_PEAK_GBS = 8000.0
VALU_SIMDS = 1024  # 256 CUs x 4 SIMDs
VALU_CLOCK_HZ = 2.4e9


def note(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algo_chain_bytes(sc, n_events: int) -> int:
    """Compulsory traffic of the fused chain for one CPI: I/Q in, taps, weights,
    Hermitian covariance out (int64 re+im), 176-bit events (22 B)."""
    return (4 * sc.CH * sc.N * sc.F + 4 * sc.N * sc.T + 4 * sc.B * sc.CH + 16 * sc.N * sc.CH * (sc.CH + 1) // 2
            + 22 * n_events)


def stage_kernels(ctx, ch, sc, lean=True):
    """Per-stage launchers (device buffers of `ch`), for the per-kernel timing.
    lean: the stages as the timed (lean) chain graph runs them, without the
    PFB and FFT flag outputs."""
    L, h, b = ctx.L, ctx.h, ch.b
    S = sc.CH * sc.N * sc.F
    cols = sc.N * sc.F
    pfb_flags = None if lean or not b["pfb_flags"] else b["pfb_flags"].ptr
    fft_flags = None if lean or not b["fft_flags"] else b["fft_flags"].ptr

    def pfb():
        ctx.lib.check("pfb", L.rd_pfb_run(h, b["iq"].ptr, b["taps"].ptr, sc.N, sc.T, sc.CH, sc.F, None, b["u"].ptr,
                                          pfb_flags))

    def fft():
        ctx.lib.check("fft", L.rd_fft_run(h, b["u"].ptr, sc.N, 2, 0xFFFFFFFF, 1, sc.CH * sc.F, None, b["X"].ptr,
                                          fft_flags))

    def bf():
        # the chain's beamformer: FFT layout in, power + aligned copy out (lean outputs)
        ctx.lib.check("bf", L.rd_bf_stage(h, b["X"].ptr, b["weights"].ptr, sc.CH, sc.B, sc.F, sc.N, None, None,
                                          b["power"].ptr, b["Xa"].ptr))

    def cov():
        ctx.lib.check("cov", L.rd_cov_run(h, b["Xa"].ptr, sc.CH, sc.F, sc.N, b["R"].ptr))

    # algorithmic bytes per launch of each stage kernel (its compulsory reads + writes)
    flag_b = 1 if pfb_flags else 0
    stages = {
        "pfb": (pfb, S * (4 + 4 + flag_b) + 4 * sc.N * sc.T),
        "fft": (fft, S * 8 + (4 * sc.CH * sc.F if fft_flags else 0)),
        "bf": (bf, S * 4 + 4 * sc.B * sc.CH + S * 4 + sc.B * cols * 4),
        "cov": (cov, S * 4 + 16 * sc.N * sc.CH * (sc.CH + 1) // 2),
    }
    return stages


def time_stage(torch, stream, fn, reps: int) -> float:
    """Mean launch duration in ms of `fn` over `reps` back-to-back launches,
    by HIP events recorded on the library's stream."""
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    fn()
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def cpu_baseline(sc, iq, taps, W, budget_s: float):
    import oracle

    # the reference headers cap the FFT at N <= 1024 (fft_ref.hpp:216-219)
    kind = "reference" if oracle.ref_available() and sc.N <= 1024 else "port"
    orc = oracle.load("ref" if kind == "reference" else "oracle")
    threads = max(1, min(16, os.cpu_count() or 1))
    t0 = time.perf_counter()
    n = 0
    out = None
    while True:
        out = orc.chain(iq, taps, W, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg(), rd_scenario.MAX_GUARD,
                        rd_scenario.MAX_REF, keep=("power", "R", "events"), threads=threads)
        n += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = time.perf_counter() - t0
    # the same chain on one thread, one CPI
    t1 = time.perf_counter()
    orc.chain(iq, taps, W, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg(), rd_scenario.MAX_GUARD, rd_scenario.MAX_REF,
              keep=("power", "R", "events"), threads=1)
    dt1 = time.perf_counter() - t1
    return dict(value=round(n * sc.samples / dt / 1e6, 3), unit="Msamples/s", cores=threads, kind=kind,
                sample=f"{n} CPI(s) of the {sc.name} workload ({sc.samples} samples each), full chain, "
                       f"{threads} threads, {dt:.1f} s",
                value_1thread=round(sc.samples / dt1 / 1e6, 3), hardware_concurrency=os.cpu_count(),
                cpu_model=cpu_model()), out


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


# ----------------------------------------------------------------------------- large leg
def profiled(stage: str, workload: str):
    """HBM bytes and VALU wave-instructions per dispatch of a stage kernel, from
    the committed rocprofv3 counter summary of that workload (None when absent)."""
    path = ROOT / "profiles" / f"chain_{workload}_hbm_counters.md"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 8 and cells[0] == stage:
            try:
                return dict(hbm_bytes=int(cells[5]), valu_insts=int(cells[7]))
            except ValueError:
                return None
    return None


def valu_frac(valu_insts: int, ms: float) -> float:
    """Fraction of the VALU issue rate a launch used: wave instructions x 4
    cycles over 1024 SIMDs x 2.4 GHz x the launch time."""
    return round(valu_insts * 4 / (VALU_SIMDS * VALU_CLOCK_HZ * ms * 1e-3), 4)


def large_leg(torch, ctx, steps: int, verify: bool, stage_reps: int):
    """The large workload (64 ch x 4096-pt x 16 taps x 512 frames, 64 beams):
    graph replay over two device-generated CPIs, per-stage times with the
    beamformer and covariance MFMA rates, and one CPI checked against the
    oracle on the host's cores."""
    sc = rd_scenario.config("large")
    taps, W = sc.taps(), sc.weights()
    ch = rd_amd.Chain(ctx, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg(), rd_scenario.MAX_GUARD, rd_scenario.MAX_REF,
                      keep_y=False, keep_flags=False, keep_R=True)
    note("large: generating 2 CPIs on the device")
    ring = [rd_workloads.device_iq(torch, sc, 7001 + i) for i in range(2)]
    iq0 = ring[0].cpu().numpy().view(np.uint32)
    ch.load(iq0, taps, W)
    totals = ctx.buf(16)
    graphs = [ch.graph(iq_ptr=ring[i].data_ptr(), total_ptr=totals.ptr + 8 * i, lean=True) for i in range(2)]
    stream = torch.cuda.ExternalStream(ctx.stream)
    for i in range(2):
        graphs[i].launch()
    ctx.sync()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for i in range(steps):
        graphs[i % 2].launch()
    e1.record(stream)
    ctx.sync()
    dt = time.perf_counter() - t0
    ms = dt * 1e3 / steps
    note("large: per-stage kernel timing")
    per = {}
    for name, (fn, nbytes) in stage_kernels(ctx, ch, sc).items():
        t = time_stage(torch, stream, fn, stage_reps)
        per[name] = dict(ms=round(t, 4), algo_bytes=nbytes, gbs=round(nbytes / (t * 1e-3) / 1e9, 1))
    pairs = sc.CH * (sc.CH + 1) // 2
    cols = sc.N * sc.F
    # int8 MFMA work as issued: 4 digit products x (re, im) x 2 ops per MAC
    mfma = {"bf": 32 * sc.B * sc.CH * cols, "cov": 32 * pairs * sc.F * sc.N}
    for k, ops in mfma.items():
        per[k]["mfma_int8_tops"] = round(ops / (per[k]["ms"] * 1e-3) / 1e12, 1)
        per[k]["mfma_frac_of_5pops"] = round(ops / (per[k]["ms"] * 1e-3) / 5e15, 4)
    for g in graphs:
        g.close()
    res = dict(workload=f"large: {sc.CH} ch x {sc.N}-pt PFB/FFT x {sc.T} taps x {sc.F} frames, {sc.B} beams, "
                        f"CA-CFAR {sc.ref}/{sc.guard}",
               cpi_samples=sc.samples, steps=steps, ms_per_cpi=round(ms, 4),
               msamples_s=round(sc.samples / (ms * 1e-3) / 1e6, 1), stages=per, mismatches=None)
    if verify:
        import oracle

        note("large: checking one CPI against the oracle on the host")
        ch.enqueue()
        ctx.sync()
        n_ev = int(ch.total.download(np.uint64, 1)[0])
        got = ch.fetch(("power", "R"), n_events=n_ev)
        threads = max(1, min(16, os.cpu_count() or 1))
        t0 = time.perf_counter()
        ref = oracle.load("oracle").chain(iq0, taps, W, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg(),
                                          rd_scenario.MAX_GUARD, rd_scenario.MAX_REF, keep=("power", "R", "events"),
                                          threads=threads)
        cpu_s = time.perf_counter() - t0
        mism = int(np.count_nonzero(got["power"] != ref["power"]) + np.count_nonzero(got["R"] != ref["R"]))
        if n_ev != ref["n_events"]:
            mism += abs(n_ev - ref["n_events"]) + 1
        else:
            mism += int(np.count_nonzero(got["events"] != ref["events"]))
        res.update(mismatches=mism, events_per_cpi=n_ev,
                   oracle_cpu=dict(seconds=round(cpu_s, 2), threads=threads,
                                   msamples_s=round(sc.samples / cpu_s / 1e6, 2), kind="port"))
    del ring
    return res


# ----------------------------------------------------------------------------- batched leg
def batched_leg(torch, ctx, n_cpi, reps, autotune, oracle_cpis):
    """256 uneven CPIs: one batched graph launch vs one CPI per launch, both
    checked against each other on every CPI and against the CPU on a sample."""
    scenes = rd_scenario.uneven_batch(n_cpi)
    note(f"batch: generating {n_cpi} uneven CPIs on the device")
    bs = rd_workloads.BatchSet(torch, scenes, seed0=77_000)
    L, h = ctx.L, ctx.h
    tot = np.zeros(1, np.uint64)
    totals_dev = torch.zeros(n_cpi, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()

    def plain_sync():  # the drop-in per-CPI call: chain + event count back to the host
        for d, b in zip(bs.descs, bs.bufs):
            ctx.lib.check("rd_chain_run", L.rd_chain_run(h, rd_amd.C.byref(d), rd_amd.C.byref(b),
                                                          tot.ctypes.data_as(rd_amd.C.POINTER(rd_amd.C.c_uint64))))

    def plain_async():  # one CPI per launch sequence, no host round trip between CPIs
        for i, (d, b) in enumerate(zip(bs.descs, bs.bufs)):
            ctx.lib.check("rd_chain_enqueue", L.rd_chain_enqueue(h, rd_amd.C.byref(d), rd_amd.C.byref(b),
                                                                  totals_dev.data_ptr() + 8 * i))
        ctx.sync()

    def wall(fn, n):
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        ctx.sync()
        return (time.perf_counter() - t0) / n

    note("batch: plain one-CPI-per-launch passes")
    plain_async()
    bs.fit_event_caps(torch, totals_dev.cpu().numpy())
    t_async = wall(plain_async, reps)
    t_sync = wall(plain_sync, max(1, reps // 2))
    note(f"batch: creating the batch graph (autotune={autotune})")
    t0 = time.perf_counter()
    batch = rd_amd.Batch(ctx, bs.descs, bs.bufs, autotune=autotune, keep=bs)
    t_create = time.perf_counter() - t0
    batch.launch()
    t_batch = wall(batch.launch, reps)
    totals = batch.events()
    note("batch: verifying every CPI against its one-CPI launch")
    bad_cpis = rd_workloads.verify_batch_vs_plain(torch, ctx, bs, batch, totals)
    note(f"batch: oracle check on {oracle_cpis} CPIs")
    orc_bad, orc_n = rd_workloads.verify_batch_vs_oracle(torch, ctx, bs, totals, oracle_cpis)
    samples = bs.samples
    plan = batch.plan().strip().splitlines()
    batch.close()
    return {
        "workload": f"{n_cpi} uneven CPIs: ch in {{8,16,32,64}}, FFT in {{256,1024,4096}}, frames U[32,512], "
                    f"beams = ch, 16 taps, 0-16 targets, sigma 256, CA-CFAR 16/4",
        "samples": samples,
        "batched": {"ms": round(t_batch * 1e3, 3), "msamples_s": round(samples / t_batch / 1e6, 1),
                    "impl": "one hipGraph launch: per-stage shape classes as persistent kernels over all CPIs"},
        "plain_async": {"ms": round(t_async * 1e3, 3), "msamples_s": round(samples / t_async / 1e6, 1),
                        "impl": "rd_chain_enqueue per CPI on one stream, one sync at the end"},
        "plain_sync": {"ms": round(t_sync * 1e3, 3), "msamples_s": round(samples / t_sync / 1e6, 1),
                       "impl": "rd_chain_run per CPI (event count returned to the host each CPI)"},
        "speedup_vs_plain_async": round(t_async / t_batch, 3),
        "speedup_vs_plain_sync": round(t_sync / t_batch, 3),
        "autotune": bool(autotune), "create_s": round(t_create, 2), "plan": plan,
        "mismatching_cpis": bad_cpis, "oracle_cpis": orc_n, "oracle_mismatches": orc_bad,
        "events": int(np.sum(totals)),
    }


# ----------------------------------------------------------------------------- FP32 shadow leg
def hard_cfar_hits(torch, power, N, guard, ref, alpha):
    """The bit-exact CA-CFAR decision (cfar_model.hpp:296-301) on int power
    rows [..., N] in exact int64 torch ops: hit iff (C * n) << 8 > alpha * S."""
    P = power.to(torch.int64)
    lo, hi = guard + ref, N - guard - ref
    S = torch.zeros_like(P[..., lo:hi])
    for d in range(guard + 1, guard + ref + 1):
        S += P[..., lo + d: hi + d] + P[..., lo - d: hi - d]
    return ((P[..., lo:hi] * (2 * ref)) << 8) > alpha * S


def shadow_leg(torch, ctx, n_cpi, steps, fd_dirs):
    """The differentiable FP32 shadow on 64 CPIs of 32 ch x 1024-pt x 128
    frames: a training step's forward + backward timed, gradients checked
    against float64 (plain autograd and central finite differences), the
    weights trained, quantised to Q1.15 (channel gains folded in) and rerun
    through the bit-exact batched chain."""
    import rd_shadow as sh

    C, N, T, F, B = 32, 1024, 16, 128, 32
    guard, ref, alpha_q = 4, 16, 0x800
    alpha, tau = alpha_q / 256.0, 0.1
    rng = np.random.default_rng(55)
    err = 10 ** (rng.uniform(-1, 1, C) / 20) * np.exp(1j * np.deg2rad(rng.uniform(-10, 10, C)))
    angles = list(np.rad2deg(np.arcsin(np.linspace(-0.9, 0.9, B))))
    sin_b = np.sin(np.deg2rad(angles))
    mask = torch.zeros((n_cpi, B, F, N), dtype=torch.float32, device="cuda")
    # a target's own response outside its cell (its bin in the other beams, and
    # the neighbouring bins the PFB channel overlap reaches) is neither a
    # detection to reward nor a false alarm to penalise
    care = torch.ones((n_cpi, B, F, N), dtype=torch.float32, device="cuda")
    scenes, iqs = [], []
    note(f"shadow: generating {n_cpi} CPIs (32 ch x 1024 x 128, channel gain/phase errors)")
    for q in range(n_cpi):
        nt = int(rng.integers(1, 9))
        tg = [(int(rng.integers(guard + ref + 1, N - guard - ref - 1)), float(rng.uniform(-55, 55)),
               float(rng.uniform(-30, -10))) for _ in range(nt)]
        sc = rd_scenario.Scene(f"c5_{q}", CH=C, N=N, T=T, F=F, B=B, sigma=256.0, guard=guard, ref=ref, targets=tg,
                               beam_angles=angles, seed=9000 + q)
        scenes.append(sc)
        iqs.append(rd_workloads.device_iq(torch, sc, 9000 + q, chan_err=err))
        for k, ang, _ in tg:
            care[q, :, :, k - 2:k + 3] = 0.0
        for k, ang, _ in tg:
            bb = int(np.argmin(np.abs(sin_b - np.sin(np.deg2rad(ang)))))
            mask[q, bb, :, k] = 1.0
            care[q, bb, :, k] = 1.0
    lo = guard + ref
    mask_v = mask[..., lo:N - lo].contiguous()
    care_v = care[..., lo:N - lo].contiguous()
    del mask, care
    taps_w, W0_w = scenes[0].taps(), scenes[0].weights()

    def to_cx(words_i32):  # packed int32 words -> complex64 / 2^15
        w = words_i32.to(torch.int32)
        re = ((w << 16) >> 16).float()
        im = (w >> 16).float()
        return torch.complex(re, im) / 32768.0

    x = torch.stack([to_cx(t).reshape(C, F * N) for t in iqs])
    h = to_cx(torch.from_numpy(taps_w.view(np.int32)).cuda()).reshape(N, T).contiguous()
    W = to_cx(torch.from_numpy(W0_w.view(np.int32)).cuda()).reshape(B, C).contiguous().requires_grad_()
    gain = torch.ones(C, dtype=torch.complex64, device="cuda").requires_grad_()

    def loss32(W_=None, g_=None):
        p = sh.FrontEnd.apply(x, W if W_ is None else W_, gain if g_ is None else g_, h)
        return sh.loss_fn(sh.soft_cfar_logits(p, N, guard, ref, alpha, tau), mask_v, care_v)

    # timing of one training step (forward + backward)
    for _ in range(2):
        loss32().backward()
    W.grad = None
    gain.grad = None
    torch.cuda.synchronize()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        W.grad = None
        gain.grad = None
        loss32().backward()
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t0) / reps
    samples = n_cpi * C * N * F

    # gradient checks against float64
    note("shadow: float64 gradient checks")
    W.grad = None
    gain.grad = None
    L32 = loss32()
    L32.backward()
    gW32, gg32 = W.grad.detach().clone(), gain.grad.detach().clone()
    x64, h64, m64, c64 = x.to(torch.complex128), h.to(torch.complex128), mask_v.double(), care_v.double()

    def loss64(W_, g_):
        p = sh.reference(x64, W_, g_, h64)
        return sh.loss_fn(sh.soft_cfar_logits(p, N, guard, ref, alpha, tau), m64, c64)

    W64 = W.detach().to(torch.complex128).requires_grad_()
    g64 = gain.detach().to(torch.complex128).requires_grad_()
    L64 = loss64(W64, g64)
    gW64, gg64 = torch.autograd.grad(L64, (W64, g64))
    g32 = torch.cat([gW32.reshape(-1), gg32]).to(torch.complex128)
    g64v = torch.cat([gW64.reshape(-1), gg64])
    autograd_rel = float((g32 - g64v).abs().norm() / g64v.abs().norm())
    fd_rel = []
    gen = torch.Generator(device="cpu").manual_seed(5)
    eps = 1e-5
    with torch.no_grad():
        for _ in range(fd_dirs):
            vW = torch.complex(torch.randn(B, C, generator=gen, dtype=torch.float64),
                               torch.randn(B, C, generator=gen, dtype=torch.float64)).cuda()
            vg = torch.complex(torch.randn(C, generator=gen, dtype=torch.float64),
                               torch.randn(C, generator=gen, dtype=torch.float64)).cuda()
            lp = loss64(W64 + eps * vW, g64 + eps * vg)
            lm = loss64(W64 - eps * vW, g64 - eps * vg)
            fd = float((lp - lm) / (2 * eps))
            # directional derivative from the gradient: Re <g, v> over re/im parts
            an = float((gW32.to(torch.complex128).conj() * vW).real.sum() + (gg32.to(torch.complex128).conj() * vg).real.sum())
            fd_rel.append(abs(an - fd) / abs(fd))
    del x64, m64, c64

    # training
    note(f"shadow: {steps} training steps")
    opt = torch.optim.Adam([W, gain], lr=1e-3)
    losses = []
    for _ in range(steps):
        opt.zero_grad()
        L = loss32()
        L.backward()
        opt.step()
        losses.append(float(L.detach()))

    # quantise back and rerun the bit-exact chain (batched) with both weight sets
    note("shadow: bit-exact rerun with the quantised weights")
    Wq = sh.quantise_weights(W, gain)
    with torch.no_grad():  # the FP32 loss at the quantised weights (gains folded in, unit gains)
        Wq_c = to_cx(torch.from_numpy(Wq.view(np.int32)).cuda()).reshape(B, C).contiguous()
        W0_c = to_cx(torch.from_numpy(W0_w.view(np.int32)).cuda()).reshape(B, C).contiguous()
        ones = torch.ones(C, dtype=torch.complex64, device="cuda")
        loss_q = [float(loss32(W0_c, ones)), float(loss32(Wq_c, ones))]
    hits = {}
    consistency = 0
    for name, wq in (("initial", W0_w), ("trained", Wq)):
        wt = torch.from_numpy(np.asarray(wq, np.uint32).view(np.int32)).cuda()
        tp = torch.from_numpy(taps_w.view(np.int32)).cuda()
        descs, bufs, keep = [], [], []
        cfg = np.array([1, 0, 0, guard, guard, ref, ref, alpha_q], np.uint32)
        for q in range(n_cpi):
            S = C * N * F
            d = dict(iq=iqs[q], taps=tp, weights=wt, u=torch.empty(S, dtype=torch.int32, device="cuda"),
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
        hit = hard_cfar_hits(torch, pw, N, guard, ref, alpha_q)
        # the decision rule reproduces the bit-exact CFAR's per-row detection counts
        rc = torch.stack([d["row_count"].reshape(B, F) for d in keep]).to(torch.int64) - 1
        consistency += int((hit.sum(-1) != rc).sum())
        t = mask_v > 0
        fa_cells = (mask_v == 0) & (care_v > 0)
        hits[name] = dict(pd=round(float(hit[t].float().mean()), 4), false_alarms=int(hit[fa_cells].sum()),
                          sidelobe_hits=int(hit[care_v == 0].sum()))
        del pw, hit, keep
    return {
        "workload": f"{n_cpi} CPIs x 32 ch x 1024-pt x 128 frames, FP32, 32x32 complex weights + 32 channel gains, "
                    f"soft CA-CFAR sigmoid((p/mean - {alpha})/{tau}), BCE vs target cells",
        "samples": samples,
        "train_step_ms": round(t_step * 1e3, 2),
        "train_step_msamples_s": round(samples / t_step / 1e6, 1),
        "grad_rel_err_vs_fp64_autograd": autograd_rel,
        "grad_rel_err_vs_fp64_finite_diff": fd_rel,
        "loss_first_last": [losses[0], losses[-1]] if losses else None,
        "loss_fp32_at_q15_weights_initial_trained": loss_q,
        "bit_exact_rerun": hits,
        "cfar_rule_vs_row_counts_mismatches": consistency,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=320)
    ap.add_argument("--warmup", type=int, default=32)
    ap.add_argument("--workload", default="medium", choices=["tiny", "medium", "large"])
    ap.add_argument("--ring", type=int, default=32, help="distinct input CPIs cycled through")
    ap.add_argument("--streams", type=int, default=3,
                    help="contexts (one HIP stream each) the timed CPIs are dealt to round-robin")
    ap.add_argument("--cpu-seconds", type=float, default=3.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--stage-reps", type=int, default=50)
    ap.add_argument("--batch-cpis", type=int, default=256, help="CPIs of the uneven batch leg (0: skip it)")
    ap.add_argument("--batch-reps", type=int, default=3)
    ap.add_argument("--batch-oracle", type=int, default=3, help="smallest batch CPIs also checked on the CPU")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--batch-only", action="store_true", help="run only the batched leg")
    ap.add_argument("--shadow-cpis", type=int, default=64, help="CPIs of the FP32 shadow leg (0: skip it)")
    ap.add_argument("--shadow-steps", type=int, default=20)
    ap.add_argument("--shadow-only", action="store_true", help="run only the FP32 shadow leg")
    ap.add_argument("--large-steps", type=int, default=10, help="graph replays of the large-workload leg (0: skip it)")
    ap.add_argument("--large-verify", type=int, default=1, help="check one large CPI against the oracle (0: skip)")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (HBM-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              f"(CPIs are independent: N GPUs would each run their own share)", file=sys.stderr)
        sys.exit(2)

    import torch

    lib = rd_amd.load()
    if lib.device_count() < 1:
        sys.exit("no GPU visible")
    ctx = lib.context(0)
    if args.shadow_only:
        res = shadow_leg(torch, ctx, args.shadow_cpis, args.shadow_steps, 2)
        print(json.dumps({"metric": "FP32 shadow training step (forward + backward)",
                          "value": res["train_step_msamples_s"], "unit": "Msamples/s", "n_gpus": 1,
                          "steps": 5, "warmup": 2, "ms_per_step": res["train_step_ms"], "higher_is_better": True,
                          "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
                          "data": "synthetic (device-generated scenes)", "config": {"workload": res["workload"]},
                          "shadow": res}), flush=True)
        ctx.close()
        return
    if args.batch_only:
        res = batched_leg(torch, ctx, args.batch_cpis, args.batch_reps, not args.no_autotune, args.batch_oracle)
        print(json.dumps({"metric": "bit-exact Msamples/s PFB->CFAR (batched uneven CPIs, one graph launch)",
                          "value": res["batched"]["msamples_s"], "unit": "Msamples/s", "n_gpus": 1,
                          "steps": args.batch_reps, "warmup": 1, "ms_per_step": res["batched"]["ms"],
                          "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                          "dtype": "int16 in, int32/int64 exact", "data": "synthetic (device-generated scenes)",
                          "config": {"workload": res["workload"]}, "batch": res}), flush=True)
        ctx.close()
        return
    sc = rd_scenario.config(args.workload)
    taps, W = sc.taps(), sc.weights()
    ch = rd_amd.Chain(ctx, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg(), rd_scenario.MAX_GUARD, rd_scenario.MAX_REF,
                      keep_y=True, keep_flags=True, keep_R=True)
    note(f"workload {sc.name}: generating {args.ring} input CPIs")
    ch.load(sc.iq(seed=0), taps, W)
    ring = []
    for i in range(args.ring):
        ring.append(ctx.put(sc.iq(seed=1000 + i)))
        if sc.samples > 50_000_000:
            note(f"  input CPI {i + 1}/{args.ring} resident")
    totals = ctx.buf(8 * args.ring)

    # correctness of one CPI against the CPU run (also the timed CPU baseline)
    iq0 = sc.iq(seed=1000)
    cpu = None
    if not args.no_cpu:
        note("CPU baseline")
        cpu, cpu_out = cpu_baseline(sc, iq0, taps, W, args.cpu_seconds)
    ch.enqueue(iq_ptr=ring[0].ptr)
    ctx.sync()
    n_ev = int(ch.total.download(np.uint64, 1)[0])
    got = ch.fetch(("power", "R"), n_events=n_ev)
    mismatches = None
    if cpu is not None:
        mismatches = int(np.count_nonzero(got["power"] != cpu_out["power"]) + np.count_nonzero(got["R"] != cpu_out["R"]))
        if n_ev != cpu_out["n_events"]:
            mismatches += abs(n_ev - cpu_out["n_events"]) + 1
        else:
            mismatches += int(np.count_nonzero(got["events"] != cpu_out["events"]))

    stream = torch.cuda.ExternalStream(ctx.stream)
    note(f"capturing {args.ring} chain graphs (lean and full-output)")
    graphs = [ch.graph(iq_ptr=ring[i].ptr, total_ptr=totals.ptr + 8 * i, lean=True) for i in range(args.ring)]
    graphs_full = [ch.graph(iq_ptr=ring[i].ptr, total_ptr=totals.ptr + 8 * i) for i in range(args.ring)]

    def timed(launch_one, steps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        ctx.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e0.record(stream)
        for i in range(steps):
            launch_one(i)
        t_host = time.perf_counter() - t0
        e1.record(stream)
        ctx.sync()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, t_host, e0.elapsed_time(e1) * 1e-3

    def run_graph(i):
        graphs[i % args.ring].launch()

    def run_graph_full(i):
        graphs_full[i % args.ring].launch()

    def run_stream(i):
        ch.enqueue(iq_ptr=ring[i % args.ring].ptr, total_ptr=totals.ptr + 8 * (i % args.ring))

    note(f"warmup {args.warmup} + {args.steps} timed steps (graph replay, one stream)")
    timed(run_graph, args.warmup)
    dt, host_dt, gpu_dt = timed(run_graph, args.steps)
    one_ms = dt * 1e3 / args.steps

    # the same lean graphs on args.streams contexts (one HIP stream each),
    # CPIs dealt round-robin, so one CPI's short tail kernels overlap the
    # next CPI's stages; every context has its own chain buffers and totals
    n_str = max(1, args.streams)
    ctxs = [ctx] + [rd_amd.Ctx(ctx.lib) for _ in range(n_str - 1)]
    chains, s_graphs, s_totals = [ch], [graphs], [totals]
    for c in ctxs[1:]:
        chk = rd_amd.Chain(c, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg(), rd_scenario.MAX_GUARD,
                           rd_scenario.MAX_REF, keep_y=False, keep_flags=False, keep_R=True)
        chk.load(sc.iq(seed=0), taps, W)
        tk = c.buf(8 * args.ring)
        chains.append(chk)
        s_totals.append(tk)
        s_graphs.append([chk.graph(iq_ptr=ring[i].ptr, total_ptr=tk.ptr + 8 * i, lean=True)
                         for i in range(args.ring)])

    def timed_streams(steps):
        for c in ctxs:
            c.sync()
        t0 = time.perf_counter()
        for i in range(steps):
            s_graphs[i % n_str][(i // n_str) % args.ring].launch()
        for c in ctxs:
            c.sync()
        return time.perf_counter() - t0

    note(f"warmup {args.warmup} + {args.steps} timed steps (graph replay over {n_str} streams)")
    timed_streams(args.warmup)
    dt_s = timed_streams(args.steps)
    ms_per_step = dt_s * 1e3 / args.steps
    value = args.steps * sc.samples / dt_s / 1e6
    # every context saw the same CPIs: their event totals must agree
    tot0 = totals.download(np.uint64, args.ring)
    used = min(args.ring, args.steps // n_str)  # ring slots every context filled
    for tk in s_totals[1:]:
        if not np.array_equal(tk.download(np.uint64, args.ring)[:used], tot0[:used]):
            sys.exit("event totals differ between the streams' contexts")
    for gl in s_graphs[1:]:
        for g in gl:
            g.close()
    for c in ctxs[1:]:
        c.close()
    note("full-output graphs")
    timed(run_graph_full, min(args.warmup, 8))
    f_dt, f_host, f_gpu = timed(run_graph_full, args.steps)
    note("stream-enqueue comparison")
    timed(run_stream, min(args.warmup, 8))
    s_dt, s_host, s_gpu = timed(run_stream, args.steps)
    totals_ok = totals.download(np.uint64, args.ring)
    if not np.all(totals_ok[:min(args.ring, args.steps)] > 0):
        sys.exit("a timed CPI produced no events (summaries are always emitted)")

    # per-kernel launch durations (HIP events on the library's stream)
    note("per-stage kernel timing")
    stages = stage_kernels(ctx, ch, sc)
    per = {}
    for name, (fn, nbytes) in stages.items():
        ms = time_stage(torch, stream, fn, args.stage_reps)
        per[name] = dict(ms=round(ms, 5), algo_bytes=nbytes, gbs=round(nbytes / (ms * 1e-3) / 1e9, 1))
    # the same PFB launch with complex taps (the imaginary part of every tap
    # set to 1): the general four-product window instead of the real-tap one
    taps_cx = ctx.put(np.asarray(sc.taps(), np.uint32) | np.uint32(1 << 16))
    u_cx = ctx.buf(4 * sc.CH * sc.N * sc.F)

    def pfb_cx():
        ctx.lib.check("pfb", ctx.L.rd_pfb_run(ctx.h, ch.b["iq"].ptr, taps_cx.ptr, sc.N, sc.T, sc.CH, sc.F, None,
                                              u_cx.ptr, None))

    pfb_cx_ms = time_stage(torch, stream, pfb_cx, args.stage_reps)
    dom = max(per, key=lambda k: per[k]["ms"])
    ach = per[dom]["gbs"]
    prof = profiled(dom, sc.name)
    chain_bytes = algo_chain_bytes(sc, n_ev)
    out = {
        "metric": "bit-exact Msamples/s PFB->CFAR (plain GPU, one CPI per graph launch)",
        "value": round(value, 2),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int16 in, int32/int64 exact",
        "data": "synthetic (seeded complex Gaussian noise + plane-wave targets, ring of distinct CPIs)",
        "config": {"workload": f"{sc.name}: {sc.CH} ch x {sc.N}-pt PFB/FFT x {sc.T} taps x {sc.F} frames, "
                               f"{sc.B} beams, CA-CFAR {sc.ref}/{sc.guard}", "cpi_samples": sc.samples,
                   "ring": args.ring, "streams": n_str,
                   "impl": "one hipGraph per CPI: PFB, FFT, BF+power MFMA (writes the aligned copy), cov MFMA, "
                           f"CFAR rows/scan/gather; CPIs dealt round-robin to {n_str} contexts (one HIP stream "
                           "each)",
                   "outputs": "events + covariance timed; every stage output verified bit-exact on a separate "
                              "CPI each run (timing.graph_full_outputs also writes them all)"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": prof["hbm_bytes"] if prof else None,
                     "valu_frac": valu_frac(prof["valu_insts"], per[dom]["ms"]) if prof else None},
        "chain": {"algo_bytes_per_cpi": chain_bytes, "achieved_gbs": round(chain_bytes / (ms_per_step * 1e-3) / 1e9, 1),
                  "frac": round(chain_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "timing": {"graph_streams": {"streams": n_str, "wall_ms": round(ms_per_step, 4)},
                   "pfb_complex_taps_ms": round(pfb_cx_ms, 5),
                   "graph": {"wall_ms": round(one_ms, 4), "host_ms": round(host_dt * 1e3 / args.steps, 4),
                             "gpu_ms": round(gpu_dt * 1e3 / args.steps, 4),
                             "msamples_s": round(args.steps * sc.samples / dt / 1e6, 2)},
                   "graph_full_outputs": {"wall_ms": round(f_dt * 1e3 / args.steps, 4),
                                          "gpu_ms": round(f_gpu * 1e3 / args.steps, 4),
                                          "msamples_s": round(args.steps * sc.samples / f_dt / 1e6, 2)},
                   "stream": {"wall_ms": round(s_dt * 1e3 / args.steps, 4), "host_ms": round(s_host * 1e3 / args.steps, 4),
                              "gpu_ms": round(s_gpu * 1e3 / args.steps, 4),
                              "msamples_s": round(args.steps * sc.samples / s_dt / 1e6, 2)}},
        "stages": per,
        "cpu_baseline": cpu,
        "mismatches": mismatches,
        "events_per_cpi": n_ev,
    }
    if args.calib_copy_mb:
        nbytes = args.calib_copy_mb << 20
        src = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        src.fill_(1)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        out["calib_copy_bytes"] = nbytes
    if args.batch_cpis:
        for g in graphs + graphs_full:
            g.close()
        graphs, graphs_full = [], []
        out["batch"] = batched_leg(torch, ctx, args.batch_cpis, args.batch_reps, not args.no_autotune,
                                   args.batch_oracle)
    if args.shadow_cpis:
        torch.cuda.empty_cache()
        out["shadow"] = shadow_leg(torch, ctx, args.shadow_cpis, args.shadow_steps, 2)
    if args.large_steps and sc.name != "large":
        torch.cuda.empty_cache()
        out["large"] = large_leg(torch, ctx, args.large_steps, bool(args.large_verify), max(2, args.stage_reps // 10))
    print(json.dumps(out), flush=True)
    for g in graphs + graphs_full:
        g.close()
    ctx.close()


if __name__ == "__main__":
    main()
