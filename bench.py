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

HBM_PEAK_GBS = 8000.0


def note(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algo_chain_bytes(sc, n_events: int) -> int:
    """Compulsory traffic of the fused chain for one CPI: I/Q in, taps, weights,
    Hermitian covariance out (int64 re+im), 176-bit events (22 B)."""
    return (4 * sc.CH * sc.N * sc.F + 4 * sc.N * sc.T + 4 * sc.B * sc.CH + 16 * sc.N * sc.CH * (sc.CH + 1) // 2
            + 22 * n_events)


def stage_kernels(ctx, ch, sc):
    """Per-stage launchers (device buffers of `ch`), for the per-kernel timing."""
    L, h, b = ctx.L, ctx.h, ch.b
    S = sc.CH * sc.N * sc.F
    cols = sc.N * sc.F

    def pfb():
        ctx.lib.check("pfb", L.rd_pfb_run(h, b["iq"].ptr, b["taps"].ptr, sc.N, sc.T, sc.CH, sc.F, None, b["u"].ptr,
                                          b["pfb_flags"].ptr if b["pfb_flags"] else None))

    def fft():
        ctx.lib.check("fft", L.rd_fft_run(h, b["u"].ptr, sc.N, 2, 0xFFFFFFFF, 1, sc.CH * sc.F, None, b["X"].ptr,
                                          b["fft_flags"].ptr if b["fft_flags"] else None))

    def bf():
        # the chain's beamformer: FFT layout in, power + aligned copy out (lean outputs)
        ctx.lib.check("bf", L.rd_bf_stage(h, b["X"].ptr, b["weights"].ptr, sc.CH, sc.B, sc.F, sc.N, None, None,
                                          b["power"].ptr, b["Xa"].ptr))

    def cov():
        ctx.lib.check("cov", L.rd_cov_run(h, b["Xa"].ptr, sc.CH, sc.F, sc.N, b["R"].ptr))

    # algorithmic bytes per launch of each stage kernel (its compulsory reads + writes)
    flag_b = 1 if b["pfb_flags"] else 0
    stages = {
        "pfb": (pfb, S * (4 + 4 + flag_b) + 4 * sc.N * sc.T),
        "fft": (fft, S * 8 + 4 * sc.CH * sc.F),
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
    return dict(value=round(n * sc.samples / dt / 1e6, 3), unit="Msamples/s", cores=threads, kind=kind,
                sample=f"{n} CPI(s) of the {sc.name} workload ({sc.samples} samples each), full chain, "
                       f"{threads} threads, {dt:.1f} s"), out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=320)
    ap.add_argument("--warmup", type=int, default=32)
    ap.add_argument("--workload", default="medium", choices=["tiny", "medium", "large"])
    ap.add_argument("--ring", type=int, default=32, help="distinct input CPIs cycled through")
    ap.add_argument("--cpu-seconds", type=float, default=3.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--stage-reps", type=int, default=50)
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

    note(f"warmup {args.warmup} + {args.steps} timed steps (graph replay)")
    timed(run_graph, args.warmup)
    dt, host_dt, gpu_dt = timed(run_graph, args.steps)
    ms_per_step = dt * 1e3 / args.steps
    value = args.steps * sc.samples / dt / 1e6
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
    dom = max(per, key=lambda k: per[k]["ms"])
    ach = per[dom]["gbs"]
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
                   "ring": args.ring, "impl": "one hipGraph per CPI: PFB, FFT, align, BF+power MFMA, cov MFMA, "
                                              "CFAR count/scan/emit",
                   "outputs": "events + covariance timed; every stage output verified bit-exact on a separate "
                              "CPI each run (timing.graph_full_outputs also writes them all)"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None},
        "chain": {"algo_bytes_per_cpi": chain_bytes, "achieved_gbs": round(chain_bytes / (ms_per_step * 1e-3) / 1e9, 1),
                  "frac": round(chain_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "timing": {"graph": {"wall_ms": round(ms_per_step, 4), "host_ms": round(host_dt * 1e3 / args.steps, 4),
                             "gpu_ms": round(gpu_dt * 1e3 / args.steps, 4)},
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
    print(json.dumps(out), flush=True)
    for g in graphs + graphs_full:
        g.close()
    ctx.close()


if __name__ == "__main__":
    main()
