#!/usr/bin/env python3
"""Throughput of the bit-exact PFB -> FFT -> align -> beamformer/power ->
covariance -> CFAR -> event-packing chain on one MI355X: the split-int8 MFMA
path (batched where there are many CPIs) against the plain GPU path (no matrix
cores, one CPI per launch) and the CPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload medium]

A step is one pass over a ring of distinct input CPIs of the workload
(default 32 medium CPIs, 537 MB, more than the 256 MB last-level cache), each
CPI one HIP graph launch of the whole chain, the CPIs of a pass dealt
round-robin to --streams contexts (one HIP stream each) so that every input
CPI is read once per step. Inputs are resident in HBM before the timed region.
One JSON line is printed:

  value       whole-job Msamples/s (one sample = one complex int16 input sample
              of one channel), wall clock over exactly K steps
  roofline    the dominant kernel of the chain (largest mean launch time, HIP
              events around each stage launch on the library's stream over the
              ring's distinct CPIs): the fused chain's algorithmic bytes per CPI
              (input, taps, weights, Hermitian covariance, events) over that
              kernel's mean launch time, against 8 TB/s; `traffic` is the
              kernel's HBM bytes per launch from the committed rocprofv3 counter
              summary, `valu_frac_chain` the chain's VALU issue fraction
  plain_gpu   the same chain with the plain (no matrix cores) beamformer and
              covariance and a stand-alone alignment, one CPI per launch
  cpu_baseline  the reference's C++ models (oracle/_ref) or the CPU
              restatement, on this host's cores, on a bounded sample of CPIs
  mismatches  bit-exact comparison of one CPI's power, covariance and events
              against the CPU run (must be 0)
  large / batch / shadow  the large workload (MFMA vs plain), the 256-CPI
              uneven batch (one graph launch vs per-CPI launches, MFMA and
              plain) and the FP32 shadow
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
INT8_PEAK_OPS = 5.0e15
VALU_SIMDS = 1024  # 256 CUs x 4 SIMDs
VALU_CLOCK_HZ = 2.4e9
SPIN_CYCLES = 20_000_000  # shader-clock cycles of the spin ahead of each timed group (about 8 ms)
PLAIN_STAGES = ("pfb", "fft", "align", "bf_plain", "cov_plain", "cfar")


def fused(sc) -> bool:
    """The chain runs the beamformer and the covariance as one kernel (at most
    16 channels, bins a multiple of 4), as rd_chain_enqueue decides."""
    return sc.CH <= 16 and sc.N % 4 == 0 and sc.B <= 128 and 16 * sc.F * sc.N * 4 < 2 ** 32


def x_layout(sc) -> int:
    """The chain's X layout: bin-quad blocked (1) where it fuses the
    beamformer and the covariance, which then read X as contiguous runs."""
    return 1 if fused(sc) else 0


def chain_stages(sc):
    """The stages of the chain as it runs the workload, in order."""
    return ("pfb", "fft", "bfcov", "cfar") if fused(sc) else ("pfb", "fft", "bf", "cov", "cfar")


def kernel_stages(sc):
    """The stages that are one kernel per launch (CFAR is rows + scan + gather)."""
    return tuple(k for k in chain_stages(sc) if k != "cfar")


def note(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algo_chain_bytes(sc, n_events: int) -> int:
    """Compulsory traffic of the fused chain for one CPI: I/Q in, taps, weights,
    Hermitian covariance out (int64 re+im), 176-bit events (22 B)."""
    return (4 * sc.CH * sc.N * sc.F + 4 * sc.N * sc.T + 4 * sc.B * sc.CH + 16 * sc.N * sc.CH * (sc.CH + 1) // 2
            + 22 * n_events)


def stage_launchers(ctx, sc, plain=False):
    """The chain's stage launches on the library stream, as the lean chain
    graph runs them (no PFB/FFT flags, no beam samples): each takes the chain
    whose buffers it uses and the input CPI's device pointer. With each
    kernel's own algorithmic bytes per launch (its compulsory reads + writes)."""
    L, h = ctx.L, ctx.h
    S = sc.CH * sc.N * sc.F
    cols = sc.N * sc.F
    cfg = rd_amd.CfarCfg.of(sc.cfar_cfg())
    chk = ctx.lib.check

    def pfb(ch, iq):
        b = ch.b
        chk("pfb", L.rd_pfb_run(h, iq, b["taps"].ptr, sc.N, sc.T, sc.CH, sc.F, None, b["u"].ptr, None))

    xl = x_layout(sc)

    def fft(ch, iq):  # the chain's FFT stage, X in the chain's layout
        b = ch.b
        chk("fft", L.rd_fft_stage(h, b["u"].ptr, sc.N, 2, 0xFFFFFFFF, sc.CH, sc.F, xl if not plain else 0,
                                  b["X"].ptr, None))

    def bf(ch, iq):  # FFT layout in, power + aligned copy out
        b = ch.b
        chk("bf", L.rd_bf_stage(h, b["X"].ptr, b["weights"].ptr, sc.CH, sc.B, sc.F, sc.N, None, None, b["power"].ptr,
                                b["Xa"].ptr))

    def cov(ch, iq):
        chk("cov", L.rd_cov_run(h, ch.b["Xa"].ptr, sc.CH, sc.F, sc.N, ch.b["R"].ptr))

    def bfcov(ch, iq):  # FFT layout in, power + covariance out, no aligned copy
        b = ch.b
        chk("bfcov", L.rd_bf_cov_stage(h, b["X"].ptr, b["weights"].ptr, sc.CH, sc.B, sc.F, sc.N, None, None,
                                       b["power"].ptr, None, b["R"].ptr, xl))

    def align(ch, iq):
        chk("align", L.rd_align_run(h, ch.b["X"].ptr, sc.CH, sc.F, sc.N, ch.b["Xa"].ptr))

    def bf_plain(ch, iq):
        b = ch.b
        chk("bf_plain", L.rd_bf_plain_run(h, b["Xa"].ptr, b["weights"].ptr, sc.CH, sc.B, cols, None, None,
                                          b["power"].ptr))

    def cov_plain(ch, iq):
        chk("cov_plain", L.rd_cov_plain_run(h, ch.b["Xa"].ptr, sc.CH, sc.F, sc.N, ch.b["R"].ptr))

    def cfar(ch, iq):  # rows, scan and gather; the total stays on the device
        b = ch.b
        chk("cfar", L.rd_cfar_enqueue(h, b["power"].ptr, 4, sc.N, sc.B * sc.F, rd_amd.C.byref(cfg),
                                      rd_scenario.MAX_GUARD, rd_scenario.MAX_REF, None, sc.F, b["events"].ptr,
                                      ch.events_cap, b["row_count"].ptr, ch.total.ptr))

    cov_bytes = S * 4 + 16 * sc.N * sc.CH * (sc.CH + 1) // 2
    bf_bytes = S * 4 + 4 * sc.B * sc.CH + sc.B * cols * 4
    table = {
        "pfb": (pfb, S * 8 + 4 * sc.N * sc.T),
        "fft": (fft, S * 8),
        "bf": (bf, bf_bytes + S * 4),
        "cov": (cov, cov_bytes),
        "bfcov": (bfcov, bf_bytes + 16 * sc.N * sc.CH * (sc.CH + 1) // 2),
        "align": (align, S * 8),
        "bf_plain": (bf_plain, bf_bytes),
        "cov_plain": (cov_plain, cov_bytes),
        "cfar": (cfar, sc.B * cols * 4),
    }
    return {k: table[k] for k in (PLAIN_STAGES if plain else chain_stages(sc))}


def slot_chains(ctx, sc, taps, W, n):
    """n lean chains of the workload (no flags, no beam samples), one per ring
    slot, so every stage launch of the per-stage timing reads inputs that no
    launch just before it touched."""
    out = []
    for _ in range(n):
        c = rd_amd.Chain(ctx, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg(), rd_scenario.MAX_GUARD,
                         rd_scenario.MAX_REF, keep_y=False, keep_flags=False, keep_R=True,
                         events_cap=sc.B * sc.F * 64, x_layout=x_layout(sc))
        c.b["taps"].upload(taps)
        c.b["weights"].upload(W)
        out.append(c)
    return out


def ring_stage_times(torch, ctx, chains, sc, iq_ptrs, plain=False, passes=1):
    """Mean launch duration (ms) of every stage: for each stage in chain order,
    HIP events on the library's stream around its launches for every ring CPI
    back to back (each CPI with its own chain buffers, so the stage reads what
    the previous stage wrote for that CPI, a whole ring's stage ago), divided
    by the number of launches; `passes` times over the ring."""
    stages = stage_launchers(ctx, sc, plain)
    stream = torch.cuda.ExternalStream(ctx.stream)
    acc = {n: 0.0 for n in stages}
    for fn, _ in stages.values():  # warm every kernel once
        fn(chains[0], iq_ptrs[0])
    ctx.sync()
    for _ in range(passes):
        evs = []
        for n, (fn, _) in stages.items():
            # a bounded clock spin first, so that every launch of the group is
            # queued before the GPU reaches it: the events then time the
            # kernels back to back, not the host's launch rate
            with torch.cuda.stream(stream):
                torch.cuda._sleep(SPIN_CYCLES)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for c, p in zip(chains, iq_ptrs):
                fn(c, p)
            e1.record(stream)
            evs.append((n, e0, e1))
            ctx.sync()
        for n, e0, e1 in evs:
            acc[n] += e0.elapsed_time(e1) / len(iq_ptrs)
    out = {}
    for n, (_, nbytes) in stages.items():
        ms = acc[n] / passes
        out[n] = dict(ms=round(ms, 5), algo_bytes=nbytes, gbs=round(nbytes / (ms * 1e-3) / 1e9, 1))
    return out


def profiled(stage: str, workload: str):
    """Counters per dispatch of a stage kernel (HBM bytes, VALU and MFMA wave
    instructions) from the committed rocprofv3 counter summary of that
    workload, `profiles/chain_<workload>_counters.md` (None when absent)."""
    path = ROOT / "profiles" / f"chain_{workload}_counters.md"
    if not path.exists():
        return None
    head = None
    for line in path.read_text().splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) < 4:
            continue
        if cells[0] == "stage":
            head = cells
            continue
        if head and cells[0] == stage:
            row = dict(zip(head, cells))
            try:
                return dict(hbm_bytes=int(row["HBM bytes"]), valu_insts=int(row["SQ_INSTS_VALU"]),
                            mfma_insts=int(row.get("SQ_INSTS_MFMA", "0") or 0), avg_us=float(row["avg us"]))
            except (KeyError, ValueError):
                return None
    return None


def valu_frac(valu_insts: float, ms: float) -> float:
    """Fraction of the VALU issue rate a launch used: wave instructions x 4
    cycles over 1024 SIMDs x 2.4 GHz x the launch time."""
    return round(valu_insts * 4 / (VALU_SIMDS * VALU_CLOCK_HZ * ms * 1e-3), 4)


def mfma_ops(sc):
    """int8 MFMA work of the beamformer and covariance per CPI as issued: four
    digit products x (re, im) GEMMs x 2 ops per MAC over the padded K."""
    pairs = sc.CH * (sc.CH + 1) // 2
    cols = sc.N * sc.F
    ops = {"bf": 32 * sc.B * sc.CH * cols, "cov": 32 * pairs * sc.F * sc.N}
    ops["bfcov"] = ops["bf"] + ops["cov"]
    return {k: v for k, v in ops.items() if k in chain_stages(sc)}


def bf_cov_ms(per) -> float:
    """Time of the beamformer and covariance of a CPI (fused or not)."""
    return per["bfcov"]["ms"] if "bfcov" in per else per["bf"]["ms"] + per["cov"]["ms"]


def lean_bufs(ch, iq_ptr=None):
    b = rd_amd.CpiBufs.from_buffer_copy(ch.bufs)
    if iq_ptr is not None:
        b.iq = iq_ptr
    b.pfb_flags = b.fft_flags = b.y = b.yflags = b.Xa = None
    return b


def check_cpi(ch, n_ev, ref):
    """Mismatching values of one CPI's power, covariance and events."""
    got = ch.fetch(("power", "R"), n_events=n_ev)
    mism = int(np.count_nonzero(got["power"] != ref["power"]) + np.count_nonzero(got["R"] != ref["R"]))
    if n_ev != ref["n_events"]:
        mism += abs(n_ev - ref["n_events"]) + 1
    else:
        mism += int(np.count_nonzero(got["events"] != ref["events"]))
    return mism


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
def large_leg(torch, ctx, steps: int, verify: bool):
    """The large workload (64 ch x 4096-pt x 16 taps x 512 frames, 64 beams):
    the MFMA chain as graph replays over two device-generated CPIs against the
    plain chain (no matrix cores) enqueued per CPI on the same CPIs, per-stage
    times of both, the beamformer's and covariance's int8 MFMA rate as issued,
    one CPI checked against the oracle on the host's cores and its covariance
    trace checked in FP64 on the GPU."""
    sc = rd_scenario.config("large")
    taps, W = sc.taps(), sc.weights()
    ch = rd_amd.Chain(ctx, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg(), rd_scenario.MAX_GUARD, rd_scenario.MAX_REF,
                      keep_y=False, keep_flags=False, keep_R=True)
    note("large: generating 2 CPIs on the device")
    ring = [rd_workloads.device_iq(torch, sc, 7001 + i) for i in range(2)]
    iq0 = ring[0].cpu().numpy().view(np.uint32)
    ch.load(iq0, taps, W)
    totals = ctx.buf(32)
    graphs = [ch.graph(iq_ptr=ring[i].data_ptr(), total_ptr=totals.ptr + 8 * i, lean=True) for i in range(2)]
    for g in graphs:
        g.launch()
    ctx.sync()

    def wall(fn, n):
        ctx.sync()
        t0 = time.perf_counter()
        for i in range(n):
            fn(i)
        ctx.sync()
        return (time.perf_counter() - t0) * 1e3 / n

    ms = wall(lambda i: graphs[i % 2].launch(), steps)
    # the plain chain, one CPI per launch sequence, on the same CPIs
    pbufs = [lean_bufs(ch, ring[i].data_ptr()) for i in range(2)]

    def plain(i):
        ctx.lib.check("rd_chain_enqueue_plain", ctx.L.rd_chain_enqueue_plain(
            ctx.h, rd_amd.C.byref(ch.desc), rd_amd.C.byref(pbufs[i % 2]), totals.ptr + 16 + 8 * (i % 2)))

    plain(0)
    plain(1)
    plain_ms = wall(plain, max(2, steps // 2))
    tot = totals.download(np.uint64, 4)
    if not np.array_equal(tot[:2], tot[2:]):
        sys.exit(f"large: plain and MFMA event totals differ: {tot}")
    note("large: per-stage kernel timing")
    ptrs = [r.data_ptr() for r in ring]
    sch = slot_chains(ctx, sc, taps, W, len(ptrs))
    per = ring_stage_times(torch, ctx, sch, sc, ptrs, passes=2)
    per_plain = ring_stage_times(torch, ctx, sch, sc, ptrs, plain=True)
    del sch
    for k, ops in mfma_ops(sc).items():
        per[k]["mfma_int8_tops_as_issued"] = round(ops / (per[k]["ms"] * 1e-3) / 1e12, 1)
        per[k]["mfma_frac_of_5pops_as_issued"] = round(ops / (per[k]["ms"] * 1e-3) / INT8_PEAK_OPS, 4)
    for k in chain_stages(sc):
        prof = profiled(k, "large")
        if prof:
            per[k]["valu_frac"] = valu_frac(prof["valu_insts"], per[k]["ms"])
            if prof["mfma_insts"]:
                per[k]["mfma_wave_insts"] = prof["mfma_insts"]
    for g in graphs:
        g.close()
    res = dict(workload=f"large: {sc.CH} ch x {sc.N}-pt PFB/FFT x {sc.T} taps x {sc.F} frames, {sc.B} beams, "
                        f"CA-CFAR {sc.ref}/{sc.guard}",
               cpi_samples=sc.samples, steps=steps, ms_per_cpi=round(ms, 4),
               msamples_s=round(sc.samples / (ms * 1e-3) / 1e6, 1),
               plain_gpu=dict(ms_per_cpi=round(plain_ms, 4), msamples_s=round(sc.samples / (plain_ms * 1e-3) / 1e6, 1),
                              impl="rd_chain_enqueue_plain per CPI: stand-alone alignment, VALU beamformer and "
                                   "covariance (int64 sums), same PFB/FFT/CFAR", stages=per_plain),
               mfma_split_over_plain=round(plain_ms / ms, 3),
               bf_cov_mfma_over_plain=round((per_plain["align"]["ms"] + per_plain["bf_plain"]["ms"]
                                             + per_plain["cov_plain"]["ms"]) / bf_cov_ms(per), 3),
               stages=per, mismatches=None)
    if verify:
        import oracle

        note("large: checking one CPI against the oracle on the host")
        ch.enqueue()
        ctx.sync()
        n_ev = int(ch.total.download(np.uint64, 1)[0])
        threads = max(1, min(16, os.cpu_count() or 1))
        t0 = time.perf_counter()
        ref = oracle.load("oracle").chain(iq0, taps, W, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg(),
                                          rd_scenario.MAX_GUARD, rd_scenario.MAX_REF, keep=("power", "R", "events"),
                                          threads=threads)
        cpu_s = time.perf_counter() - t0
        res.update(mismatches=check_cpi(ch, n_ev, ref), events_per_cpi=n_ev,
                   oracle_cpu=dict(seconds=round(cpu_s, 2), threads=threads,
                                   msamples_s=round(sc.samples / cpu_s / 1e6, 2), kind="port"))
        f = ch.fetch(("Xa", "R"))
        res["covariance_trace_fp64"] = rd_workloads.covariance_trace_check(torch, f["Xa"], f["R"], sc.CH, sc.F, sc.N)
    del ring
    return res


# ----------------------------------------------------------------------------- batched leg
def batched_leg(torch, ctx, n_cpi, reps, autotune, oracle_cpis):
    """256 uneven CPIs: one batched graph launch of the MFMA chain vs one CPI
    per launch (MFMA chain, and the plain chain without matrix cores), checked
    against each other on every CPI and against the CPU on a sample."""
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

    plain_totals = torch.zeros(n_cpi, dtype=torch.int64, device="cuda")
    def lean():  # the current buffers (event buffers may have been regrown) without the optional outputs
        out = []
        for b in bs.bufs:
            lb = rd_amd.CpiBufs.from_buffer_copy(b)
            lb.pfb_flags = lb.fft_flags = lb.y = lb.yflags = None
            out.append(lb)
        return out

    def plain_gpu():  # no matrix cores, one CPI per launch sequence
        for i, (d, b) in enumerate(zip(bs.descs, plain_bufs)):
            ctx.lib.check("rd_chain_enqueue_plain", L.rd_chain_enqueue_plain(h, rd_amd.C.byref(d), rd_amd.C.byref(b),
                                                                              plain_totals.data_ptr() + 8 * i))
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
    note("batch: plain GPU (no matrix cores) one-CPI-per-launch passes")
    plain_bufs = lean()
    t_plain_gpu = wall(plain_gpu, max(1, reps // 2))
    if not torch.equal(plain_totals, totals_dev):
        sys.exit("batch: the plain chain's event totals differ from the MFMA chain's")
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
        "mfma_per_cpi_async": {"ms": round(t_async * 1e3, 3), "msamples_s": round(samples / t_async / 1e6, 1),
                        "impl": "rd_chain_enqueue per CPI on one stream, one sync at the end"},
        "mfma_per_cpi_sync": {"ms": round(t_sync * 1e3, 3), "msamples_s": round(samples / t_sync / 1e6, 1),
                       "impl": "rd_chain_run per CPI (event count returned to the host each CPI)"},
        "plain_gpu": {"ms": round(t_plain_gpu * 1e3, 3), "msamples_s": round(samples / t_plain_gpu / 1e6, 1),
                      "impl": "rd_chain_enqueue_plain per CPI (no matrix cores), one sync at the end"},
        "speedup_vs_plain_gpu": round(t_plain_gpu / t_batch, 3),
        "speedup_vs_mfma_per_cpi_async": round(t_async / t_batch, 3),
        "speedup_vs_mfma_per_cpi_sync": round(t_sync / t_batch, 3),
        "autotune": bool(autotune), "create_s": round(t_create, 2), "plan": plan,
        "mismatching_cpis": bad_cpis, "oracle_cpis": orc_n, "oracle_mismatches": orc_bad,
        "events": int(np.sum(totals)),
    }


# ----------------------------------------------------------------------------- FP32 shadow leg
def shadow_leg(torch, ctx, n_cpi, steps, fd_dirs, n_train=None, pos_weight=512.0, lr=3e-3):
    """The differentiable FP32 shadow on 64 CPIs of 32 ch x 1024-pt x 128
    frames (rd_workloads.ShadowProblem): one training step (forward + backward
    over every CPI) timed, gradients checked against float64 (plain autograd
    and central finite differences), the weights and per-channel gains trained
    on the first 3/4 of the CPIs, quantised to Q1.15 (gains folded in) and both
    weight sets rerun through the bit-exact batched chain, scored on the
    training and the held-out CPIs."""
    import rd_shadow as sh

    n_train = n_cpi * 3 // 4 if n_train is None else n_train
    note(f"shadow: generating {n_cpi} CPIs (32 ch x 1024 x 128, channel gain/phase errors)")
    pb = rd_workloads.ShadowProblem(torch, n_cpi, n_train, pos_weight=pos_weight)
    W = pb.W0.clone().requires_grad_()
    gain = torch.ones(pb.C, dtype=torch.complex64, device="cuda").requires_grad_()
    for _ in range(2):
        pb.loss32(W, gain).backward()
    torch.cuda.synchronize()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        W.grad = None
        gain.grad = None
        pb.loss32(W, gain).backward()
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t0) / reps
    note("shadow: float64 gradient checks")
    autograd_rel, fd_rel = pb.gradient_check(pb.W0, torch.ones(pb.C, dtype=torch.complex64, device="cuda"), fd_dirs)
    note(f"shadow: {steps} training steps on {pb.n_train} CPIs")
    Wt, gt, losses = pb.train(steps, lr)
    note("shadow: bit-exact rerun with the quantised weights")
    Wq = sh.quantise_weights(Wt, gt)
    with torch.no_grad():  # the FP32 loss at the quantised weights (gains folded in, unit gains)
        ones = torch.ones(pb.C, dtype=torch.complex64, device="cuda")
        held = slice(pb.n_train, pb.n_cpi)
        loss_q = {"initial": float(pb.loss32(pb.W0, ones, held)),
                  "trained": float(pb.loss32(pb.words_to_weights(Wq), ones, held))}
    hits, consistency = {}, 0
    for name, wq in (("initial", pb.W0_w), ("trained", Wq)):
        hits[name], c = pb.bit_exact_hits(ctx, wq)
        consistency += c
    samples = pb.samples
    del pb
    return {
        "workload": f"{n_cpi} CPIs x 32 ch x 1024-pt x 128 frames, FP32, 32x32 complex weights + 32 channel gains, "
                    f"soft CA-CFAR sigmoid((p/mean - 8.0)/0.1), weighted BCE (targets x {pos_weight:g}) against "
                    f"target cells, sidelobes and noise; trained on {n_train} CPIs, {n_cpi - n_train} held out",
        "samples": samples,
        "train_step_ms": round(t_step * 1e3, 2),
        "train_step_msamples_s": round(samples / t_step / 1e6, 1),
        "grad_rel_err_vs_fp64_autograd": autograd_rel,
        "grad_rel_err_vs_fp64_finite_diff": fd_rel,
        "train_steps": steps, "lr": lr,
        "loss_first_last": [losses[0], losses[-1]] if losses else None,
        "held_out_loss_fp32_at_q15_weights": loss_q,
        "bit_exact_rerun": hits,
        "cfar_rule_vs_row_counts_mismatches": consistency,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed steps (passes over the ring)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="medium", choices=["tiny", "medium", "large"])
    ap.add_argument("--ring", type=int, default=32, help="distinct input CPIs; one step is one pass over them")
    ap.add_argument("--streams", type=int, default=3,
                    help="contexts (one HIP stream each) the CPIs of a pass are dealt to round-robin")
    ap.add_argument("--cpu-seconds", type=float, default=3.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--stage-passes", type=int, default=2, help="passes over the ring for the per-stage timing")
    ap.add_argument("--batch-cpis", type=int, default=256, help="CPIs of the uneven batch leg (0: skip it)")
    ap.add_argument("--batch-reps", type=int, default=3)
    ap.add_argument("--batch-oracle", type=int, default=3, help="smallest batch CPIs also checked on the CPU")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--batch-only", action="store_true", help="run only the batched leg")
    ap.add_argument("--shadow-cpis", type=int, default=64, help="CPIs of the FP32 shadow leg (0: skip it)")
    ap.add_argument("--shadow-steps", type=int, default=60)
    ap.add_argument("--shadow-only", action="store_true", help="run only the FP32 shadow leg")
    ap.add_argument("--large-steps", type=int, default=10, help="graph replays of the large-workload leg (0: skip it)")
    ap.add_argument("--large-verify", type=int, default=1, help="check one large CPI against the oracle (0: skip)")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-GPU timing of the main workload")
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
    R = args.ring
    n_str = max(1, min(args.streams, R))
    ch = rd_amd.Chain(ctx, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg(), rd_scenario.MAX_GUARD, rd_scenario.MAX_REF,
                      keep_y=True, keep_flags=True, keep_R=True, x_layout=x_layout(sc))
    note(f"workload {sc.name}: generating {R} input CPIs")
    ch.b["taps"].upload(taps)
    ch.b["weights"].upload(W)
    if sc.samples > 50_000_000:  # large CPIs are generated on the device (same construction as Scene.iq)
        ring = [rd_workloads.device_iq(torch, sc, 1000 + i) for i in range(R)]
        ring_ptrs = [t.data_ptr() for t in ring]
        iq0 = None if args.no_cpu else ring[0].cpu().numpy().view(np.uint32)
    else:
        ring = [ctx.put(sc.iq(seed=1000 + i)) for i in range(R)]
        ring_ptrs = [r.ptr for r in ring]
        iq0 = sc.iq(seed=1000)
    totals = ctx.buf(8 * R)

    # correctness of one CPI against the CPU run (also the timed CPU baseline)
    cpu = None
    if not args.no_cpu:
        note("CPU baseline")
        cpu, cpu_out = cpu_baseline(sc, iq0, taps, W, args.cpu_seconds)
    ch.enqueue(iq_ptr=ring_ptrs[0])
    ctx.sync()
    n_ev = int(ch.total.download(np.uint64, 1)[0])
    mismatches = check_cpi(ch, n_ev, cpu_out) if cpu is not None else None
    f = ch.fetch(("Xa", "R"))
    trace = rd_workloads.covariance_trace_check(torch, f["Xa"], f["R"], sc.CH, sc.F, sc.N)
    del f

    def timed(step, steps):
        """Wall clock of `steps` calls of step(), bracketed by a synchronisation
        of every context and of torch on both sides."""
        for c in ctxs:
            c.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        for c in ctxs:
            c.sync()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    # the headline: one lean graph per ring slot, slot j captured on context
    # j % n_str with its own chain buffers, so a step reads each input CPI once
    ctxs = [ctx] + [rd_amd.Ctx(ctx.lib) for _ in range(n_str - 1)]
    chains, s_totals = [ch], [totals]
    for c in ctxs[1:]:
        chk = rd_amd.Chain(c, sc.CH, sc.N, sc.T, sc.F, sc.B, sc.cfar_cfg(), rd_scenario.MAX_GUARD,
                           rd_scenario.MAX_REF, keep_y=False, keep_flags=False, keep_R=True, x_layout=x_layout(sc))
        chk.b["taps"].upload(taps)
        chk.b["weights"].upload(W)
        chains.append(chk)
        s_totals.append(c.buf(8 * R))
    note(f"capturing {R} chain graphs over {n_str} streams")
    slot_graphs = [chains[j % n_str].graph(iq_ptr=ring_ptrs[j], total_ptr=s_totals[j % n_str].ptr + 8 * j, lean=True)
                   for j in range(R)]

    def step_streams():
        for g in slot_graphs:
            g.launch()

    note(f"warmup {args.warmup} + {args.steps} timed steps of {R} CPIs (graph replay over {n_str} streams)")
    timed(step_streams, args.warmup)
    dt = timed(step_streams, args.steps)
    ms_per_step = dt * 1e3 / args.steps
    ms_per_cpi = ms_per_step / R
    value = args.steps * R * sc.samples / dt / 1e6
    slot_tot = np.array([s_totals[j % n_str].download(np.uint64, R)[j] for j in range(R)], np.uint64)
    for g in slot_graphs:
        g.close()
    for c in ctxs[1:]:
        c.close()
    ctxs = [ctx]

    # one stream: lean graphs, full-output graphs, lean stream enqueue
    note("one-stream comparisons: lean graphs, full-output graphs, stream enqueue")
    g1 = [ch.graph(iq_ptr=ring_ptrs[j], total_ptr=totals.ptr + 8 * j, lean=True) for j in range(R)]
    timed(lambda: [g.launch() for g in g1], 1)
    t_graph = timed(lambda: [g.launch() for g in g1], args.steps)
    one_tot = totals.download(np.uint64, R)
    if not np.array_equal(one_tot, slot_tot) or not np.all(one_tot > 0):
        sys.exit("event totals differ between the streams' contexts and the one-stream run")
    for g in g1:
        g.close()
    gf = [ch.graph(iq_ptr=ring_ptrs[j], total_ptr=totals.ptr + 8 * j) for j in range(R)]
    timed(lambda: [g.launch() for g in gf], 1)
    t_full = timed(lambda: [g.launch() for g in gf], args.steps)
    for g in gf:
        g.close()
    bufs_l = [lean_bufs(ch, p) for p in ring_ptrs]

    natural = rd_amd.CpiDesc.from_buffer_copy(ch.desc)
    natural.x_layout = 0  # the plain chain keeps X in natural order

    def enqueue_pass(fn):
        d = natural if fn == "rd_chain_enqueue_plain" else ch.desc
        for j in range(R):
            ctx.lib.check(fn, getattr(ctx.L, fn)(ctx.h, rd_amd.C.byref(d), rd_amd.C.byref(bufs_l[j]),
                                                 totals.ptr + 8 * j))

    timed(lambda: enqueue_pass("rd_chain_enqueue"), 1)
    t_stream = timed(lambda: enqueue_pass("rd_chain_enqueue"), args.steps)
    plain = None
    if not args.no_plain:
        note("plain GPU chain (no matrix cores), one CPI per launch")
        timed(lambda: enqueue_pass("rd_chain_enqueue_plain"), 1)
        if not np.array_equal(totals.download(np.uint64, R), one_tot):
            sys.exit("the plain chain's event totals differ from the MFMA chain's")
        t_plain = timed(lambda: enqueue_pass("rd_chain_enqueue_plain"), max(1, args.steps // 2))
        plain_ms = t_plain * 1e3 / max(1, args.steps // 2) / R
        plain = dict(ms_per_cpi=round(plain_ms, 5), msamples_s=round(sc.samples / (plain_ms * 1e-3) / 1e6, 1),
                     impl="rd_chain_enqueue_plain per CPI on one stream over the ring: PFB, FFT, stand-alone "
                          "alignment, VALU beamformer + power and covariance (int64 sums), CFAR",
                     mfma_split_over_plain=round(plain_ms / ms_per_cpi, 3),
                     mfma_over_plain_same_launch=round(t_plain / max(1, args.steps // 2) / (t_stream / args.steps),
                                                       3))

    # per-kernel launch durations over the ring's distinct CPIs
    note("per-stage kernel timing over the ring")
    sch = slot_chains(ctx, sc, taps, W, R)
    per = ring_stage_times(torch, ctx, sch, sc, ring_ptrs, passes=args.stage_passes)
    if plain is not None:
        plain["stages"] = ring_stage_times(torch, ctx, sch, sc, ring_ptrs, plain=True, passes=1)
    del sch
    for k, ops in mfma_ops(sc).items():
        per[k]["mfma_int8_tops_as_issued"] = round(ops / (per[k]["ms"] * 1e-3) / 1e12, 2)
        per[k]["mfma_frac_of_5pops_as_issued"] = round(ops / (per[k]["ms"] * 1e-3) / INT8_PEAK_OPS, 4)
    chain_valu = 0
    profs = {k: profiled(k, sc.name) for k in chain_stages(sc)}
    for k, prof in profs.items():
        if prof:
            per[k]["valu_frac"] = valu_frac(prof["valu_insts"], per[k]["ms"])
            per[k]["hbm_bytes_counted"] = prof["hbm_bytes"]
            if prof["mfma_insts"]:
                per[k]["mfma_wave_insts"] = prof["mfma_insts"]
            chain_valu += prof["valu_insts"]
    dom = max(kernel_stages(sc), key=lambda k: per[k]["ms"])
    chain_bytes = algo_chain_bytes(sc, n_ev)
    ach = chain_bytes / (per[dom]["ms"] * 1e-3) / 1e9
    prof = profs.get(dom)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": prof["hbm_bytes"] if prof else None,
                "algo_bytes_per_launch": chain_bytes, "kernel_ms": per[dom]["ms"],
                "kernel_own_bytes": per[dom]["algo_bytes"],
                "kernel_own_frac": round(per[dom]["gbs"] / HBM_PEAK_GBS, 4),
                "valu_frac_kernel": per[dom].get("valu_frac"),
                "valu_frac_chain": (valu_frac(chain_valu, ms_per_cpi) if all(profs.values()) else None),
                "chain_valu_wave_insts_per_cpi": chain_valu or None}
    out = {
        "metric": "bit-exact Msamples/s PFB→CFAR: MFMA-split+batched vs plain GPU vs CPU",
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
                   "cpis_per_step": R, "ms_per_cpi": round(ms_per_cpi, 5), "streams": n_str,
                   "impl": "one lean hipGraph per CPI (PFB, FFT, "
                           + ("beamformer+power and covariance as one split-int8 MFMA kernel reading X once, "
                              "X kept bin-quad blocked by the FFT"
                              if fused(sc) else "BF+power split-int8 MFMA writing the aligned copy, covariance "
                                                "split-int8 MFMA")
                           + f", CFAR rows/scan/gather); a step is one pass over {R} distinct input CPIs dealt "
                             f"round-robin to {n_str} contexts (one HIP stream each)",
                   "outputs": "events + covariance timed; every stage output verified bit-exact on a separate "
                              "CPI each run (timing.graph_full_outputs also writes them all)"},
        "roofline": roofline,
        "chain": {"algo_bytes_per_cpi": chain_bytes,
                  "achieved_gbs": round(chain_bytes / (ms_per_cpi * 1e-3) / 1e9, 1),
                  "frac": round(chain_bytes / (ms_per_cpi * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "plain_gpu": plain,
        "timing": {"graph_one_stream": {"ms_per_cpi": round(t_graph * 1e3 / args.steps / R, 5),
                                        "msamples_s": round(args.steps * R * sc.samples / t_graph / 1e6, 1)},
                   "graph_full_outputs": {"ms_per_cpi": round(t_full * 1e3 / args.steps / R, 5),
                                          "msamples_s": round(args.steps * R * sc.samples / t_full / 1e6, 1)},
                   "stream_enqueue": {"ms_per_cpi": round(t_stream * 1e3 / args.steps / R, 5),
                                      "msamples_s": round(args.steps * R * sc.samples / t_stream / 1e6, 1)}},
        "stages": per,
        "cpu_baseline": cpu,
        "mismatches": mismatches,
        "covariance_trace_fp64": trace,
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
        del src, dst
    del ring
    if args.batch_cpis:
        out["batch"] = batched_leg(torch, ctx, args.batch_cpis, args.batch_reps, not args.no_autotune,
                                   args.batch_oracle)
    if args.shadow_cpis:
        torch.cuda.empty_cache()
        out["shadow"] = shadow_leg(torch, ctx, args.shadow_cpis, args.shadow_steps, 2)
    if args.large_steps and sc.name != "large":
        torch.cuda.empty_cache()
        out["large"] = large_leg(torch, ctx, args.large_steps, bool(args.large_verify))
    out["three_way"] = {
        "medium_msamples_s": {"mfma_split": out["value"], "plain_gpu": plain["msamples_s"] if plain else None,
                              "cpu": cpu["value"] if cpu else None},
        "large_msamples_s": ({"mfma_split": out["large"]["msamples_s"],
                              "plain_gpu": out["large"]["plain_gpu"]["msamples_s"],
                              "cpu_oracle": out["large"].get("oracle_cpu", {}).get("msamples_s")}
                             if "large" in out else None),
        "large_mfma_split_over_plain": out["large"]["mfma_split_over_plain"] if "large" in out else None,
        "batch_msamples_s": ({"mfma_split_batched": out["batch"]["batched"]["msamples_s"],
                              "mfma_split_per_cpi": out["batch"]["mfma_per_cpi_async"]["msamples_s"],
                              "plain_gpu_per_cpi": out["batch"]["plain_gpu"]["msamples_s"]}
                             if "batch" in out else None),
        "batch_mfma_batched_over_plain_per_cpi": out["batch"]["speedup_vs_plain_gpu"] if "batch" in out else None,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
