#!/usr/bin/env python3
"""Per-stage summary of rocprofv3 output for the chain's kernels: the table
bench.py reads back (`profiles/chain_<workload>_counters.md`).

    python rd_profsum.py --trace <dir> --pmc <dir> [--pmc <dir> ...] \\
        --stage 'fft=fft_kernel<10, 2, true, false>' --algo fft=33554432 ... \\
        --copy-mib 256 --title '...' --command '...' [--command '...']

--trace   a `rocprofv3 --kernel-trace --stats` output directory (average
          durations from its *kernel_stats.csv)
--pmc     `rocprofv3 --pmc ...` output directories (*counter_collection.csv);
          each counter is averaged over the dispatches of a kernel
--stage   stage=substring[+substring...]: the kernels of a stage (their
          counters and durations summed per launch); an optional @grid after a
          substring keeps only dispatches of that grid size (work-items)
--algo    stage=bytes: the stage's algorithmic bytes per launch
--copy-mib  size of the device-to-device copy (`__amd_rocclr_copyBuffer`) run
          in the same program: FETCH_SIZE and WRITE_SIZE are scaled by what
          they read for it, so HBM bytes = fscale * FETCH_SIZE KiB * 1024 +
          wscale * WRITE_SIZE KiB * 1024
MFMA columns: SQ_INSTS_VALU_MFMA_I8 (wave instructions of
v_mfma_i32_16x16x32_i8, 16384 int8 ops each) gives the int8 rate over the
traced average duration, against 5 POP/s; `MFMA busy` is
SQ_VALU_MFMA_BUSY_CYCLES / (1024 SIMDs x 2.4 GHz x the traced average
duration) (GRBM_GUI_ACTIVE, which rocprofv3's MfmaUtil divides by, spans the
whole counter-collection window of a dispatch rather than the kernel).
"""
from __future__ import annotations

import argparse
import csv
import glob
import gzip
import io
import os
from collections import defaultdict

INT8_OPS_PER_MFMA = 16 * 16 * 32 * 2
SIMDS = 1024


def _files(d, pattern):
    return sorted(glob.glob(os.path.join(d, "**", pattern), recursive=True) +
                  glob.glob(os.path.join(d, "**", pattern + ".gz"), recursive=True))


def _open(f):
    return io.TextIOWrapper(gzip.open(f), newline="") if f.endswith(".gz") else open(f, newline="")


def read_counters(dirs):
    """{(kernel name, grid size): {counter: mean per dispatch}, dispatches}."""
    per = defaultdict(lambda: defaultdict(float))  # dispatch -> counter -> value
    meta = {}
    for d in dirs:
        for f in _files(d, "*counter_collection.csv"):
            with _open(f) as fh:
                for row in csv.DictReader(fh):
                    key = (f, row["Dispatch_Id"])
                    per[key][row["Counter_Name"]] += float(row["Counter_Value"])
                    meta[key] = (row["Kernel_Name"], int(row["Grid_Size"]))
    groups = defaultdict(list)
    for key, vals in per.items():
        groups[meta[key]].append(vals)
    out = {}
    for k, lst in groups.items():
        names = set().union(*[v.keys() for v in lst])
        out[k] = ({n: sum(v.get(n, 0.0) for v in lst) / max(1, sum(1 for v in lst if n in v)) for n in names},
                  len(lst))
    return out


def read_stats(d, full=False):
    """{kernel name: (calls, average ns)} from *kernel_stats.csv (full: also
    min, max and total ns)."""
    out = {}
    for f in _files(d, "*kernel_stats.csv"):
        with _open(f) as fh:
            for row in csv.DictReader(fh):
                v = (int(row["Calls"]), float(row["AverageNs"]))
                if full:
                    v += (float(row["MinNs"]), float(row["MaxNs"]), float(row["TotalDurationNs"]))
                out[row["Name"]] = v
    return out


def list_table(d, title, commands):
    """The library's kernels of a kernel trace, by total time."""
    st = read_stats(d, full=True)
    rows = sorted(((n, v) for n, v in st.items() if "rd::" in n), key=lambda kv: -kv[1][4])
    lines = [f"# {title}", ""]
    if commands:
        lines += ["Command (run from `/tmp` with `TMPDIR=/tmp` on one MI355X; `<repo>` is the repository root, "
                  "`<out>` an output directory):", "", "```"] + commands + ["```", "",
                  "Per-kernel durations from that run's `kernel_stats.csv` (every dispatch of the run).", ""]
    lines += ["| kernel | calls | avg us | min us | max us | total ms |", "|---|---|---|---|---|---|"]
    for n, (c, avg, mn, mx, tot) in rows:
        lines.append(f"| `{short(n)}` | {c} | {avg / 1e3:.2f} | {mn / 1e3:.2f} | {mx / 1e3:.2f} | {tot / 1e6:.3f} |")
    return "\n".join(lines)


def short(name: str) -> str:
    n = name.split("(")[0].replace("void ", "").replace("rd::", "").replace("unsigned int", "u32")
    return n.replace("unsigned long", "u64").strip()


def match(spec: str, name: str, grid: int | None = None):
    sub, _, g = spec.partition("@")
    return sub in name and (not g or grid is None or int(g) == grid)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--trace", required=True)
    ap.add_argument("--pmc", action="append", default=[])
    ap.add_argument("--stage", action="append", default=[])
    ap.add_argument("--algo", action="append", default=[])
    ap.add_argument("--copy-mib", type=int, default=256)
    ap.add_argument("--title", default="")
    ap.add_argument("--command", action="append", default=[])
    ap.add_argument("--list", action="store_true", help="only the kernel-trace table of the library's kernels")
    a = ap.parse_args()
    if a.list:
        print(list_table(a.trace, a.title, a.command))
        return
    stats = read_stats(a.trace)
    ctr = read_counters(a.pmc)
    copy = [v for (n, g), (v, c) in ctr.items() if "copyBuffer" in n]
    fscale = wscale = None
    if copy:
        nbytes = a.copy_mib << 20
        fe = max(c.get("FETCH_SIZE", 0.0) for c in copy)
        wr = max(c.get("WRITE_SIZE", 0.0) for c in copy)
        fscale = nbytes / (fe * 1024) if fe else None
        wscale = nbytes / (wr * 1024) if wr else None
    stages = [s.split("=", 1) for s in a.stage]
    algo = dict(s.split("=", 1) for s in a.algo)
    cols = ["stage", "kernel", "dispatches", "avg us", "FETCH_SIZE KiB", "WRITE_SIZE KiB", "HBM bytes",
            "algorithmic bytes", "HBM / algorithmic", "SQ_INSTS_VALU", "SQ_INSTS_MFMA", "SQ_INSTS_VALU_MFMA_I8",
            "int8 MFMA TOP/s", "of 5 POP/s", "MFMA busy", "VALU issue"]
    lines = [f"# {a.title}", ""]
    if a.command:
        lines += ["Commands (run from `/tmp` with `TMPDIR=/tmp` on one MI355X; `<repo>` is the repository root, "
                  "`<out>` an output directory):", "", "```"] + a.command + ["```", ""]
    if fscale:
        lines += [f"HBM bytes per launch = {fscale:.3f} x FETCH_SIZE KiB x 1024 + {wscale:.3f} x WRITE_SIZE KiB x "
                  f"1024: the factors make the {a.copy_mib} MiB device copy in the same runs read and write "
                  f"{a.copy_mib} MiB. Counters are means per dispatch; `avg us` is the kernel-trace average. "
                  "A stage of several kernels sums them per launch. VALU issue = SQ_INSTS_VALU x 4 cycles / "
                  "(1024 SIMDs x 2.4 GHz x avg us); MFMA busy = SQ_VALU_MFMA_BUSY_CYCLES / (1024 SIMDs x 2.4 GHz x avg us); "
                  "int8 MFMA TOP/s = SQ_INSTS_VALU_MFMA_I8 x 16384 ops (one v_mfma_i32_16x16x32_i8) / avg us.", ""]
    lines += ["| " + " | ".join(cols) + " |", "|" + "---|" * len(cols)]

    def row(stage, label, parts):
        calls = min((c for _, c, _ in parts), default=0)
        avg_ns = sum(ns for _, _, ns in parts)
        tot = defaultdict(float)
        for vals, _, _ in parts:
            for k, v in vals.items():
                tot[k] += v
        fe, wr = tot.get("FETCH_SIZE"), tot.get("WRITE_SIZE")
        hbm = int(round(fscale * fe * 1024 + wscale * wr * 1024)) if fscale and fe is not None and wr is not None else None
        alg = int(algo[stage]) if stage in algo else None
        mf8 = tot.get("SQ_INSTS_VALU_MFMA_I8")
        tops = mf8 * INT8_OPS_PER_MFMA / (avg_ns * 1e-9) / 1e12 if mf8 and avg_ns else None
        busy = (tot["SQ_VALU_MFMA_BUSY_CYCLES"] / (SIMDS * 2.4e9 * avg_ns * 1e-9)
                if avg_ns and "SQ_VALU_MFMA_BUSY_CYCLES" in tot else None)
        valu = tot.get("SQ_INSTS_VALU")
        issue = valu * 4 / (SIMDS * 2.4e9 * avg_ns * 1e-9) if valu and avg_ns else None

        def f(v, fmt):
            return "-" if v is None else format(v, fmt)

        return "| " + " | ".join([stage, label, str(calls), f(avg_ns / 1e3 if avg_ns else None, ".2f"),
                                  f(fe, ".1f"), f(wr, ".1f"), f(hbm, "d"), f(alg, "d"),
                                  f(hbm / alg if hbm and alg else None, ".2f"), f(valu, ".0f"),
                                  f(tot.get("SQ_INSTS_MFMA"), ".0f"), f(mf8, ".0f"), f(tops, ".1f"),
                                  f(tops * 1e12 / 5e15 if tops else None, ".4f"), f(busy, ".4f"),
                                  f(issue, ".3f")]) + " |"

    for stage, spec in stages:
        parts, labels = [], []
        for sub in spec.split("+"):
            hits = [(k, v) for k, v in ctr.items() if match(sub, k[0], k[1])]
            if not hits:
                continue
            (name, grid), (vals, n) = max(hits, key=lambda kv: kv[1][1])
            calls, ns = next(((c, t) for nm, (c, t) in stats.items() if match(sub, nm)), (n, 0.0))
            parts.append((vals, calls, ns))
            labels.append(f"`{short(name)}`")
        if parts:
            lines.append(row(stage, " + ".join(labels), parts))
    print("\n".join(lines))


if __name__ == "__main__":
    main()
