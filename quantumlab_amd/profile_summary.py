"""Turn rocprofv3 CSV output into the short summaries kept under profiles/.

Usage (from the repository root, after the rocprofv3 runs named in the
summary headers):
    python3 -m quantumlab_amd.profile_summary --name brick32_fp64 \\
        --trace <trace dir> --fetch <FETCH_SIZE dir> --write <WRITE_SIZE dir> \\
        --out profiles

The rocprofv3 commands recorded in a summary write into the directories
given here (their -d argument).

FETCH_SIZE / WRITE_SIZE (KiB per dispatch) are calibrated against the one
kernel of known traffic in the run: kScale, the in-place renormalisation
sweep, which reads and writes every amplitude exactly once (16 B each for an
FP64 register). The calibrated factors are written with the summary.
"""
from __future__ import annotations

import argparse
import csv
import json
import os
import re
from collections import defaultdict


def short(name: str) -> str:
    name = re.sub(r"qmx::\(anonymous namespace\)::", "", name)
    name = re.sub(r"qmx::", "", name)
    return name.split("(")[0].replace("void ", "")


def trace_table(d: str, cmd: str) -> str:
    rows = list(csv.DictReader(open(os.path.join(d, "bench_kernel_stats.csv"))))
    out = [f"Command: `{cmd}`", "", "| kernel | calls | avg ms | min ms | max ms | % time |",
           "|---|---:|---:|---:|---:|---:|"]
    for r in rows:
        if float(r["Percentage"]) < 0.01:
            continue
        out.append(f"| {short(r['Name'])} | {r['Calls']} | {float(r['AverageNs']) / 1e6:.3f} | "
                   f"{float(r['MinNs']) / 1e6:.3f} | {float(r['MaxNs']) / 1e6:.3f} | {float(r['Percentage']):.2f} |")
    return "\n".join(out)


def counter_means(d: str):
    agg = defaultdict(list)
    for r in csv.DictReader(open(os.path.join(d, "bench_counter_collection.csv"))):
        agg[(short(r["Kernel_Name"]), int(r["Grid_Size"]))].append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in agg.items()}, {k: len(v) for k, v in agg.items()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--name", required=True)
    ap.add_argument("--trace", required=True)
    ap.add_argument("--fetch", default=None)
    ap.add_argument("--write", default=None)
    ap.add_argument("--out", default="profiles")
    ap.add_argument("--qubits", type=int, default=32)
    ap.add_argument("--trace-cmd", default=None, help="the rocprofv3 command of the trace (kernel-trace-only summaries)")
    args = ap.parse_args()
    if args.fetch is None:
        # kernel durations only
        os.makedirs(args.out, exist_ok=True)
        with open(os.path.join(args.out, f"{args.name}_kernel_trace.md"), "w") as f:
            f.write(f"# {args.name}: kernel durations (rocprofv3 kernel trace)\n\n")
            f.write(trace_table(args.trace, args.trace_cmd) + "\n")
        return
    # the state-vector legs only (per-gate FP64 and fused FP32), so the sweep
    # kernels' averages are those of the 32-qubit registers
    trace_cmd = (f"rocprofv3 --kernel-trace --stats --output-format csv -d {args.trace} -o bench -- "
                 "python3 bench.py --steps 1 --warmup 0 --no-cpu --no-swap --no-density --no-trotter --fp32-qubits 0")
    pmc_cmd = ("rocprofv3 --pmc {c} --output-format csv -d {d} -o bench -- "
               "python3 bench.py --steps 1 --warmup 0 --no-cpu --no-fused --no-swap --no-density --no-trotter")
    fetch, calls = counter_means(args.fetch)
    write, _ = counter_means(args.write)
    n_amps = 1 << args.qubits
    known_kib = n_amps * 16 / 1024.0
    cal = [k for k in fetch if k[0].startswith("kScale<double>")]
    if not cal:
        raise SystemExit("no kScale<double> dispatch in the PMC runs: cannot calibrate")
    f_fac = known_kib / fetch[cal[0]]
    w_fac = known_kib / write[cal[0]]
    lines = [f"# {args.name}: HBM traffic per dispatch (rocprofv3 PMC)", "",
             "Commands:", f"- `{pmc_cmd.format(c='FETCH_SIZE', d=args.fetch)}`",
             f"- `{pmc_cmd.format(c='WRITE_SIZE', d=args.write)}`",
             f"- `python3 -m quantumlab_amd.profile_summary --name {args.name} --trace {args.trace} "
             f"--fetch {args.fetch} --write {args.write} --out {args.out}`", "",
             f"Calibration: kScale<double> (in-place sweep of 2^{args.qubits} FP64 amplitudes, known "
             f"{known_kib / 2**20:.1f} GiB read and {known_kib / 2**20:.1f} GiB written) reads "
             f"FETCH_SIZE {fetch[cal[0]] / 2**20:.3f} GiB and WRITE_SIZE {write[cal[0]] / 2**20:.3f} GiB, so "
             f"fetch factor = {f_fac:.4f}, write factor = {w_fac:.4f}.", "",
             "| kernel | grid | dispatches | fetch GiB (calibrated) | write GiB (calibrated) | total GiB |",
             "|---|---:|---:|---:|---:|---:|"]
    traffic = {}
    for k in sorted(fetch, key=lambda k: -fetch[k]):
        if k not in write or fetch[k] < 1024:
            continue
        fg, wg = fetch[k] * f_fac / 2**20, write[k] * w_fac / 2**20
        lines.append(f"| {k[0]} | {k[1]} | {calls[k]} | {fg:.3f} | {wg:.3f} | {fg + wg:.3f} |")
        traffic[f"{k[0]}|{k[1]}"] = (fg + wg) * 2**30
    os.makedirs(args.out, exist_ok=True)
    with open(os.path.join(args.out, f"{args.name}_pmc_traffic.md"), "w") as f:
        f.write("\n".join(lines) + "\n")
    with open(os.path.join(args.out, f"{args.name}_kernel_trace.md"), "w") as f:
        f.write(f"# {args.name}: kernel durations (rocprofv3 kernel trace)\n\n")
        f.write(trace_table(args.trace, trace_cmd) + "\n")
    with open(os.path.join(args.out, f"{args.name}_traffic.json"), "w") as f:
        json.dump({"qubits": args.qubits, "fetch_factor": f_fac, "write_factor": w_fac,
                   "bytes_per_dispatch": traffic}, f, indent=1, sort_keys=True)
    print("\n".join(lines))


if __name__ == "__main__":
    main()
