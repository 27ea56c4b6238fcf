"""Turn rocprofv3 CSV output into the short summaries kept under profiles/.

One summary per bench leg (`bench.py --only LEG`), from up to four rocprofv3
runs of the same command, each into its own directory (-d), all with
`--output-format csv -o bench`:
  --trace  the kernel trace (`--kernel-trace --stats`): per-kernel durations;
  --fetch  `--pmc FETCH_SIZE`, --write `--pmc WRITE_SIZE` (KiB per dispatch):
           HBM traffic per dispatch;
  --sq     `--pmc` with the SQ_* counters below: instruction mix and waits.

Usage (from the repository root, after the rocprofv3 runs):
    python3 -m quantumlab_amd.profile_summary --name LEG_NAME --bench-args "--only fused ..." \\
        --trace DIR [--fetch DIR --write DIR] [--sq DIR] [--factors profiles/<stream>_traffic.json]

FETCH_SIZE / WRITE_SIZE are calibrated against a kernel of known traffic:
kScale, the in-place renormalisation sweep, which reads and writes every
amplitude of its register exactly once (bench.py --only stream runs it on a
32-qubit FP64 register). Its factors are written to <name>_traffic.json and
other legs take them from there (--factors), since they are properties of the
device's counters, not of the kernel.
"""
from __future__ import annotations

import argparse
import csv
import json
import os
import re
from collections import defaultdict

# one pass holds at most 8 SQ counters
SQ_COUNTERS = ["SQ_INSTS_VALU", "SQ_INSTS_MFMA", "SQ_INSTS_LDS", "SQ_LDS_BANK_CONFLICT", "SQ_WAIT_ANY",
               "SQ_WAVE_CYCLES", "SQ_BUSY_CYCLES", "SQ_VALU_MFMA_BUSY_CYCLES"]


def short(name: str) -> str:
    name = re.sub(r"qmx::\(anonymous namespace\)::", "", name)
    name = re.sub(r"qmx::", "", name)
    return name.split("(")[0].replace("void ", "")


def rocprof_cmd(flags: str, d: str, bench_args: str) -> str:
    return f"rocprofv3 {flags} --output-format csv -d {d} -o bench -- python3 bench.py {bench_args}"


def trace_rows(d: str):
    rows = list(csv.DictReader(open(os.path.join(d, "bench_kernel_stats.csv"))))
    return [r for r in rows if float(r["Percentage"]) >= 0.01]


def trace_table(d: str, cmd: str) -> str:
    out = [f"Command: `{cmd}`", "", "| kernel | calls | avg ms | min ms | max ms | % time |",
           "|---|---:|---:|---:|---:|---:|"]
    for r in trace_rows(d):
        out.append(f"| {short(r['Name'])} | {r['Calls']} | {float(r['AverageNs']) / 1e6:.3f} | "
                   f"{float(r['MinNs']) / 1e6:.3f} | {float(r['MaxNs']) / 1e6:.3f} | {float(r['Percentage']):.2f} |")
    return "\n".join(out)


def counters(d: str):
    """(kernel, grid) -> counter -> mean value per dispatch; and dispatch counts."""
    agg = defaultdict(lambda: defaultdict(list))
    for r in csv.DictReader(open(os.path.join(d, "bench_counter_collection.csv"))):
        agg[(short(r["Kernel_Name"]), int(r["Grid_Size"]))][r.get("Counter_Name", "value")].append(
            float(r["Counter_Value"]))
    means = {k: {c: sum(v) / len(v) for c, v in cs.items()} for k, cs in agg.items()}
    calls = {k: max(len(v) for v in cs.values()) for k, cs in agg.items()}
    return means, calls


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--name", required=True)
    ap.add_argument("--bench-args", required=True, help="the bench.py arguments every run used")
    ap.add_argument("--trace", required=True)
    ap.add_argument("--fetch", default=None)
    ap.add_argument("--write", default=None)
    ap.add_argument("--sq", default=None)
    ap.add_argument("--factors", default=None, help="a _traffic.json with fetch/write factors (calibration)")
    ap.add_argument("--calibration-qubits", type=int, default=32)
    ap.add_argument("--qubits", type=int, default=None, help="the workload's register size (recorded)")
    ap.add_argument("--out", default="profiles")
    args = ap.parse_args()
    os.makedirs(args.out, exist_ok=True)
    cmds = [rocprof_cmd("--kernel-trace --stats", args.trace, args.bench_args)]
    with open(os.path.join(args.out, f"{args.name}_kernel_trace.md"), "w") as f:
        f.write(f"# {args.name}: kernel durations (rocprofv3 kernel trace)\n\n")
        f.write(trace_table(args.trace, cmds[0]) + "\n")
    if not (args.fetch or args.sq):
        return
    lines = [f"# {args.name}: counters per dispatch (rocprofv3 PMC)", "", "Commands:"]
    summary = {"name": args.name, "bench_args": args.bench_args, "qubits": args.qubits}
    if args.fetch and args.write:
        fetch, calls = counters(args.fetch)
        write, _ = counters(args.write)
        cal = [k for k in fetch if k[0].startswith("kScale<double>")]
        if cal:
            known_kib = (1 << args.calibration_qubits) * 16 / 1024.0
            f_fac = known_kib / fetch[cal[0]]["FETCH_SIZE"]
            w_fac = known_kib / write[cal[0]]["WRITE_SIZE"]
            cal_text = (f"Calibration: kScale<double> (in-place sweep of 2^{args.calibration_qubits} FP64 "
                        f"amplitudes, {known_kib / 2**20:.1f} GiB read and written) reads FETCH_SIZE "
                        f"{fetch[cal[0]]['FETCH_SIZE'] / 2**20:.3f} GiB and WRITE_SIZE "
                        f"{write[cal[0]]['WRITE_SIZE'] / 2**20:.3f} GiB: fetch factor {f_fac:.4f}, "
                        f"write factor {w_fac:.4f}.")
        elif args.factors:
            fac = json.load(open(args.factors))
            f_fac, w_fac = fac["fetch_factor"], fac["write_factor"]
            cal_text = (f"Calibration: fetch factor {f_fac:.4f}, write factor {w_fac:.4f}, from the kScale<double> "
                        f"sweep of {os.path.basename(args.factors)}.")
        else:
            raise SystemExit("no kScale<double> dispatch and no --factors: cannot calibrate")
        cmds += [rocprof_cmd("--pmc FETCH_SIZE", args.fetch, args.bench_args),
                 rocprof_cmd("--pmc WRITE_SIZE", args.write, args.bench_args)]
        table = ["| kernel | grid | dispatches | fetch GiB | write GiB | total GiB |", "|---|---:|---:|---:|---:|---:|"]
        traffic = {}
        for k in sorted(fetch, key=lambda k: -fetch[k]["FETCH_SIZE"]):
            if k not in write or fetch[k]["FETCH_SIZE"] + write[k]["WRITE_SIZE"] < 64:
                continue
            fg = fetch[k]["FETCH_SIZE"] * f_fac / 2**20
            wg = write[k]["WRITE_SIZE"] * w_fac / 2**20
            table.append(f"| {k[0]} | {k[1]} | {calls[k]} | {fg:.4f} | {wg:.4f} | {fg + wg:.4f} |")
            traffic[f"{k[0]}|{k[1]}"] = (fg + wg) * 2**30
        summary.update({"fetch_factor": f_fac, "write_factor": w_fac, "bytes_per_dispatch": traffic})
        traffic_text = [cal_text, "", "HBM traffic per dispatch (calibrated):", ""] + table
    else:
        traffic_text = []
    if args.sq:
        sq, calls = counters(args.sq)
        present = [c for c in SQ_COUNTERS if any(c in v for v in sq.values())]
        cmds.append(rocprof_cmd("--pmc " + " ".join(present), args.sq, args.bench_args))
        table = ["| kernel | grid | dispatches | " + " | ".join(present) +
                 " | VALU/MFMA | wait share | LDS conflict cycles/LDS instr |",
                 "|---|---:|---:|" + "---:|" * (len(present) + 3)]
        sqsum = {}
        for k in sorted(sq, key=lambda k: -sq[k].get("SQ_WAVE_CYCLES", 0)):
            v = sq[k]
            if v.get("SQ_WAVE_CYCLES", 0) < 1e4:
                continue
            ratio = v.get("SQ_INSTS_VALU", 0) / v["SQ_INSTS_MFMA"] if v.get("SQ_INSTS_MFMA") else None
            wait = v.get("SQ_WAIT_ANY", 0) / v["SQ_WAVE_CYCLES"] if v.get("SQ_WAVE_CYCLES") else None
            lds = v.get("SQ_LDS_BANK_CONFLICT", 0) / v["SQ_INSTS_LDS"] if v.get("SQ_INSTS_LDS") else None
            fmt = lambda x: "-" if x is None else f"{x:.3g}"
            table.append(f"| {k[0]} | {k[1]} | {calls[k]} | " + " | ".join(f"{v.get(c, 0):.4g}" for c in present)
                         + f" | {fmt(ratio)} | {fmt(wait)} | {fmt(lds)} |")
            sqsum[f"{k[0]}|{k[1]}"] = dict(v, valu_per_mfma=ratio, wait_share=wait, lds_conflict_per_instr=lds)
        summary["sq"] = sqsum
        sq_text = ["", "Instruction mix and waits per dispatch (SQ counters; wait share = SQ_WAIT_ANY / "
                   "SQ_WAVE_CYCLES):", ""] + table
    else:
        sq_text = []
    cmds.append(f"python3 -m quantumlab_amd.profile_summary --name {args.name} --bench-args \"{args.bench_args}\" "
                f"--trace {args.trace}" + (f" --fetch {args.fetch} --write {args.write}" if args.fetch else "")
                + (f" --sq {args.sq}" if args.sq else "") + (f" --factors {args.factors}" if args.factors else "")
                + (f" --qubits {args.qubits}" if args.qubits else "") + f" --out {args.out}")
    lines += [f"- `{c}`" for c in cmds] + [""] + traffic_text + sq_text
    with open(os.path.join(args.out, f"{args.name}_pmc.md"), "w") as f:
        f.write("\n".join(lines) + "\n")
    with open(os.path.join(args.out, f"{args.name}_traffic.json"), "w") as f:
        json.dump(summary, f, indent=1, sort_keys=True)
    print("\n".join(lines))


if __name__ == "__main__":
    main()
