#!/usr/bin/env python3
"""Benchmark of the QuEST state-vector hot path on one MI355X.

Workload: the seeded brick circuit at 32 qubits, complex FP64 (64 GiB state),
depth in 5-layer periods: every layer is compactUnitary on all 32 qubits,
controlledNot on alternating pairs, rotateZ on one random qubit; every 5th layer
adds multiControlledUnitary (3 controls) and a Haar 5-qubit multiQubitUnitary.
One step = one 5-layer period (~250 gates), applied through QuEST's C API on
libquest_mi355x, one kernel per gate, state resident in HBM from the start.

Reported: `value` = algorithmic GB/s of the per-gate FP64 sweeps (bytes each
gate must read and write: 2*N*16 B for a dense 1-target gate, N*16 B for
controlledNot, 2*N*16/2^c B for c controls) over the timed steps; gates/s;
the roofline of the dominant kernel (compactUnitary's 2x2 sweep) timed per
launch with HIP events on the library stream; the fused-FP32 leg (same
circuit, dense <=5-qubit blocks, FP32 amplitudes, FP64 reductions) with its
speed-up and fidelity / max-abs error against the FP64 run; and the CPU
baseline: QuEST 3.4's own OpenMP CPU backend (oracle/_ref, compiled from the
reference sources) or, where that is absent, the oracle's C restatement,
timed on this host on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GB/s+gates/s at 32q FP64; swap-tests/s batched; Trotter solve time at F≥1-1e-4"
HBM_PEAK_GBS = 8000.0
PERIOD = 5


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def period_gates(n, step, seed):
    from quantumlab_amd import workloads as W
    return W.brick_circuit(n, PERIOD, seed, extra_every=PERIOD, first_layer=step * PERIOD)


def run_fp64_per_gate(qe, n, steps, warmup, seed):
    from quantumlab_amd import workloads as W
    from quantumlab_amd.quest import cplx
    q = qe.qureg(n)
    plans = [period_gates(n, s, seed) for s in range(warmup + steps)]
    for s in range(warmup):
        W.apply_gates(qe, q, plans[s])
        log(f"fp64 warmup step {s} done")
    qe.qmxSync()
    slots = []  # (slot, target)
    slot = 0
    t0 = time.perf_counter()
    qe.qmxTimerBegin()
    for s in range(warmup, warmup + steps):
        for g in plans[s]:
            if g[0] == "cu":
                qe.lib.qmxTimerSlotBegin(slot)
                qe.lib.compactUnitary(q, g[1], cplx(g[2]), cplx(g[3]))
                qe.lib.qmxTimerSlotEnd(slot)
                slots.append((slot, g[1]))
                slot += 1
            else:
                W.apply_gates(qe, q, [g])
    ev_ms = qe.qmxTimerEndMs()
    qe.qmxSync()
    wall = time.perf_counter() - t0
    qe.check()
    # FP64 norm check + renormalisation (outside the timed region): one reduction
    # and one in-place scale sweep of known size (read + write 16 B per amplitude),
    # which also calibrates the HBM traffic counters against a plain stream
    norm2 = qe.qmxRenormalise(q, 1)
    log(f"fp64 norm^2 after timed steps {norm2:.15f}")
    cu = [(qe.lib.qmxTimerSlotMs(sl), t) for sl, t in slots]
    timed = [g for s in range(warmup, warmup + steps) for g in plans[s]]
    log(f"fp64 timed {steps} steps: wall {wall:.3f}s events {ev_ms / 1e3:.3f}s")
    return q, plans, timed, wall, ev_ms, cu


def run_fp32_fused(qe, n, plans, warmup, steps, k):
    from quantumlab_amd import workloads as W
    from quantumlab_amd.quest import QMX_PREC_FP32
    q = qe.qureg(n, precision=QMX_PREC_FP32)
    circuits = []
    for gates in plans:
        c = W.build_circuit(qe, n, gates)
        qe.qmxCircuitFuse(c, k)
        circuits.append(c)
    blocks = sum(qe.qmxCircuitNumBlocks(c) for c in circuits[warmup:])
    for s in range(warmup):
        qe.qmxCircuitApply(q, circuits[s], 1)
    qe.qmxSync()
    t0 = time.perf_counter()
    for s in range(warmup, warmup + steps):
        qe.qmxCircuitApply(q, circuits[s], 1)
    qe.qmxSync()
    wall = time.perf_counter() - t0
    norm2 = qe.qmxRenormalise(q, 1)  # FP64 norm, periodic renormalisation
    for c in circuits:
        qe.qmxCircuitDestroy(c)
    log(f"fp32 fused timed {steps} steps: wall {wall:.3f}s blocks {blocks} norm^2 before renorm {norm2:.12f}")
    return q, wall, blocks, norm2


def cpu_baseline(n_cpu, seed):
    """QuEST's own CPU backend (oracle/_ref) on one 5-layer period at n_cpu qubits."""
    from quantumlab_amd import workloads as W
    from quantumlab_amd.quest import QuEST
    ref = os.path.join(ROOT, "oracle", "_ref", "libquest_ref.so")
    port = os.path.join(ROOT, "oracle", "build", "libcpuref.so")
    kind, path = ("reference", ref) if os.path.exists(ref) else ("port", port)
    lib = QuEST(path)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    gates = period_gates(n_cpu, 0, seed)
    q = lib.qureg(n_cpu)
    t0 = time.perf_counter()
    W.apply_gates(lib, q, gates)
    dt = time.perf_counter() - t0
    lib.destroy(q)
    gbytes = W.gate_bytes(gates, n_cpu, 16) / 1e9
    return {"value": round(gbytes / dt, 3), "unit": "GB/s", "cores": threads, "kind": kind,
            "sample": f"{n_cpu}-qubit FP64 brick period ({PERIOD} layers, {len(gates)} gates, same generator) "
                      f"through QuEST's C API; {'QuEST 3.4 CPU backend compiled from the reference sources' if kind == 'reference' else 'C restatement of it'}, "
                      f"OpenMP {threads} threads; {dt:.2f} s",
            "gates_per_s": round(len(gates) / dt, 2)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=32)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--fuse-qubits", type=int, default=5)
    ap.add_argument("--cpu-qubits", type=int, default=27)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fused", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (the state-vector "
              f"path here has no multi-GPU sharding; see DESIGN.md for how it would shard)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import quantumlab_amd
    from quantumlab_amd import workloads as W

    qe = quantumlab_amd.load()
    buf = __import__("ctypes").create_string_buffer(512)
    qe.qmxDeviceInfo(buf, 512)
    log(buf.value.decode())
    n = args.qubits
    N = 1 << n

    q64, plans, timed, wall, ev_ms, cu = run_fp64_per_gate(qe, n, args.steps, args.warmup, args.seed)
    alg_bytes = W.gate_bytes(timed, n, 16)
    value = alg_bytes / wall / 1e9
    cu_ms = np.array([m for m, _ in cu])
    cu_bytes = 2.0 * N * 16
    achieved = cu_bytes / (cu_ms.mean() / 1e3) / 1e9
    per_target = {}
    for m, t in cu:
        per_target.setdefault(t, []).append(m)
    tmeans = {t: float(np.mean(v)) for t, v in per_target.items()}
    # HBM bytes per launch of the dominant kernel, from the calibrated PMC runs
    # summarised in profiles/brick32_fp64_pmc_traffic.md (same workload)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "brick32_fp64_traffic.json")
    if os.path.exists(tpath):
        tj = json.load(open(tpath))
        key = f"kMatrix2<double, false, true>|{N // 2}"
        if tj.get("qubits") == n and key in tj["bytes_per_dispatch"]:
            traffic = tj["bytes_per_dispatch"][key]

    result = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"brick circuit, {n} qubits, complex FP64 state ({N * 16 / 2**30:.0f} GiB), "
                               f"one step = {PERIOD} layers incl. multiControlledUnitary(3 ctrls) + "
                               f"5-qubit multiQubitUnitary, per-gate QuEST C-API kernels",
                   "qubits": n, "gates_per_step": len(timed) // args.steps, "seed": args.seed},
        "gates_per_s": round(len(timed) / wall, 2),
        "roofline": {"bound": "hbm", "kernel": "kMatrix2<double> (compactUnitary 2x2 pair sweep)",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "avg_launch_ms": round(float(cu_ms.mean()), 4), "launches": int(len(cu_ms)),
                     "algorithmic_bytes_per_launch": cu_bytes,
                     "traffic_source": "profiles/brick32_fp64_pmc_traffic.md" if traffic else None,
                     "per_target_ms_min": round(min(tmeans.values()), 4),
                     "per_target_ms_max": round(max(tmeans.values()), 4)},
        "event_time_s": round(ev_ms / 1e3, 4),
    }
    log(f"fp64 per-gate: {value:.1f} GB/s, {result['gates_per_s']:.1f} gates/s, compactUnitary "
        f"{cu_ms.mean():.3f} ms avg -> {achieved:.1f} GB/s ({achieved / HBM_PEAK_GBS:.1%} of peak)")

    if not args.no_fused:
        q32, wall32, blocks, norm2 = run_fp32_fused(qe, n, plans, args.warmup, args.steps, args.fuse_qubits)
        ip = qe.calcInnerProduct(q64, q32)
        # normalised fidelity: norms from FP64 reductions (the FP32-stored vector
        # re-rounds its norm^2 by ~1e-9 even right after renormalisation)
        fid = (ip.real ** 2 + ip.imag ** 2) / (qe.calcTotalProb(q64) * qe.calcTotalProb(q32))
        maxabs = qe.qmxMaxAbsDiff(q64, q32)
        p_err = max(abs(qe.calcProbOfOutcome(q64, k, 0) - qe.calcProbOfOutcome(q32, k, 0)) for k in range(0, n, 4))
        result["fused_fp32"] = {"speedup_vs_fp64_per_gate": round(wall / wall32, 3),
                                "ms_per_step": round(wall32 / args.steps * 1e3, 3),
                                "blocks_per_step": blocks / args.steps, "max_block_qubits": args.fuse_qubits,
                                "gates_per_s": round(len(timed) / wall32, 2),
                                "fidelity_error": float(1 - fid), "max_abs_amp_error": float(maxabs),
                                "max_prob_error": float(p_err), "norm2_before_renorm": float(norm2)}
        log(f"fused fp32: {wall / wall32:.2f}x vs fp64 per-gate, 1-F={1 - fid:.3e}, max|da|={maxabs:.3e}")
        qe.destroy(q32)
    qe.destroy(q64)

    if not args.no_cpu:
        log(f"cpu baseline at {args.cpu_qubits} qubits ...")
        result["cpu_baseline"] = cpu_baseline(args.cpu_qubits, args.seed)
        log(f"cpu baseline: {result['cpu_baseline']['value']} GB/s")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
