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

`swap_tests`: the batched QKMeans workload (16384 points of dimension 2^q,
q uniform in 4..10, 16 centroids, 10 k-means iterations; amplitude encoding,
swap test of every point against every centroid, assignment and centroid
update per iteration, FP32 registers and FP64 sums, one HIP graph per
iteration) in swap-tests/s, with its distance error against exact FP64 inner
products, assignment agreement with the FP64 run, and its own CPU baseline:
the reference's gate-by-gate swap-test circuits on QuEST's CPU backend.

`trotter`: Hamiltonian simulation with the learned Trotter planner at 24
qubits (FP32): solve time (features + network + one applyTrotterCircuit at
the planner's order and repetitions) against plain step doubling at fixed
order 2 and 4, the learned result's infidelity against the 4x-finer FP64
run, and QuEST's CPU backend per Pauli exponential scaled to the same solves.

`density`: the noisy brick circuit (60 layers, dephasing + depolarising +
damping after every gate on its qubits) on a 14-qubit density matrix in FP64
and FP32, with purity, per-qubit P(0), fidelity with the noiseless run, the
FP32 form's errors against FP64, and QuEST's CPU backend at 13 qubits.
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
FP32_MATRIX_PEAK_TFLOPS = 157.3
# FP32 FMA on the VALU without packed instructions: half of the published
# 157.3 TFLOP/s vector peak, which counts packed FP32 (not used here)
FP32_VALU_PEAK_TFLOPS = 78.6


def profiled_traffic(name, kernel, qubits=None):
    """HBM bytes per launch of `kernel` (name prefix) from the calibrated
    FETCH_SIZE / WRITE_SIZE summary profiles/<name>_traffic.json that
    quantumlab_amd.profile_summary wrote for the same workload; None without it."""
    path = os.path.join(ROOT, "profiles", f"{name}_traffic.json")
    if not os.path.exists(path):
        return None, None
    tj = json.load(open(path))
    if qubits is not None and tj.get("qubits") not in (None, qubits):
        return None, None
    for key, b in sorted(tj.get("bytes_per_dispatch", {}).items()):
        if key.split("|")[0].startswith(kernel):
            return b, f"profiles/{name}_pmc.md"
    return None, None
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


def fused_bytes(qe, c, n, amp_bytes=8):
    """Algorithmic HBM bytes of one application of a fused brick-circuit
    period, and the number of full sweeps in it. A step of several gates
    (padded to 5 qubits on registers of >= 9 qubits) or a lone 5-qubit gate
    reads and writes the register once; a lone gate on fewer qubits runs its
    per-gate kernel and moves what that moves: a one-qubit gate 2NB, a
    controlledNot NB, the 3-control unitary 2NB/8."""
    import ctypes as C
    N = float(1 << n)
    targs = (C.c_int * 64)()
    total, sweeps = 0.0, 0
    for s in range(qe.qmxCircuitNumBlocks(c)):
        k = qe.qmxCircuitStep(c, s, targs, None, None)
        if k >= 5 or k == 1:
            total += 2 * N * amp_bytes
            sweeps += k >= 5
        elif k == 2:
            total += N * amp_bytes
        elif k == 4:
            total += 2 * N * amp_bytes / 8
        else:
            total += 2 * N * amp_bytes
    return total, sweeps


def run_fp32_fused(qe, n, plans, warmup, steps, k):
    """Fused periods on an n-qubit FP32 register: (register, wall time of the
    timed periods, their steps, the FP64 norm^2 before renormalising, their
    algorithmic bytes, their full 5-qubit sweeps)."""
    from quantumlab_amd import workloads as W
    from quantumlab_amd.quest import QMX_PREC_FP32
    q = qe.qureg(n, precision=QMX_PREC_FP32)
    circuits = []
    for gates in plans:
        c = W.build_circuit(qe, n, gates)
        qe.qmxCircuitFuse(c, k)
        circuits.append(c)
    blocks = sum(qe.qmxCircuitNumBlocks(c) for c in circuits[warmup:])
    moved = [fused_bytes(qe, c, n) for c in circuits[warmup:]]
    nbytes, sweeps = sum(m[0] for m in moved), sum(m[1] for m in moved)
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
    log(f"fp32 fused timed {steps} steps: wall {wall:.3f}s blocks {blocks} ({sweeps} full sweeps) "
        f"norm^2 before renorm {norm2:.12f}")
    return q, wall, blocks, norm2, nbytes, sweeps


def run_fp32_large(qe, n, seed, k):
    """One period of the brick circuit on an n-qubit FP32 register (33 qubits:
    64 GiB) through fused blocks, timed; then the same gates one by one on an
    n-qubit FP64 register (128 GiB at 33 qubits) as its reference. The fused
    period's time is the median of `timed` applications: the first (from |0>)
    gives the fidelity, the others run on the evolved register afterwards."""
    import numpy as np
    from quantumlab_amd import workloads as W
    from quantumlab_amd.quest import QMX_PREC_FP32
    timed = 3
    gates = period_gates(n, 0, seed)
    q32 = qe.qureg(n, precision=QMX_PREC_FP32)
    c = W.build_circuit(qe, n, gates)
    qe.qmxCircuitFuse(c, k)
    blocks = qe.qmxCircuitNumBlocks(c)
    nbytes, sweeps = fused_bytes(qe, c, n)
    qe.qmxSync()
    t0 = time.perf_counter()
    qe.qmxCircuitApply(q32, c, 1)
    qe.qmxSync()
    walls = [time.perf_counter() - t0]
    q64 = qe.qureg(n)
    qe.qmxSync()
    t0 = time.perf_counter()
    W.apply_gates(qe, q64, gates)
    qe.qmxSync()
    wall64 = time.perf_counter() - t0
    ip = qe.calcInnerProduct(q64, q32)
    fid = (ip.real ** 2 + ip.imag ** 2) / (qe.calcTotalProb(q64) * qe.calcTotalProb(q32))
    maxabs = qe.qmxMaxAbsDiff(q64, q32)
    qe.destroy(q64)
    for _ in range(timed - 1):
        qe.qmxSync()
        t0 = time.perf_counter()
        qe.qmxCircuitApply(q32, c, 1)
        qe.qmxSync()
        walls.append(time.perf_counter() - t0)
    qe.qmxCircuitDestroy(c)
    qe.destroy(q32)
    wall = float(np.median(walls))
    N = 1 << n
    gbs = nbytes / wall / 1e9
    log(f"fused fp32 at {n} qubits: {wall * 1e3:.1f} ms (median of {[round(w * 1e3, 1) for w in walls]}) for "
        f"{len(gates)} gates in {blocks} blocks ({gbs:.0f} GB/s); "
        f"fp64 per-gate {wall64:.2f} s; 1-F={1 - fid:.3e}, max|da|={maxabs:.3e}")
    return {"qubits": n, "state_gib": N * 8 / 2**30, "gates": len(gates), "blocks": blocks, "full_sweeps": sweeps,
            "ms_per_step": round(wall * 1e3, 3), "ms_per_step_each": [round(w * 1e3, 3) for w in walls],
            "gates_per_s": round(len(gates) / wall, 2),
            "fp64_per_gate_seconds": round(wall64, 3),
            "fp64_per_gate_gb_per_s": round(W.gate_bytes(gates, n, 16) / wall64 / 1e9, 1),
            "speedup_vs_fp64_per_gate": round(wall64 / wall, 3),
            "fidelity_error": float(1 - fid), "max_abs_amp_error": float(maxabs),
            "roofline": {"bound": "hbm", "kernel": "kFused5Tile", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
                         "mfma_tflops": round(sweeps * N * 192 / wall / 1e12, 2)}}


# CPUs this process may use, read before any OpenMP runtime starts (with
# OMP_PROC_BIND set it pins the calling thread to its first place)
CPUS_ALLOWED = len(os.sched_getaffinity(0))


def host_cpu():
    """The host the CPU baselines run on: model, physical cores, logical CPUs,
    the CPUs this process may use, and the OpenMP settings in force."""
    model, cores = None, set()
    phys = core = None
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name" and model is None:
                model = v
            elif k == "physical id":
                phys = v
            elif k == "core id":
                core = v
            elif not k and phys is not None and core is not None:
                cores.add((phys, core))
                phys = core = None
    except OSError:
        pass
    return {"model": model, "physical_cores": len(cores) or None, "logical_cpus": os.cpu_count(),
            "cpus_allowed": CPUS_ALLOWED, "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "omp_proc_bind": os.environ.get("OMP_PROC_BIND"), "omp_places": os.environ.get("OMP_PLACES")}


def cpu_baseline(n_cpu, seed):
    """QuEST's own CPU backend (oracle/_ref) on one 5-layer period at n_cpu
    qubits (30 by default: a 32-qubit period would take minutes; the per-gate
    cost of these bandwidth-bound sweeps scales with the amplitude count, so
    the 32-qubit figure is the 30-qubit time x 4)."""
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
            "gates_per_s": round(len(gates) / dt, 2),
            "s_per_period_scaled_to_32q": round(dt * (1 << (32 - n_cpu)), 2) if n_cpu <= 32 else None,
            "host": host_cpu()}


def run_stream(qe, n, reps=3):
    """The achievable HBM rate on this box, from the library's own streaming
    kernels on an n-qubit FP64 register (2^n x 16 B): calcTotalProb (kSumSq,
    one read sweep), qmxRenormalise (kSumSq then kScale, an in-place read +
    write sweep: kScale's time is the difference) and cloneQureg (the runtime's
    device-to-device copy into a second register). Best of `reps`, HIP events on
    the library stream. kScale is also the calibration kernel of the PMC
    traffic summaries (it reads and writes every amplitude exactly once)."""
    N = 1 << n
    q, q2 = qe.qureg(n), qe.qureg(n)
    qe.initPlusState(q)
    best = {"read": 1e30, "renorm": 1e30, "copy": 1e30}
    for _ in range(reps):
        qe.qmxSync()
        qe.qmxTimerBegin()
        qe.calcTotalProb(q)
        best["read"] = min(best["read"], qe.qmxTimerEndMs())
        qe.qmxTimerBegin()
        qe.qmxRenormalise(q, 1)
        best["renorm"] = min(best["renorm"], qe.qmxTimerEndMs())
        qe.qmxTimerBegin()
        qe.cloneQureg(q2, q)
        best["copy"] = min(best["copy"], qe.qmxTimerEndMs())
    qe.destroy(q)
    qe.destroy(q2)
    nbytes = N * 16.0
    scale_ms = best["renorm"] - best["read"]
    res = {"qubits": n, "bytes_per_register": nbytes,
           "read_sweep_ms": round(best["read"], 3), "read_GBps": round(nbytes / best["read"] / 1e6, 1),
           "in_place_rw_ms": round(scale_ms, 3), "in_place_rw_GBps": round(2 * nbytes / scale_ms / 1e6, 1),
           "copy_ms": round(best["copy"], 3), "copy_GBps": round(2 * nbytes / best["copy"] / 1e6, 1),
           "peak_GBps": HBM_PEAK_GBS}
    log(f"stream at {n} qubits: read {res['read_GBps']} GB/s, in-place r+w {res['in_place_rw_GBps']} GB/s, "
        f"copy {res['copy_GBps']} GB/s")
    return res


def run_density(qe, n, layers, seed):
    """Noisy brick circuit on an n-qubit density matrix in FP64 and in FP32
    (FP64 reductions): wall time, algorithmic GB/s, purity, per-qubit P(0),
    fidelity with the noiseless pure-state run, and the FP32 form's errors
    against the FP64 form."""
    from quantumlab_amd import workloads as W
    from quantumlab_amd.quest import QMX_PREC_FP32, QMX_PREC_FP64
    gates = W.noisy_brick_circuit(n, layers, seed)
    pure = qe.qureg(n)
    W.apply_gates(qe, pure, W.unitary_part(gates))
    res = {"workload": f"{n}-qubit density matrix (4^{n} entries, {(1 << 2 * n) * 16 / 2**30:.0f} GiB FP64) from |0><0|, "
                       f"{layers} brick layers, mixDephasing + mixDepolarising + mixDamping after every gate on "
                       f"each qubit it touches",
           "operations": len(gates), "channels": sum(1 for g in gates if g[0] in W.CHANNELS)}
    import ctypes as C

    def fusion_stats():
        sw, ops = C.c_longlong(0), C.c_longlong(0)
        qe.lib.qmxDensityFusionStats(C.byref(sw), C.byref(ops))
        return sw.value, ops.value

    outs = {}
    entries = float(1 << (2 * n))
    for name, prec, B in (("fp64", QMX_PREC_FP64, 16), ("fp32", QMX_PREC_FP32, 8)):
        rho = qe.qureg(n, density=True, precision=prec)
        qe.qmxSync()
        s0 = fusion_stats()
        t0 = time.perf_counter()
        W.apply_gates(qe, rho, gates)
        qe.qmxSync()
        wall = time.perf_counter() - t0
        s1 = fusion_stats()
        sweeps, absorbed = s1[0] - s0[0], s1[1] - s0[1]
        probs = [qe.calcProbOfOutcome(rho, k, 0) for k in range(n)]
        outs[name] = {"seconds": round(wall, 4), "ms_per_layer": round(wall / layers * 1e3, 3),
                      "sweeps": sweeps, "operations_composed": absorbed,
                      "gb_per_s": round(sweeps * 2 * entries * B / wall / 1e9, 1),
                      "per_operation_gb_per_s": round(W.density_bytes(gates, n, B) / wall / 1e9, 1),
                      "purity": qe.calcPurity(rho), "fidelity_vs_noiseless": qe.calcFidelity(rho, pure),
                      "trace": qe.calcTotalProb(rho), "prob0": probs}
        qe.destroy(rho)
        log(f"density {n}q {name}: {wall:.2f} s, {sweeps} sweeps for {absorbed} operations, "
            f"{outs[name]['gb_per_s']} GB/s, purity {outs[name]['purity']:.6f}, "
            f"F {outs[name]['fidelity_vs_noiseless']:.6f}")
    qe.destroy(pure)
    a, b = outs["fp64"], outs["fp32"]
    res.update(outs)
    res["roofline"] = {"bound": "hbm", "kernel": "kDensePairMfma<double> (two composed 2-qubit superoperators per sweep of the 2n-qubit vector, on the matrix cores)",
                       "achieved": a["gb_per_s"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": round(a["gb_per_s"] / HBM_PEAK_GBS, 4),
                       "traffic": profiled_traffic("density14_fp64", "kDensePairMfma<double>", n)[0],
                       "traffic_source": profiled_traffic("density14_fp64", "kDensePairMfma<double>", n)[1],
                       "note": "read + write of every entry per sweep (2 * 4^n * 16 B) over the leg's wall time; "
                               "all operations compose, so the sweeps are the only device traffic; a sweep applies "
                               "two 2-qubit superoperators (32 complex FMAs per entry). "
                               "per_operation_gb_per_s counts each gate and channel as its own sweep, as the "
                               "reference executes them"}
    res["fp32_vs_fp64"] = {"purity_error": abs(a["purity"] - b["purity"]),
                           "fidelity_error": abs(a["fidelity_vs_noiseless"] - b["fidelity_vs_noiseless"]),
                           "max_prob_error": max(abs(x - y) for x, y in zip(a["prob0"], b["prob0"])),
                           "speedup": round(a["seconds"] / b["seconds"], 3)}
    return res


def cpu_density_baseline(n_cpu, layers, seed):
    """QuEST's CPU backend on the noisy brick circuit at n_cpu qubits (density)."""
    from quantumlab_amd import workloads as W
    from quantumlab_amd.quest import QuEST
    ref = os.path.join(ROOT, "oracle", "_ref", "libquest_ref.so")
    port = os.path.join(ROOT, "oracle", "build", "libcpuref.so")
    kind, path = ("reference", ref) if os.path.exists(ref) else ("port", port)
    lib = QuEST(path)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    gates = W.noisy_brick_circuit(n_cpu, layers, seed)
    rho = lib.qureg(n_cpu, density=True)
    t0 = time.perf_counter()
    W.apply_gates(lib, rho, gates)
    dt = time.perf_counter() - t0
    lib.destroy(rho)
    return {"value": round(W.density_bytes(gates, n_cpu, 16) / dt / 1e9, 3), "unit": "GB/s", "cores": threads,
            "kind": kind, "ms_per_layer": round(dt / layers * 1e3, 2),
            "sample": f"{n_cpu}-qubit FP64 density matrix, {layers} noisy brick layers ({len(gates)} calls, same "
                      f"generator); {'QuEST 3.4 CPU backend compiled from the reference sources' if kind == 'reference' else 'C restatement of it'}, "
                      f"OpenMP {threads} threads; {dt:.2f} s"}


def clopper_pearson_lower(successes, trials, alpha=0.05):
    """One-sided (1 - alpha) lower confidence bound on a success probability
    (Clopper-Pearson, exact binomial)."""
    from scipy.stats import beta
    return 0.0 if successes == 0 else float(beta.ppf(alpha, successes, trials - successes + 1))


def run_trotter(qe, n, count, seed, eps=1e-4, sd_count=16):
    """Hamiltonian simulation at n qubits (FP32 registers): per instance, the
    learned planner's (order, reps) and one applyTrotterCircuit, timed from the
    Hamiltonian to the evolved state (features, network, evolution), and its
    normalised infidelity against the 4x-finer FP64 run at the same order; on
    the first sd_count instances also plain step doubling at fixed order 2 and
    4, timed the same way (the speed-up compares the same instances)."""
    import ctypes as C

    import numpy as np
    from quantumlab_amd import hamiltonians as H
    from quantumlab_amd import planner as PL
    from quantumlab_amd.quest import QMX_PREC_FP32
    net = PL.Planner(qe, path=os.path.join(ROOT, "quantumlab_amd", "data", "trotter_planner.txt"))
    rows = []
    for idx in range(count):
        inst = H.trotter_instance(n, seed, idx)
        hb = H.HamilBatch([inst])
        q = qe.qureg(n, precision=QMX_PREC_FP32)
        H.prepare(qe, q, inst)
        qe.qmxSync()
        o, r = C.c_int(), C.c_int()
        t0 = time.perf_counter()
        qe.lib.qmxApplyTrotterLearned(q, hb[0], inst.time, net.h, C.byref(o), C.byref(r))
        qe.qmxSync()
        t_learned = time.perf_counter() - t0
        ref = qe.qureg(n)
        H.prepare(qe, ref, inst)
        qe.lib.applyTrotterCircuit(ref, hb[0], inst.time, o.value, 4 * r.value)
        ip = qe.calcInnerProduct(ref, q)
        infid = 1 - (ip.real ** 2 + ip.imag ** 2) / (qe.calcTotalProb(ref) * qe.calcTotalProb(q))
        qe.destroy(ref)
        sd = {}
        for order in ((2, 4) if idx < sd_count else ()):
            H.prepare(qe, q, inst)
            qe.qmxSync()
            reps = C.c_int()
            t0 = time.perf_counter()
            ok = qe.lib.qmxApplyTrotterStepDoubling(q, hb[0], inst.time, order, eps, 1 << 14, C.byref(reps))
            qe.qmxSync()
            sd[order] = (time.perf_counter() - t0, reps.value, ok)
        qe.destroy(q)
        sweeps = qe.lib.qmxTrotterSweepsPerRep(hb[0], o.value, 0, QMX_PREC_FP32) * r.value
        rows.append((inst, o.value, r.value, t_learned, infid, sd, sweeps))
        if sd:
            log(f"trotter {idx}: {inst.model} {inst.terms} terms t={inst.time:.2f}: learned order {o.value} "
                f"x{r.value} {t_learned * 1e3:.0f} ms (1-F {infid:.1e}); step doubling o2 x{sd[2][1]} "
                f"{sd[2][0] * 1e3:.0f} ms, o4 x{sd[4][1]} {sd[4][0] * 1e3:.0f} ms")
        elif idx % 16 == 15 or count <= 16:
            log(f"trotter {idx}: learned order {o.value} x{r.value} {t_learned * 1e3:.0f} ms (1-F {infid:.1e})")
    net.close()
    tl_all = np.array([x[3] for x in rows])
    inf = np.array([x[4] for x in rows])
    succ = int(np.sum(inf <= eps))
    timed = [x for x in rows if x[5]]
    tl = np.array([x[3] for x in timed])
    t2 = np.array([x[5][2][0] for x in timed])
    t4 = np.array([x[5][4][0] for x in timed])
    exps = sum(H.stages(x[1]) * x[2] * x[0].terms for x in rows)
    sweep_bytes = sum(x[6] for x in rows) * 2.0 * (1 << n) * 8
    gbs = sweep_bytes / tl_all.sum() / 1e9
    return {"metric": "Trotter solve time at F>=1-1e-4", "unit": "ms", "higher_is_better": False,
            "workload": f"{count} seeded Hamiltonians (TFIM/Heisenberg chains + 32 random 2-4-local strings) on "
                        f"{n} qubits, random product states, t ~ U(0.5, 4); FP32 registers; none of them in the "
                        f"planner's training or calibration sets (other seeds)",
            "qubits": n, "instances": count, "dtype": "f32",
            "successes": succ, "success_lower_95": round(clopper_pearson_lower(succ, count), 4),
            "step_doubling_instances": len(timed),
            "learned_ms_mean_all": round(float(tl_all.mean() * 1e3), 2),
            "learned_ms_mean": round(float(tl.mean() * 1e3), 2) if timed else None,
            "step_doubling_o2_ms_mean": round(float(t2.mean() * 1e3), 2) if timed else None,
            "step_doubling_o4_ms_mean": round(float(t4.mean() * 1e3), 2) if timed else None,
            "speedup_vs_step_doubling_o2": round(float(t2.sum() / tl.sum()), 3) if timed else None,
            "speedup_vs_step_doubling_o4": round(float(t4.sum() / tl.sum()), 3) if timed else None,
            "speedup_vs_best_fixed_order": round(float(min(t2.sum(), t4.sum()) / tl.sum()), 3) if timed else None,
            "success_rate": float(np.mean(inf <= eps)), "max_infidelity_vs_4x_finer_fp64": float(inf.max()),
            "step_doubling_met_eps": bool(all(x[5][2][2] and x[5][4][2] for x in timed)),
            "pauli_exponentials_per_s_learned": round(exps / tl_all.sum(), 1),
            "roofline": {"bound": "hbm", "kernel": "kPauliCoset<float> (fused Pauli-rotation sweeps)",
                         "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(gbs / HBM_PEAK_GBS, 4),
                         "traffic": profiled_traffic("trotter24_fp32", "kPauliCoset<float>", n)[0],
                         "traffic_source": profiled_traffic("trotter24_fp32", "kPauliCoset<float>", n)[1],
                         "sweeps": int(sum(x[6] for x in rows)), "bytes_per_sweep": 2.0 * (1 << n) * 8,
                         "note": "read + write of the FP32 register per fused block, over the learned solves' wall "
                                 "time, host planning and launch gaps included; commuting rotations are packed into "
                                 "fewer, fuller blocks, so a sweep carries more rotations (and LDS round trips, one "
                                 "per coset group) per HBM read + write"},
            "choices": [[x[1], x[2]] for x in rows],
            "step_doubling_reps": [[x[5][2][1], x[5][4][1]] for x in timed]}, rows


def cpu_trotter_baseline(rows, n):
    """QuEST's CPU backend: one order-1 repetition (every term once) of the
    first instance at n qubits (FP64), scaled to each learned choice."""
    from quantumlab_amd import hamiltonians as H
    from quantumlab_amd.quest import QuEST
    ref = os.path.join(ROOT, "oracle", "_ref", "libquest_ref.so")
    port = os.path.join(ROOT, "oracle", "build", "libcpuref.so")
    kind, path = ("reference", ref) if os.path.exists(ref) else ("port", port)
    lib = QuEST(path)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    inst = rows[0][0]
    hb = H.HamilBatch([inst])
    q = lib.qureg(n)
    H.prepare(lib, q, inst)
    t0 = time.perf_counter()
    lib.lib.applyTrotterCircuit(q, hb[0], inst.time, 1, 1)
    dt = time.perf_counter() - t0
    lib.destroy(q)
    per_exp = dt / inst.terms
    solve = [H.stages(x[1]) * x[2] * x[0].terms * per_exp for x in rows]
    return {"value": round(1e3 * sum(solve) / len(solve), 1), "unit": "ms", "cores": threads, "kind": kind,
            "sample": f"one order-1 repetition ({inst.terms} Pauli exponentials) of the first instance at {n} "
                      f"qubits, FP64, {dt:.2f} s; the solve time is that per-exponential cost times the learned "
                      f"choices' exponentials; {'QuEST 3.4 CPU backend compiled from the reference sources' if kind == 'reference' else 'C restatement of it'}, "
                      f"OpenMP {threads} threads"}


def exact_p0(data, cents):
    """P(ancilla = 0) = (1 + <a|b>^2)/2 from FP64 numpy inner products of the
    normalised points and the normalised truncated centroids."""
    import numpy as np
    out = np.empty((len(data.qubits), cents.shape[0]))
    for q in np.unique(data.qubits):
        idx = np.nonzero(data.qubits == q)[0]
        D = 1 << int(q)
        A = np.stack([data.point(p) for p in idx])
        A /= np.linalg.norm(A, axis=1, keepdims=True)
        B = cents[:, :D]
        nb = np.linalg.norm(B, axis=1)
        B = B / np.where(nb > 0, nb, 1.0)[:, None]
        out[idx] = 0.5 * (1.0 + (A @ B.T) ** 2)
    return out


def run_swap_tests(qe, points, centroids, iterations, warmup, seed):
    """Batched QKMeans: amplitude encoding + swap tests of every point against
    every centroid, `iterations` k-means iterations, FP32 registers with FP64
    sums, one captured graph per iteration. Checked iteration by iteration
    against the FP64 plan (assignments) and exact FP64 inner products
    (distances), then timed from the same initial centroids."""
    import numpy as np
    from quantumlab_amd import qkmeans as K
    from quantumlab_amd.quest import QMX_PREC_FP32, QMX_PREC_FP64
    data = K.make_dataset(num_points=points, num_clusters=centroids, q_lo=4, q_hi=10, seed=seed)
    tests = points * centroids
    p64 = K.SwapPlan(qe, data, QMX_PREC_FP64)
    p32 = K.SwapPlan(qe, data, QMX_PREC_FP32)
    same_assign, dist_err, dist_err64 = True, 0.0, 0.0
    for it in range(iterations):
        cents = p64.centroids()
        exact = exact_p0(data, cents)
        p64.run(1, graph=False)
        p32.run(1, graph=False)
        a64, a32 = p64.assignments(), p32.assignments()
        same_assign &= bool(np.array_equal(a64, a32))
        # distance the swap test reports: <a|b>^2 = 2 P0 - 1
        dist_err = max(dist_err, float(np.max(np.abs(2 * p32.p0() - 2 * exact))))
        dist_err64 = max(dist_err64, float(np.max(np.abs(2 * p64.p0() - 2 * exact))))
    final64 = (p64.assignments(), p64.p0())
    p32.close()
    plan = K.SwapPlan(qe, data, QMX_PREC_FP32)
    timings = plan.autotune()
    rows = {q: plan.rows_per_tile(q) for q in range(4, 11)}
    plan.run(warmup, graph=True)
    plan.set_centroids(data.init)
    qe.qmxSync()
    qe.qmxTimerBegin()
    t0 = time.perf_counter()
    plan.run(iterations, graph=True)
    qe.qmxSync()
    wall = time.perf_counter() - t0
    ev_ms = qe.qmxTimerEndMs()
    timed_same = bool(np.array_equal(plan.assignments(), final64[0]))
    timed_err = float(np.max(np.abs(2 * plan.p0() - 2 * final64[1])))
    # the swap-test tile kernel alone: MFMA blocks per launch / device time
    blocks = int(qe.lib.qmxSwapPlanMfmaBlocks(plan.plan))
    tile_ms = float(qe.lib.qmxSwapPlanTimeTiles(plan.plan, 5))
    tflops = blocks * 4096 / (tile_ms / 1e3) / 1e12
    valu_tflops = blocks * 2048 / (tile_ms / 1e3) / 1e12
    evals = K.swap_test_amplitude_evaluations(data)
    res = {"swap_tests_per_s": round(tests * iterations / wall, 1), "unit": "swap-tests/s",
           "workload": f"{points} points x {centroids} centroids, q in 4..10 (registers 2q+1 = 9..21 qubits), "
                       f"{iterations} k-means iterations, FP32 amplitudes + FP64 sums, one HIP graph per iteration",
           "ms_per_iteration": round(wall / iterations * 1e3, 3), "event_ms_per_iteration": round(ev_ms / iterations, 3),
           "tests_per_iteration": tests, "ancilla0_amplitudes_per_iteration": evals,
           "amplitudes_per_s": evals * iterations / wall,
           "assignments_identical_to_fp64_every_iteration": same_assign,
           "max_distance_error_vs_exact_fp64": dist_err, "max_distance_error_fp64_plan": dist_err64,
           "timed_run_assignments_identical_to_fp64": timed_same, "timed_run_max_distance_diff_vs_fp64": timed_err,
           "autotune_timings": timings, "rows_per_tile": rows, "tiles": plan.num_tiles(),
           "method": "closed-form evaluation of the circuit's ancilla-0 amplitudes (a_x b_y + a_y b_x)/2 over all "
                     "4^q index pairs, from the encoded registers; the CSWAP ladder and the Hadamards are not applied "
                     "gate by gate (the gate-level circuit is the oracle's and the cpu_reference_method's)",
           "roofline": {"bound": "valu", "kernel": "kSwapTilesMfma (32x32x2 FP32 MFMA blocks of the amplitude matrix, "
                                                   "squares summed on the VALU)",
                        "achieved": round(valu_tflops, 2), "peak": FP32_VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": round(valu_tflops / FP32_VALU_PEAK_TFLOPS, 4),
                        "traffic": profiled_traffic("qkmeans_swap", "kSwapTilesMfma")[0],
                        "traffic_source": profiled_traffic("qkmeans_swap", "kSwapTilesMfma")[1],
                        "avg_launch_ms": round(tile_ms, 4), "mfma_blocks_per_launch": blocks,
                        "valu_flops_per_block": 2048, "mfma_tflops": round(tflops, 2),
                        "mfma_frac": round(tflops / FP32_MATRIX_PEAK_TFLOPS, 4),
                        "note": "registers are L2-resident (19 MB of FP32 point registers), so HBM does not bound it; "
                                "the matrix cores form each 32x32 block of amplitudes (4096 flops) and the VALU "
                                "squares and sums its 1024 entries (one FP32 FMA each, 2048 flops), so the VALU, "
                                "at its non-packed FP32 rate (packed FP32 is not used), is the bound"}}
    plan.close()
    p64.close()
    return res, data


def cpu_swap_baseline(data, seconds_per_size=2.0):
    """The reference's own method on the CPU: each swap test built gate by gate
    (CreateSwapTestReal's list: H, Ry/CX encoders, CX + Toffoli CSWAPs, H) and
    run through QuEST 3.4's CPU backend (oracle/_ref) or the C restatement,
    one register of 2q+1 qubits per test, P(ancilla = 0) read with
    calcProbOfOutcome. Sample: per register size q = 4..10 (the workload draws q
    uniformly), tests until `seconds_per_size` is spent (3 to 200 tests);
    rate = 1 / mean over q of the per-test time."""
    import numpy as np
    from oracle import swaptest_oracle as O
    from quantumlab_amd.quest import QuEST
    ref = os.path.join(ROOT, "oracle", "_ref", "libquest_ref.so")
    port = os.path.join(ROOT, "oracle", "build", "libcpuref.so")
    kind, path = ("reference", ref) if os.path.exists(ref) else ("port", port)
    lib = QuEST(path)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    per_q, count, gates_total, err = {}, {}, 0, 0.0
    for q in range(4, 11):
        times = []
        for k, p in enumerate(np.nonzero(data.qubits == q)[0]):
            if len(times) >= 3 and (sum(times) >= seconds_per_size or len(times) >= 200):
                break
            a, b = data.point(int(p)), data.init[k % data.init.shape[0], :1 << q]
            gates = O.swap_test_gates(a, b)
            reg = lib.qureg(2 * q + 1)
            t0 = time.perf_counter()
            O.run_gates_quest(lib, reg, gates)
            p0 = lib.lib.calcProbOfOutcome(reg, 0, 0)
            times.append(time.perf_counter() - t0)
            lib.destroy(reg)
            gates_total += len(gates)
            ov = (a @ b) / (np.linalg.norm(a) * np.linalg.norm(b))
            err = max(err, abs(p0 - 0.5 * (1 + ov ** 2)))
        per_q[q] = float(np.mean(times))
        count[q] = len(times)
    mean_s = float(np.mean(list(per_q.values())))
    return {"value": round(1.0 / mean_s, 3), "unit": "swap-tests/s", "cores": threads, "kind": kind,
            "sample": f"{sum(count.values())} swap tests, {min(count.values())} to {max(count.values())} per q = 4..10 ({gates_total} QuEST calls in all) on {2 * 4 + 1}..{2 * 10 + 1}-qubit "
                      f"FP64 registers, gate by gate as the reference builds them; "
                      f"{'QuEST 3.4 CPU backend compiled from the reference sources' if kind == 'reference' else 'C restatement of it'}, "
                      f"OpenMP {threads} threads; {sum(per_q[q] * count[q] for q in per_q):.2f} s",
            "seconds_per_test_by_q": {q: round(t, 4) for q, t in per_q.items()},
            "max_p0_error_vs_closed_form": err}


def cpu_swap_same_algorithm(data, seconds=10.0, seed=7):
    """The same closed-form evaluation on the host cores: for each swap test,
    all 4^q ancilla-0 amplitudes (a_x b_y + a_y b_x)/2 squared and summed in
    FP64 (oracle/cpuref/swap_closed_form.c, OpenMP over tests), on the
    normalised point and centroid registers. Sample: tests drawn uniformly from
    the workload's 262144 (so q follows its distribution), in batches of 512
    until `seconds` is spent."""
    import ctypes as C

    import numpy as np
    from quantumlab_amd.quest import QuEST
    lib = QuEST(os.path.join(ROOT, "oracle", "build", "libcpuref.so"))
    fn = lib.lib.cpurefSwapP0ClosedForm
    dp, lp = C.POINTER(C.c_double), C.POINTER(C.c_longlong)
    fn.argtypes = [C.c_int, C.POINTER(C.c_int), dp, lp, dp, lp, dp]
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    rng = np.random.default_rng(seed)
    C_ = data.init.shape[0]
    done, spent, err = 0, 0.0, 0.0
    while spent < seconds:
        picks = rng.integers(0, len(data.qubits) * C_, size=512)
        qs = np.array([int(data.qubits[t // C_]) for t in picks], dtype=np.int32)
        a_parts, b_parts = [], []
        for t, q in zip(picks, qs):
            a = data.point(int(t // C_))
            b = data.init[int(t % C_), :1 << int(q)]
            a_parts.append(a / np.linalg.norm(a))
            nb = np.linalg.norm(b)
            b_parts.append(b / nb if nb > 0 else b)
        off = np.concatenate([[0], np.cumsum(1 << qs.astype(np.int64))[:-1]]).astype(np.int64)
        A, B = np.concatenate(a_parts), np.concatenate(b_parts)
        out = np.zeros(len(picks))
        t0 = time.perf_counter()
        fn(len(picks), qs.ctypes.data_as(C.POINTER(C.c_int)), A.ctypes.data_as(dp), off.ctypes.data_as(lp),
           B.ctypes.data_as(dp), off.ctypes.data_as(lp), out.ctypes.data_as(dp))
        spent += time.perf_counter() - t0
        done += len(picks)
        err = max(err, float(np.max(np.abs(out - 0.5 * (1 + np.array([a @ b for a, b in zip(a_parts, b_parts)]) ** 2)))))
    return {"value": round(done / spent, 1), "unit": "swap-tests/s", "cores": threads, "kind": "port",
            "sample": f"{done} swap tests drawn uniformly from the workload (q = 4..10), every ancilla-0 amplitude "
                      f"evaluated in FP64 as the GPU plan does (C/OpenMP, {threads} threads); {spent:.2f} s",
            "max_p0_error_vs_inner_product": err}


def statevec_leg(qe, args, n, N):
    """The headline: per-gate FP64 brick circuit, value, roofline."""
    import numpy as np
    from quantumlab_amd import workloads as W
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
    traffic, traffic_src = profiled_traffic("brick32_fp64", "kMatrix2<double, false, true>", n)

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
        "roofline": {"bound": "hbm", "kernel": "compactUnitary 2x2 sweep: kMatrix2Lane<double> (targets 0-5), kMatrix2Span<double> (6-10), kMatrix2<double> (11-31)",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "avg_launch_ms": round(float(cu_ms.mean()), 4), "launches": int(len(cu_ms)),
                     "algorithmic_bytes_per_launch": cu_bytes,
                     "traffic_source": traffic_src,
                     "per_target_ms_min": round(min(tmeans.values()), 4),
                     "per_target_ms_max": round(max(tmeans.values()), 4)},
        "event_time_s": round(ev_ms / 1e3, 4),
    }
    log(f"fp64 per-gate: {value:.1f} GB/s, {result['gates_per_s']:.1f} gates/s, compactUnitary "
        f"{cu_ms.mean():.3f} ms avg -> {achieved:.1f} GB/s ({achieved / HBM_PEAK_GBS:.1%} of peak)")
    return result, q64, plans, timed, wall


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=32)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--fuse-qubits", type=int, default=5)
    ap.add_argument("--fp32-qubits", type=int, default=33,
                    help="register size of the large FP32 fused run (0: skip)")
    ap.add_argument("--cpu-qubits", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fused", action="store_true")
    ap.add_argument("--no-swap", action="store_true")
    ap.add_argument("--no-density", action="store_true")
    ap.add_argument("--no-trotter", action="store_true")
    ap.add_argument("--trotter-qubits", type=int, default=24)
    ap.add_argument("--trotter-count", type=int, default=128)
    ap.add_argument("--trotter-step-doubling", type=int, default=16,
                    help="instances (the first ones) that also run the step-doubling baselines")
    ap.add_argument("--density-qubits", type=int, default=14)
    ap.add_argument("--density-layers", type=int, default=60)
    ap.add_argument("--cpu-density-qubits", type=int, default=13)
    ap.add_argument("--swap-points", type=int, default=16384)
    ap.add_argument("--swap-centroids", type=int, default=16)
    ap.add_argument("--swap-iterations", type=int, default=10)
    ap.add_argument("--only", choices=["statevec", "fused", "fused-large", "fused-kernel", "density", "trotter", "swap",
                                       "stream"],
                    help="run one leg (for profiling it alone); 'fused' includes the FP64 per-gate run it is "
                         "compared with, 'fused-kernel' is the fused FP32 periods alone, the others skip the "
                         "headline leg; no CPU baselines")
    args = ap.parse_args()
    if args.only:
        args.no_cpu = True
        args.no_fused = args.only not in ("fused", "fused-large")
        args.no_density = args.only != "density"
        args.no_trotter = args.only != "trotter"
        args.no_swap = args.only != "swap"
        if args.only == "fused-large":
            args.fused_large_only = True
        elif args.only not in ("statevec", "fused"):
            args.fused_steps = args.steps
            args.steps = 0
        if args.only == "statevec":
            args.fp32_qubits = 0
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (the state-vector "
              f"path here has no multi-GPU sharding: one register lives on one GPU)", file=sys.stderr)
        sys.exit(2)

    if not args.no_cpu:
        # OpenMP reads these when its runtime loads (with the first CPU library)
        os.environ.setdefault("OMP_PROC_BIND", "close")
        os.environ.setdefault("OMP_PLACES", "cores")

    import numpy as np
    import quantumlab_amd
    from quantumlab_amd import workloads as W

    qe = quantumlab_amd.load()
    buf = __import__("ctypes").create_string_buffer(512)
    qe.qmxDeviceInfo(buf, 512)
    log(buf.value.decode())
    n = args.qubits
    N = 1 << n

    if args.steps > 0 and not getattr(args, "fused_large_only", False):
        result, q64, plans, timed, wall = statevec_leg(qe, args, n, N)
    else:
        # profiling runs of the other legs only: no headline value
        q64 = None
        result = {"metric": METRIC, "value": None, "unit": "GB/s", "n_gpus": 1, "steps": 0,
                  "warmup": args.warmup, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
                  "vs_baseline": None, "dtype": "f64", "data": "synthetic",
                  "config": {"workload": "state-vector leg skipped (--steps 0 or --only)", "qubits": n}}

    if not args.no_fused and q64 is not None and not getattr(args, "fused_large_only", False):
        q32, wall32, blocks, norm2, fbytes, fsweeps = run_fp32_fused(qe, n, plans, args.warmup, args.steps,
                                                                     args.fuse_qubits)
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
        fgbs = fbytes / wall32 / 1e9
        result["fused_fp32"]["full_sweeps_per_step"] = fsweeps / args.steps
        result["fused_fp32"]["roofline"] = {
            "bound": "hbm", "kernel": "kFused5Tile (5-qubit blocks on LDS tiles, v_mfma_f32_16x16x4f32, "
                                      "three real products per complex product)", "achieved": round(fgbs, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(fgbs / HBM_PEAK_GBS, 4),
            "traffic": profiled_traffic("fused32_fp32", "kFused5Tile", n)[0],
            "traffic_source": profiled_traffic("fused32_fp32", "kFused5Tile", n)[1],
            "bytes_per_block": 2.0 * N * 8,
            "mfma_tflops": round(fsweeps * N * 192 / wall32 / 1e12, 2),
            "note": "a block of several gates (padded to 5 qubits) reads and writes the FP32 register once, with "
                    "192 MFMA flops per amplitude; a lone gate on fewer qubits runs its per-gate kernel and counts "
                    "what that moves; over the fused steps' wall time (launch gaps included)"}
        log(f"fused fp32: {wall / wall32:.2f}x vs fp64 per-gate, 1-F={1 - fid:.3e}, max|da|={maxabs:.3e}")
        qe.destroy(q32)
    if q64 is not None:
        qe.destroy(q64)
    if not args.no_fused and (args.steps > 0 or getattr(args, "fused_large_only", False)) and args.fp32_qubits > 0:
        result["fused_fp32_large"] = run_fp32_large(qe, args.fp32_qubits, args.seed, args.fuse_qubits)

    if args.only == "stream":
        result["stream"] = run_stream(qe, n)
    if args.only == "fused-kernel":
        plans = [period_gates(n, s, args.seed) for s in range(args.warmup + args.fused_steps)]
        q32, wall32, blocks, _, fbytes, fsweeps = run_fp32_fused(qe, n, plans, args.warmup, args.fused_steps,
                                                                 args.fuse_qubits)
        qe.destroy(q32)
        gbs = fbytes / wall32 / 1e9
        result["fused_kernel"] = {"qubits": n, "periods": args.fused_steps, "blocks": blocks,
                                  "full_sweeps": fsweeps, "ms_per_block": round(wall32 / blocks * 1e3, 3),
                                  "GBps": round(gbs, 1)}

    if not args.no_density:
        result["density"] = run_density(qe, args.density_qubits, args.density_layers, args.seed)

    trotter_rows = None
    if not args.no_trotter:
        result["trotter"], trotter_rows = run_trotter(qe, args.trotter_qubits, args.trotter_count, args.seed,
                                                    sd_count=args.trotter_step_doubling)
        tr = result["trotter"]
        log(f"trotter: learned {tr['learned_ms_mean']} ms vs step doubling o2 {tr['step_doubling_o2_ms_mean']} ms, "
            f"o4 {tr['step_doubling_o4_ms_mean']} ms; success {tr['success_rate']:.2f}")

    swap_data = None
    if not args.no_swap:
        result["swap_tests"], swap_data = run_swap_tests(qe, args.swap_points, args.swap_centroids,
                                                         args.swap_iterations, args.warmup, args.seed)
        st = result["swap_tests"]
        log(f"swap tests: {st['swap_tests_per_s']:.4g}/s, {st['ms_per_iteration']} ms/iteration, "
            f"distance err {st['max_distance_error_vs_exact_fp64']:.2e}, "
            f"assignments identical {st['assignments_identical_to_fp64_every_iteration']}")

    if not args.no_cpu:
        log(f"cpu baseline at {args.cpu_qubits} qubits ...")
        result["cpu_baseline"] = cpu_baseline(args.cpu_qubits, args.seed)
        log(f"cpu baseline: {result['cpu_baseline']['value']} GB/s")
        if "density" in result:
            result["density"]["cpu_baseline"] = cpu_density_baseline(args.cpu_density_qubits, 2, args.seed)
            log(f"cpu density baseline: {result['density']['cpu_baseline']['value']} GB/s")
        if trotter_rows is not None:
            result["trotter"]["cpu_baseline"] = cpu_trotter_baseline(trotter_rows, args.trotter_qubits)
            log(f"cpu trotter baseline: {result['trotter']['cpu_baseline']['value']} ms per solve")
        if swap_data is not None:
            result["swap_tests"]["cpu_baseline"] = cpu_swap_same_algorithm(swap_data)
            log(f"cpu swap-test baseline (same closed form): {result['swap_tests']['cpu_baseline']['value']} "
                f"swap-tests/s")
            result["swap_tests"]["cpu_reference_method"] = cpu_swap_baseline(swap_data)
            log(f"cpu swap-test baseline (reference gate-level method): "
                f"{result['swap_tests']['cpu_reference_method']['value']} swap-tests/s")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
