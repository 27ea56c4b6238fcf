"""Operation lists at the circuit-capture boundary (qmxCircuit*, qmxApplyOperations)
against the oracle of QuantumLab's own loop (oracle/ops_oracle.py: one QuEST
call per EBasicOperation, QLGate.cpp:652-803, and the density simulator's
after-gate noise, QLSimulatorDensityMatrix.cpp:88-153) run on the C restatement
of QuEST's CPU backend.

CPU tests: record-time validation (no GPU). GPU tests: FP64 state vectors and
density matrices within 1e-12 of the oracle, fused or not; FP32 fused MFMA
blocks within FP32 resolution; collapse probabilities equal."""
import ctypes as C

import numpy as np
import pytest

from oracle import ops_oracle as O
from quantumlab_amd import workloads as W
from quantumlab_amd.quest import QMX_PREC_FP32, BasicOperation, NoiseAfterGate, QuESTError
from tests.cases import ints

TOL = 1e-12


def _lib_cpu():
    import os

    import quantumlab_amd
    lib = C.CDLL(quantumlab_amd.library_path(), mode=os.RTLD_LOCAL)
    lib.qmxSetErrorMode(1)
    lib.qmxCircuitCreate.restype = C.c_void_p
    lib.qmxCircuitCreate.argtypes = [C.c_int]
    lib.qmxCircuitDestroy.argtypes = [C.c_void_p]
    lib.qmxCircuitNumGates.argtypes = [C.c_void_p]
    lib.qmxCircuitAddOperation.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.c_int, C.c_double]
    lib.qmxCircuitFuse.argtypes = [C.c_void_p, C.c_int]
    lib.qmxCircuitSetNoiseAfterGate.argtypes = [C.c_void_p, NoiseAfterGate]
    lib.qmxLastError.argtypes = [C.c_char_p, C.c_int]
    lib.qmxClearError()
    return lib


def _last_error(lib):
    buf = C.create_string_buffer(1024)
    had = lib.qmxLastError(buf, 1024)
    lib.qmxClearError()
    return buf.value.decode() if had else None


@pytest.mark.parametrize("code,qubits,param,substr", [
    (O.H, [5], 0.0, "Invalid target"),
    (O.H, [-1], 0.0, "Invalid target"),
    (O.CX, [2, 2], 0.0, "Control qubit cannot equal target"),
    (O.CRZ, [7, 1], 0.3, "Invalid control"),
    (O.CX, [1], 0.0, "exactly two qubits"),
    (O.RX, [0, 1], 0.1, "exactly one qubit"),
    (O.CC, [1], 2.0, "Invalid measurement outcome"),
    (O.NOISE_DEPHASING, [1], 0.6, "cannot exceed 1/2"),
    (O.NOISE_DEPOLARISING, [1], 0.8, "cannot exceed 3/4"),
    (O.NOISE_TWO_DEPOLARISING, [1, 1], 0.1, "target qubits must be unique"),
    (O.NOISE_DAMPING, [0], -0.1, "Probabilities must be in [0, 1]"),
    (99, [0], 0.0, "Invalid operation code"),
])
def test_circuit_records_reject_invalid_operations(code, qubits, param, substr):
    lib = _lib_cpu()
    c = lib.qmxCircuitCreate(5)
    lib.qmxCircuitAddOperation(c, code, ints(qubits), len(qubits), param)
    err = _last_error(lib)
    assert err is not None and substr in err, err
    assert lib.qmxCircuitNumGates(c) == 0
    lib.qmxCircuitDestroy(c)


def test_circuit_width_and_block_size_validated():
    lib = _lib_cpu()
    assert lib.qmxCircuitCreate(0) is None
    assert "Invalid number of qubits" in _last_error(lib)
    c = lib.qmxCircuitCreate(4)
    lib.qmxCircuitAddOperation(c, O.H, ints([3]), 1, 0.0)
    lib.qmxCircuitAddOperation(c, O.CX, ints([3, 0]), 2, 0.0)
    assert _last_error(lib) is None and lib.qmxCircuitNumGates(c) == 2
    for bad in (0, 11):
        lib.qmxCircuitFuse(c, bad)
        assert "Invalid block size" in _last_error(lib)
    assert lib.qmxCircuitFuse(c, 2) == 1
    lib.qmxCircuitSetNoiseAfterGate(c, NoiseAfterGate(0.1, 0.7, 0, 0, 0, 0))
    assert "cannot exceed 1/2" in _last_error(lib)
    lib.qmxCircuitDestroy(c)


# ---- GPU parity --------------------------------------------------------------------------

def resolve_collapses(qe, q, ops, noise=None):
    """Run ops on the oracle, fixing each collapse's outcome to the likelier one
    (so its probability is well away from 0); returns the completed list."""
    out = []
    for code, qs, t in ops:
        if code == O.CC and t < 0:
            p0 = qe.calcProbOfOutcome(q, qs[0], 0)
            t = 0.0 if p0 >= 0.5 else 1.0
        op = (code, qs, t)
        out.append(op)
        O.perform(qe, q, op)
        if noise and q.isDensityMatrix:
            O.noise_after_gate(qe, q, op, noise)
    return out


def record(gpu, n, ops, noise=None):
    c = gpu.lib.qmxCircuitCreate(n)
    for code, qs, t in ops:
        gpu.call("qmxCircuitAddOperation", c, code, ints(qs), len(qs), t)
    if noise:
        gpu.call("qmxCircuitSetNoiseAfterGate", c, NoiseAfterGate(**noise))
    return c


def oracle_state(cpuref, n, ops, density=False, noise=None, init=None):
    q = cpuref.qureg(n, density=density)
    if init is not None:
        init(cpuref, q)
    prob = O.run_ops(cpuref, q, ops, noise)
    st = cpuref.state(q)
    cpuref.destroy(q)
    return st, prob


def _init_plus(lib, q):
    lib.initPlusState(q)


@pytest.mark.gpu
@pytest.mark.parametrize("fuse", [0, 3, 5])
def test_state_vector_op_list_matches_oracle(gpu, cpuref, fuse):
    n = 12
    raw = O.random_ops(n, 400, seed=21, noise=True, collapses=3)
    raw.insert(7, (O.PHASE, [2], 0.4))
    q0 = cpuref.qureg(n)
    cpuref.initPlusState(q0)
    ops = resolve_collapses(cpuref, q0, raw)
    cpuref.destroy(q0)
    ref, pref = oracle_state(cpuref, n, ops, init=_init_plus)
    c = record(gpu, n, ops)
    q = gpu.qureg(n)
    gpu.initPlusState(q)
    if fuse:
        assert gpu.qmxCircuitFuse(c, fuse) > 0
    prob = gpu.qmxCircuitApply(q, c, 1 if fuse else 0)
    assert abs(prob - pref) < TOL
    assert np.max(np.abs(gpu.state(q) - ref)) < TOL
    gpu.destroy(q)
    gpu.qmxCircuitDestroy(c)


@pytest.mark.gpu
def test_apply_operations_entry_matches_circuit(gpu, cpuref):
    n = 10
    q0 = cpuref.qureg(n)
    ops = resolve_collapses(cpuref, q0, O.random_ops(n, 150, seed=5, collapses=2))
    cpuref.destroy(q0)
    ref, pref = oracle_state(cpuref, n, ops)
    arr = (BasicOperation * len(ops))()
    for i, (code, qs, t) in enumerate(ops):
        arr[i].operation = code
        arr[i].numQubits = len(qs)
        for j, x in enumerate(qs):
            arr[i].qubits[j] = x
        arr[i].param = t
    for fuse in (0, 4):
        q = gpu.qureg(n)
        prob = gpu.qmxApplyOperations(q, arr, len(ops), None, fuse)
        assert abs(prob - pref) < TOL
        assert np.max(np.abs(gpu.state(q) - ref)) < TOL
        gpu.destroy(q)


NOISES = [
    dict(damping=0.03, dephasing=0.02, depolarising=0.01, mixPauli=0.005),
    dict(damping=0.01, dephasing=0.02, depolarising=0.01, twoQubitDephasing=0.03, twoQubitDepolarising=0.02),
]


@pytest.mark.gpu
@pytest.mark.parametrize("noise_idx", [0, 1])
def test_density_op_list_with_noise_matches_oracle(gpu, cpuref, noise_idx):
    """Noise operations and the simulator's after-gate noise on a density register."""
    n = 6
    noise = NOISES[noise_idx]
    q0 = cpuref.qureg(n, density=True)
    ops = resolve_collapses(cpuref, q0, O.random_ops(n, 120, seed=30 + noise_idx, noise=True, collapses=2),
                            noise)
    cpuref.destroy(q0)
    ref, pref = oracle_state(cpuref, n, ops, density=True, noise=noise)
    c = record(gpu, n, ops, noise)
    q = gpu.qureg(n, density=True)
    prob = gpu.qmxCircuitApply(q, c, 1)  # fused on a density register: the ops in order
    assert abs(prob - pref) < TOL
    assert np.max(np.abs(gpu.state(q) - ref)) < TOL
    gpu.destroy(q)
    gpu.qmxCircuitDestroy(c)


@pytest.mark.gpu
def test_circuit_apply_rejects_other_widths(gpu):
    c = record(gpu, 6, [(O.H, [5], 0.0)])
    q = gpu.qureg(5)
    with pytest.raises(QuESTError, match="different numbers of qubits"):
        gpu.qmxCircuitApply(q, c, 0)
    gpu.destroy(q)
    q = gpu.qureg(3, density=True)
    with pytest.raises(QuESTError, match="different numbers of qubits"):
        gpu.qmxCircuitApply(q, c, 1)
    gpu.destroy(q)
    gpu.qmxCircuitDestroy(c)


@pytest.mark.gpu
@pytest.mark.parametrize("n,k", [(10, 5), (14, 5), (16, 5), (16, 3)])
def test_fused_brick_fp64_matches_oracle(gpu, cpuref, n, k):
    """Brick circuit with the 3-control and 5-target extras, fused into dense
    FP64 blocks (5-qubit blocks on the FP64 MFMA tiles from 11 qubits, kDense
    for 2-4 qubits and below 11), against the oracle's per-gate run."""
    gates = W.brick_circuit(n, 10, seed=40 + n, extra_every=5)
    qc = cpuref.qureg(n)
    W.apply_gates(cpuref, qc, gates)
    ref = cpuref.state(qc)
    cpuref.destroy(qc)
    c = W.build_circuit(gpu, n, gates)
    assert gpu.qmxCircuitFuse(c, k) < len(gates)
    q = gpu.qureg(n)
    gpu.qmxCircuitApply(q, c, 1)
    assert np.max(np.abs(gpu.state(q) - ref)) < TOL
    gpu.destroy(q)
    gpu.qmxCircuitDestroy(c)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [14, 18])
def test_fused_brick_fp32_mfma_matches_oracle(gpu, cpuref, n):
    """The same fused into 5-qubit MFMA blocks on FP32 amplitudes (kFused5):
    max-abs amplitude error <= 2e-6 and 1 - F <= 1e-6 against the oracle."""
    gates = W.brick_circuit(n, 10, seed=60 + n, extra_every=5)
    qc = cpuref.qureg(n)
    W.apply_gates(cpuref, qc, gates)
    ref = cpuref.state(qc)
    cpuref.destroy(qc)
    c = W.build_circuit(gpu, n, gates)
    gpu.qmxCircuitFuse(c, 5)
    q = gpu.qureg(n, precision=QMX_PREC_FP32)
    gpu.qmxCircuitApply(q, c, 1)
    got = gpu.state(q)
    assert np.max(np.abs(got - ref)) < 2e-6
    f = abs(np.vdot(ref, got)) ** 2 / (np.vdot(got, got).real * np.vdot(ref, ref).real)
    assert 1 - f < 1e-6
    gpu.destroy(q)
    gpu.qmxCircuitDestroy(c)
