"""The hot path at the workloads' full sizes, against the oracle.

- 32 qubits FP64 (64 GiB) and 33 qubits FP32 (64 GiB): gates of every kernel
  family (compactUnitary, controlledNot, the 3-control unitary, the 5-target
  unitary, rotateZ; at 33 qubits also a fused FP32 MFMA period) act on six
  qubits that include the top ones, so amplitude indices pass 2^32. The state
  stays in the span of those qubits: its 64 amplitudes there must equal the
  oracle's 6-qubit run of the relabelled gates, and the total probability must
  be 1 (so nothing leaks elsewhere).
- C3 at its full 14 qubits (a 2^28-entry density matrix), 3 noisy brick
  layers, FP64, every entry within 1e-12 of the oracle.
- One 24-qubit FP64 repetition of a Hamiltonian of the Trotter workload
  through the fused coset sweeps, within 1e-12 of the oracle.
- Outcome histograms of all 20 qubits of a 20-qubit register in one pass."""
import time

import numpy as np
import pytest

from quantumlab_amd import hamiltonians as H
from quantumlab_amd import workloads as W
from quantumlab_amd.quest import QMX_PREC_FP32, cplx, matrix2
from tests.cases import ints

pytestmark = pytest.mark.gpu


def _span_gates(S, seed):
    """A gate list on the qubits S (S[0] lowest), with relabelled copies for a
    len(S)-qubit register (qubit S[j] -> j)."""
    rng = np.random.default_rng(seed)
    big, small = [], []
    j = {q: i for i, q in enumerate(S)}

    def add(g):
        big.append(g)
        k = g[0]
        if k == "cu":
            small.append((k, j[g[1]], g[2], g[3]))
        elif k == "cnot":
            small.append((k, j[g[1]], j[g[2]]))
        elif k == "rz":
            small.append((k, j[g[1]], g[2]))
        elif k == "mcu":
            small.append((k, [j[c] for c in g[1]], j[g[2]], g[3]))
        elif k == "mqu":
            small.append((k, [j[t] for t in g[1]], g[2]))

    for q in S:
        a, b = W.compact_pair(rng)
        add(("cu", q, a, b))
    top, low = S[-1], S[0]
    add(("cnot", top, low))
    add(("cnot", low, top))
    add(("cnot", S[-2], top))
    add(("cnot", S[1], S[3]))
    add(("mcu", [top, S[2], low], S[-2], W.haar_unitary(rng, 2)))
    add(("mqu", [top, low, S[3], S[2], S[-2]], W.haar_unitary(rng, 32)))
    add(("rz", top, float(rng.uniform(0, 2 * np.pi))))
    for q in S:
        a, b = W.compact_pair(rng)
        add(("cu", q, a, b))
    add(("mcu", [S[1], S[3], S[4]], top, W.haar_unitary(rng, 2)))
    return big, small


def _span_amps(gpu, q, S):
    out = np.zeros(1 << len(S), dtype=complex)
    for v in range(1 << len(S)):
        idx = 0
        for b, qb in enumerate(S):
            if (v >> b) & 1:
                idx |= 1 << qb
        z = gpu.getAmp(q, idx)
        out[v] = z.real + 1j * z.imag
    return out


@pytest.mark.timeout(600)
@pytest.mark.parametrize("n,prec,tol", [(32, None, 1e-12), (33, QMX_PREC_FP32, 2e-6)])
def test_gates_past_two_to_the_32_match_oracle(gpu, cpuref, n, prec, tol):
    S = [0, 2, 9, 20, n - 2, n - 1]
    big, small = _span_gates(S, seed=n)
    c = None
    if prec == QMX_PREC_FP32:
        # one fused period on the same qubits (5-qubit MFMA blocks padded from the top)
        fb, fs = _span_gates(S, seed=100 + n)
        c = W.build_circuit(gpu, n, fb)
        gpu.qmxCircuitFuse(c, 5)
        small = small + fs
    qs = cpuref.qureg(len(S))
    W.apply_gates(cpuref, qs, small)
    want = cpuref.state(qs)
    cpuref.destroy(qs)
    print(f"[full-size] {n} qubits, precision {'fp32' if prec else 'fp64'}", flush=True)
    q = gpu.qureg(n, precision=prec)
    W.apply_gates(gpu, q, big)
    if c is not None:
        gpu.qmxCircuitApply(q, c, 1)
        gpu.qmxCircuitDestroy(c)
    got = _span_amps(gpu, q, S)
    total = gpu.calcTotalProb(q)
    gpu.destroy(q)
    assert np.max(np.abs(got - want)) < tol
    assert abs(total - 1) < (1e-12 if prec is None else 1e-5)
    assert abs(np.vdot(got, got).real - total) < (1e-12 if prec is None else 1e-5)


@pytest.mark.timeout(1200)
def test_noisy_density_14_qubits_three_layers_match_oracle(gpu, cpuref):
    """C3 at full size: 14-qubit density matrix, 3 noisy brick layers, FP64."""
    n = 14
    gates = W.noisy_brick_circuit(n, 3, seed=314)
    print(f"[full-size] density {n} qubits, {len(gates)} operations", flush=True)
    rho = gpu.qureg(n, density=True)
    W.apply_gates(gpu, rho, gates)
    purity = gpu.calcPurity(rho)
    p0 = [gpu.calcProbOfOutcome(rho, k, 0) for k in range(n)]
    got = gpu.state(rho)
    gpu.destroy(rho)
    ref = cpuref.qureg(n, density=True)
    t0 = time.time()
    W.apply_gates(cpuref, ref, gates)
    print(f"[full-size] oracle took {time.time() - t0:.1f} s", flush=True)
    want = cpuref.state(ref)
    assert abs(purity - cpuref.calcPurity(ref)) < 1e-12
    assert max(abs(p0[k] - cpuref.calcProbOfOutcome(ref, k, 0)) for k in range(n)) < 1e-12
    cpuref.destroy(ref)
    err = 0.0
    for s in range(0, got.size, 1 << 24):
        err = max(err, float(np.max(np.abs(got[s:s + (1 << 24)] - want[s:s + (1 << 24)]))))
    assert err < 1e-12


@pytest.mark.timeout(600)
@pytest.mark.parametrize("order", [1, 2])
def test_trotter_24_qubits_fp64_coset_matches_oracle(gpu, cpuref, order):
    """A Hamiltonian of the 24-qubit Trotter workload, one repetition, through
    the coset-form tile sweeps (FP64 tiles of 11 qubits)."""
    inst = H.trotter_instance(24, seed=2024, index=order)
    hb = H.HamilBatch([inst])
    outs = []
    for lib in (gpu, cpuref):
        q = lib.qureg(24)
        H.prepare(lib, q, inst)
        lib.lib.applyTrotterCircuit(q, hb[0], inst.time / 8, order, 1)
        outs.append(lib.state(q))
        lib.destroy(q)
    assert np.max(np.abs(outs[0] - outs[1])) < 1e-12


@pytest.mark.parametrize("density", [False, True])
def test_all_outcomes_of_twenty_qubits_in_one_pass(gpu, density):
    """calcProbOfAllOutcomes over all 20 qubits (2^20 outcomes) of a 20-qubit
    state vector (or a 10-qubit density matrix over its 10 qubits), qubits in a
    scrambled order: equal to the permuted |amplitude|^2 (diagonal), and fast."""
    from tests.cases import random_state
    n = 10 if density else 20
    rng = np.random.default_rng(7)
    qubits = [int(x) for x in rng.permutation(n)]
    q = gpu.qureg(n, density=density)
    if density:
        psi = random_state(n, 8)
        pure = gpu.qureg(n)
        gpu.set_state(pure, psi)
        gpu.initPureState(q, pure)
        gpu.destroy(pure)
        probs = np.abs(psi) ** 2
    else:
        psi = random_state(n, 9)
        gpu.set_state(q, psi)
        probs = np.abs(psi) ** 2
    gpu.calcProbOfAllOutcomes(q, qubits)  # warm-up
    t0 = time.perf_counter()
    got = gpu.calcProbOfAllOutcomes(q, qubits)
    dt = time.perf_counter() - t0
    idx = np.arange(1 << n)
    out_index = np.zeros(1 << n, dtype=np.int64)
    for j, qb in enumerate(qubits):
        out_index |= ((idx >> qb) & 1) << j
    want = np.zeros(1 << n)
    want[out_index] = probs
    assert np.max(np.abs(got - want)) < 1e-15
    assert dt < 0.5, dt
    gpu.destroy(q)
