"""Shared seeded test cases, applied identically to every QuEST-ABI library
(the MI355X library, the oracle restatement and the reference build) so their
outputs can be compared element by element. Modelled on QuEST's own tests
(QuEST340/tests/test_unitaries.cpp etc.), which act on a small register
initialised with initDebugState."""
from __future__ import annotations

import ctypes as C

import numpy as np

from quantumlab_amd import workloads as W
from quantumlab_amd.quest import QuEST


def ints(xs):
    xs = list(xs)
    return (C.c_int * len(xs))(*xs)


def random_state(n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    return v / np.linalg.norm(v)


def random_hamil(n: int, terms: int, seed: int):
    rng = np.random.default_rng(seed)
    coeffs = rng.standard_normal(terms)
    codes = rng.integers(0, 4, size=(terms, n)).astype(np.int32)
    return coeffs, codes


def pauli_rotation_cases(n: int, seed: int):
    rng = np.random.default_rng(seed)
    cases = []
    for _ in range(12):
        k = int(rng.integers(1, n + 1))
        targs = [int(x) for x in rng.choice(n, size=k, replace=False)]
        paulis = [int(x) for x in rng.integers(0, 4, size=k)]
        cases.append((targs, paulis, float(rng.uniform(-np.pi, np.pi))))
    cases.append(([0], [0], 0.3))  # identity string: no-op
    return cases


def misc_ops(qe: QuEST, q, n: int):
    """A fixed sequence over most of the gate API (needs n >= 3)."""
    rng = np.random.default_rng(11)
    t = qe.t  # struct types of this library's qreal
    a, b = W.compact_pair(rng)
    u2 = W.haar_unitary(rng, 2)
    u2b = W.haar_unitary(rng, 2)
    u2c = W.haar_unitary(rng, 2)
    u4 = W.haar_unitary(rng, 4)
    u4b = W.haar_unitary(rng, 4)
    u8 = W.haar_unitary(rng, 8)
    qe.call("hadamard", q, 0)
    qe.call("pauliX", q, 1)
    qe.call("pauliY", q, 2)
    qe.call("pauliZ", q, 0)
    qe.call("sGate", q, 1)
    qe.call("tGate", q, 2)
    qe.call("phaseShift", q, 0, 0.3)
    qe.call("controlledPhaseShift", q, 0, 2, 0.7)
    qe.call("controlledPhaseFlip", q, 1, 2)
    qe.call("rotateX", q, 0, 0.1)
    qe.call("rotateY", q, 1, 0.2)
    qe.call("rotateZ", q, 2, 0.3)
    qe.call("rotateAroundAxis", q, 1, 0.4, t.Vector(1.0, 2.0, 3.0))
    qe.call("controlledRotateX", q, 0, 1, 0.5)
    qe.call("controlledRotateY", q, 1, 2, 0.6)
    qe.call("controlledRotateZ", q, 2, 0, 0.7)
    qe.call("controlledCompactUnitary", q, 0, 2, t.cplx(a), t.cplx(b))
    qe.call("unitary", q, 1, t.matrix2(u2))
    qe.call("controlledUnitary", q, 2, 0, t.matrix2(u2b))
    qe.call("multiStateControlledUnitary", q, ints([0, 2]), ints([0, 1]), 2, 1, t.matrix2(u2c))
    qe.call("swapGate", q, 0, 2)
    qe.call("twoQubitUnitary", q, 0, 2, t.matrix4(u4))
    qe.call("multiControlledMultiQubitNot", q, ints([1]), 1, ints([0, 2]), 2)
    qe.call("multiQubitNot", q, ints([0, 1]), 2)
    qe.multiRotateZ(q, [0, 1, 2], 0.9)
    qe.call("controlledNot", q, 2, 1)
    qe.call("multiControlledPhaseShift", q, ints([0, 1, 2]), 3, 1.1)
    qe.call("multiControlledPhaseFlip", q, ints([0, 2]), 2)
    qe.multiQubitUnitary(q, [2, 0, 1], u8)
    m = qe.matrixN(u4b)
    qe.call("controlledMultiQubitUnitary", q, n - 1, ints([0, 1]), 2, m)
    qe.lib.destroyComplexMatrixN(m)
    if n >= 4:
        qe.multiControlledUnitary(q, [0, 1, 3], 2, u2)


def channel_circuit(qe: QuEST, rho, pure, n: int, layers: int, seed: int):
    """Brick layers on a density matrix with dephasing, depolarising and damping
    after every gate on each qubit it touches; the same gates, noiselessly, on
    `pure` (a state vector)."""
    gates = W.noisy_brick_circuit(n, layers, seed)
    W.apply_gates(qe, rho, gates)
    W.apply_gates(qe, pure, W.unitary_part(gates))