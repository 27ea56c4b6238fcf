"""Seeded synthetic workloads (gate lists) for the state-vector hot path.

Brick circuit (the reference configurations' shape): each layer applies
compactUnitary on every qubit with Haar-random (alpha, beta) — |alpha|^2 ~ U(0,1),
arg alpha, arg beta ~ U(0, 2pi) — then controlledNot on pairs (i, i+1), even i
on even layers and odd i on odd layers, then rotateZ(theta ~ U(0, 2pi)) on one
uniformly random qubit. Optionally, after every `extra_every`-th layer it adds
multiControlledUnitary (3 distinct random controls, a target outside them,
Haar 2x2) and multiQubitUnitary on 5 distinct random qubits (Haar 32x32: QR of
a complex Ginibre matrix with R's diagonal phases divided out, Mezzadri 2007).

A gate list is a list of tuples:
    ("cu", target, alpha, beta)          compactUnitary
    ("cnot", control, target)            controlledNot
    ("rz", target, theta)                rotateZ
    ("mcu", controls, target, u2x2)      multiControlledUnitary
    ("mqu", targets, u)                  multiQubitUnitary (targets[0] = LSB)
    ("deph", target, p)                  mixDephasing       (density matrices)
    ("depol", target, p)                 mixDepolarising
    ("damp", target, gamma)              mixDamping
The same list feeds the MI355X library, the oracle and the reference build.

Noisy brick circuit (density-matrix workload): the brick circuit with
mixDephasing(p ~ U(0, 0.02)), mixDepolarising(p ~ U(0, 0.01)) and
mixDamping(gamma ~ U(0, 0.02)) after every gate on each qubit it touches.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Sequence

import numpy as np

from .quest import QuEST


def haar_unitary(rng: np.random.Generator, dim: int) -> np.ndarray:
    z = (rng.standard_normal((dim, dim)) + 1j * rng.standard_normal((dim, dim))) / np.sqrt(2.0)
    q, r = np.linalg.qr(z)
    d = np.diagonal(r)
    return q * (d / np.abs(d))[np.newaxis, :]


def compact_pair(rng: np.random.Generator):
    u = rng.uniform(0.0, 1.0)
    pa, pb = rng.uniform(0.0, 2 * np.pi, size=2)
    alpha = np.sqrt(u) * np.exp(1j * pa)
    beta = np.sqrt(1.0 - u) * np.exp(1j * pb)
    return complex(alpha), complex(beta)


def brick_layer(rng: np.random.Generator, n: int, layer: int) -> List[tuple]:
    gates = []
    for q in range(n):
        a, b = compact_pair(rng)
        gates.append(("cu", q, a, b))
    start = 0 if layer % 2 == 0 else 1
    for i in range(start, n - 1, 2):
        gates.append(("cnot", i, i + 1))
    gates.append(("rz", int(rng.integers(0, n)), float(rng.uniform(0.0, 2 * np.pi))))
    return gates


def extra_gates(rng: np.random.Generator, n: int) -> List[tuple]:
    qs = rng.choice(n, size=4, replace=False)
    ctrls, target = [int(x) for x in qs[:3]], int(qs[3])
    mcu = ("mcu", ctrls, target, haar_unitary(rng, 2))
    targs = [int(x) for x in rng.choice(n, size=5, replace=False)]
    mqu = ("mqu", targs, haar_unitary(rng, 32))
    return [mcu, mqu]


def brick_circuit(n: int, layers: int, seed: int, extra_every: int = 0, first_layer: int = 0) -> List[tuple]:
    """Layers [first_layer, first_layer + layers) of the seeded brick circuit.

    The generator is re-seeded per layer (seed, layer), so any window of layers
    is reproducible on its own and identical to the same window of a longer run.
    """
    gates: List[tuple] = []
    for layer in range(first_layer, first_layer + layers):
        rng = np.random.default_rng([seed, layer])
        gates += brick_layer(rng, n, layer)
        if extra_every and (layer + 1) % extra_every == 0 and n >= 5:
            gates += extra_gates(rng, n)
    return gates


CHANNELS = ("deph", "depol", "damp")


def noisy_brick_circuit(n: int, layers: int, seed: int) -> List[tuple]:
    rng = np.random.default_rng([seed, 99])
    out: List[tuple] = []
    for g in brick_circuit(n, layers, seed):
        out.append(g)
        touched = [g[1]] if g[0] in ("cu", "rz") else [g[1], g[2]]
        for t in touched:
            out.append(("deph", t, float(rng.uniform(0, 0.02))))
            out.append(("depol", t, float(rng.uniform(0, 0.01))))
            out.append(("damp", t, float(rng.uniform(0, 0.02))))
    return out


def unitary_part(gates: Sequence[tuple]) -> List[tuple]:
    return [g for g in gates if g[0] not in CHANNELS]


def apply_gates(qe: QuEST, q, gates: Sequence[tuple]):
    """Apply a gate list through the QuEST C API of any loaded library."""
    for g in gates:
        kind = g[0]
        if kind == "cu":
            qe.call("compactUnitary", q, g[1], qe.t.cplx(g[2]), qe.t.cplx(g[3]))
        elif kind == "cnot":
            qe.call("controlledNot", q, g[1], g[2])
        elif kind == "rz":
            qe.call("rotateZ", q, g[1], g[2])
        elif kind == "mcu":
            qe.multiControlledUnitary(q, g[1], g[2], g[3])
        elif kind == "mqu":
            qe.multiQubitUnitary(q, g[1], g[2])
        elif kind == "deph":
            qe.call("mixDephasing", q, g[1], g[2])
        elif kind == "depol":
            qe.call("mixDepolarising", q, g[1], g[2])
        elif kind == "damp":
            qe.call("mixDamping", q, g[1], g[2])
        else:
            raise ValueError(f"unknown gate {kind}")


def build_circuit(qe: QuEST, n: int, gates: Sequence[tuple]):
    """Record a gate list into a qmxCircuit of the MI355X library (fusion input)."""
    lib = qe.lib
    c = lib.qmxCircuitCreate(n)
    for g in gates:
        kind = g[0]
        if kind == "cu":
            lib.qmxCircuitCompactUnitary(c, g[1], qe.t.cplx(g[2]), qe.t.cplx(g[3]))
        elif kind == "cnot":
            lib.qmxCircuitControlledNot(c, g[1], g[2])
        elif kind == "rz":
            lib.qmxCircuitRotateZ(c, g[1], g[2])
        elif kind == "mcu":
            arr = (C.c_int * len(g[1]))(*g[1])
            lib.qmxCircuitMultiControlledUnitary(c, arr, len(g[1]), g[2], qe.t.matrix2(g[3]))
        elif kind == "mqu":
            arr = (C.c_int * len(g[1]))(*g[1])
            u = np.ascontiguousarray(g[2], dtype=np.complex128)
            re = qe.reals(u.real).ravel()
            im = qe.reals(u.imag).ravel()
            lib.qmxCircuitMultiQubitUnitary(c, arr, len(g[1]), re.ctypes.data_as(qe.real_ptr),
                                            im.ctypes.data_as(qe.real_ptr))
        else:
            raise ValueError(f"unknown gate {kind}")
    return c


def gate_bytes(gates: Sequence[tuple], n: int, amp_bytes: int) -> float:
    """Algorithmic HBM bytes of a per-gate sweep list (read + write of the
    amplitudes each gate changes): dense 1-target 2*N*B; controlledNot N*B;
    c-controlled 2*N*B/2^c; multiQubitUnitary 2*N*B."""
    N = float(1 << n)
    tot = 0.0
    for g in gates:
        if g[0] in ("cu", "rz", "mqu"):
            tot += 2 * N * amp_bytes
        elif g[0] == "cnot":
            tot += N * amp_bytes
        elif g[0] == "mcu":
            tot += 2 * N * amp_bytes / (1 << len(g[1]))
    return tot


def density_bytes(gates: Sequence[tuple], n: int, amp_bytes: int) -> float:
    """Algorithmic HBM bytes of a gate list on an n-qubit density matrix
    (N_rho = 4^n entries): a unitary is the same gate on the 2n-qubit vector
    twice (rows, then conjugated on columns); mixDephasing reads and rewrites
    the half of the entries whose row and column bits differ (N_rho*B);
    mixDepolarising and mixDamping rewrite every entry (2*N_rho*B)."""
    Nr = float(1 << (2 * n))
    tot = 0.0
    for g in gates:
        if g[0] == "deph":
            tot += Nr * amp_bytes
        elif g[0] in ("depol", "damp"):
            tot += 2 * Nr * amp_bytes
        else:
            tot += 2 * gate_bytes([g], 2 * n, amp_bytes)
    return tot
