"""Seeded Pauli-sum Hamiltonians and product states for Trotter evolution.

Instance family (the Hamiltonian-simulation workload): a 1D chain, transverse-
field Ising (J Z_i Z_{i+1} bonds, h X_i field) or Heisenberg (J (X X + Y Y +
Z Z) bonds, h X_i field), with one J ~ U(0.5, 1.5) and one h ~ U(0, 2) per
Hamiltonian, followed by `random_strings` random 2-4-local Pauli strings
(k ~ U{2,3,4} distinct qubits anywhere on the chain, each factor uniform in
{X, Y, Z}, coefficient N(0, 0.3^2)). Term order is bonds, field, strings (the
order applyTrotterCircuit applies them in). The initial state is a product of
single-qubit states uniform on the Bloch sphere, prepared from |0..0> by
rotateY(theta_q) then rotateZ(phi_q) on every qubit. Evolution time t ~ U(0.5, 4).

At 24 qubits the family has 32 random strings; smaller training registers
keep the same strings per qubit (round(32 n / 24)).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

PAULI_I, PAULI_X, PAULI_Y, PAULI_Z = 0, 1, 2, 3


@dataclass
class TrotterInstance:
    n: int
    coeffs: np.ndarray   # (T,) float64
    codes: np.ndarray    # (T, n) int32, codes[t, q] = Pauli on qubit q
    time: float
    theta: np.ndarray    # (n,) rotateY angles of the product state
    phi: np.ndarray      # (n,) rotateZ angles
    model: str
    J: float
    h: float

    @property
    def terms(self) -> int:
        return len(self.coeffs)


def default_strings(n: int) -> int:
    return max(2, int(round(32 * n / 24)))


def trotter_instance(n: int, seed: int, index: int, random_strings: int | None = None) -> TrotterInstance:
    rng = np.random.default_rng([seed, index, 0x7107])
    k_rand = default_strings(n) if random_strings is None else random_strings
    model = "tfim" if rng.random() < 0.5 else "heisenberg"
    J = float(rng.uniform(0.5, 1.5))
    h = float(rng.uniform(0.0, 2.0))
    rows, coeffs = [], []

    def term(paulis: dict, c: float):
        row = np.zeros(n, dtype=np.int32)
        for q, p in paulis.items():
            row[q] = p
        rows.append(row)
        coeffs.append(c)

    bond_paulis = [PAULI_Z] if model == "tfim" else [PAULI_X, PAULI_Y, PAULI_Z]
    for i in range(n - 1):
        for p in bond_paulis:
            term({i: p, i + 1: p}, J)
    for i in range(n):
        term({i: PAULI_X}, h)
    for _ in range(k_rand):
        k = int(rng.integers(2, 5))
        qs = rng.choice(n, size=min(k, n), replace=False)
        ps = rng.integers(1, 4, size=len(qs))
        term({int(q): int(p) for q, p in zip(qs, ps)}, float(rng.normal(0.0, 0.3)))
    t = float(rng.uniform(0.5, 4.0))
    theta = np.arccos(1.0 - 2.0 * rng.uniform(0.0, 1.0, size=n))
    phi = rng.uniform(0.0, 2 * np.pi, size=n)
    return TrotterInstance(n, np.asarray(coeffs, dtype=np.float64), np.stack(rows), t, theta, phi, model, J, h)


def product_state(inst: TrotterInstance) -> np.ndarray:
    """Amplitudes of the prepared product state (qubit q = bit q)."""
    amp = np.ones(1, dtype=complex)
    for q in range(inst.n):
        f0 = np.cos(inst.theta[q] / 2) * np.exp(-0.5j * inst.phi[q])
        f1 = np.sin(inst.theta[q] / 2) * np.exp(0.5j * inst.phi[q])
        amp = np.concatenate([amp * f0, amp * f1])   # new qubit is the next-higher bit
    return amp


def prepare(qe, qureg, inst: TrotterInstance):
    """initZeroState, then rotateY(theta_q), rotateZ(phi_q) on every qubit."""
    qe.lib.initZeroState(qureg)
    for q in range(inst.n):
        qe.lib.rotateY(qureg, q, float(inst.theta[q]))
        qe.lib.rotateZ(qureg, q, float(inst.phi[q]))


def stages(order: int) -> int:
    """Exponentiated-Hamiltonian passes per repetition of an order's formula."""
    return 1 if order == 1 else 2 * 5 ** (order // 2 - 1)


class HamilBatch:
    """ctypes PauliHamil records (and the arrays they point into) for a list of
    instances, plus the packed times and product-state angles."""

    def __init__(self, insts):
        import ctypes as C

        from .quest import PauliHamil, qreal
        self.insts = list(insts)
        self._keep = []
        self.hamils = (PauliHamil * len(self.insts))()
        for j, inst in enumerate(self.insts):
            codes = np.ascontiguousarray(inst.codes, dtype=np.int32).ravel()
            coeffs = np.ascontiguousarray(inst.coeffs, dtype=np.float64)
            self._keep += [codes, coeffs]
            h = self.hamils[j]
            h.pauliCodes = codes.ctypes.data_as(C.POINTER(C.c_int))
            h.termCoeffs = coeffs.ctypes.data_as(C.POINTER(qreal))
            h.numSumTerms = inst.terms
            h.numQubits = inst.n
        self.times = np.array([i.time for i in self.insts], dtype=np.float64)
        self.theta = np.concatenate([i.theta for i in self.insts]).astype(np.float64)
        self.phi = np.concatenate([i.phi for i in self.insts]).astype(np.float64)

    def __len__(self):
        return len(self.insts)

    def __getitem__(self, j):
        return self.hamils[j]
