"""Dense-matrix model of QuEST operations for small registers.

The method of QuEST's own test-suite (QuEST340/tests/utilities.cpp:304-370,
760-886): build the full 2^n x 2^n operator and multiply it onto the vector, or
form U rho U^dagger / Kraus sums for density matrices. Independent of any
QuEST code path, it pins the oracle by mathematics."""
from __future__ import annotations

import numpy as np

X = np.array([[0, 1], [1, 0]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
Z = np.array([[1, 0], [0, -1]], dtype=complex)
I2 = np.eye(2, dtype=complex)
PAULI = [I2, X, Y, Z]


def full_op(n: int, targets, U, controls=(), ctrl_state=None) -> np.ndarray:
    """Operator on n qubits; targets[0] is the least significant bit of U's index."""
    U = np.asarray(U, dtype=complex)
    k = len(targets)
    dim = 1 << n
    M = np.zeros((dim, dim), dtype=complex)
    cs = [1] * len(controls) if ctrl_state is None else list(ctrl_state)
    for i in range(dim):
        if any(((i >> c) & 1) != s for c, s in zip(controls, cs)):
            M[i, i] = 1
            continue
        col = 0
        for j, t in enumerate(targets):
            col |= ((i >> t) & 1) << j
        base = i
        for t in targets:
            base &= ~(1 << t)
        for r in range(1 << k):
            out = base
            for j, t in enumerate(targets):
                if (r >> j) & 1:
                    out |= 1 << t
            M[out, i] += U[r, col]
    return M


def rz(theta):
    return np.diag([np.exp(-0.5j * theta), np.exp(0.5j * theta)])


def compact(a, b):
    return np.array([[a, -np.conj(b)], [b, np.conj(a)]], dtype=complex)


def pauli_string(n: int, targets, paulis) -> np.ndarray:
    M = np.eye(1 << n, dtype=complex)
    for t, p in zip(targets, paulis):
        M = full_op(n, [t], PAULI[p]) @ M
    return M


def expm_herm(H: np.ndarray, t: float) -> np.ndarray:
    """exp(-i t H) for Hermitian H by eigen-decomposition."""
    w, V = np.linalg.eigh(H)
    return (V * np.exp(-1j * t * w)) @ V.conj().T


def hamil_matrix(n: int, coeffs, codes) -> np.ndarray:
    H = np.zeros((1 << n, 1 << n), dtype=complex)
    for c, row in zip(coeffs, codes):
        H += c * pauli_string(n, list(range(n)), list(row))
    return H


def to_rho(flat: np.ndarray) -> np.ndarray:
    dim = int(round(np.sqrt(flat.size)))
    return flat.reshape(dim, dim).T


def from_rho(rho: np.ndarray) -> np.ndarray:
    return rho.T.reshape(-1)


def kraus(rho: np.ndarray, ops) -> np.ndarray:
    return sum(K @ rho @ K.conj().T for K in ops)
