"""Oracle for the uniformly controlled rotations (QuantumLab's FR gates) and the
complex-vector amplitude encoding.

TEST INFRASTRUCTURE ONLY (tests/ may use it; the product never does). numpy
FP64 and gate lists for a QuEST-ABI binding.

The reference builds these as gate lists (QuantumLab, Circuit/Gates):
  FRy, FRz  FRGate.cu:181-330. n local qubits, target = local qubit n-1,
            2^(n-1) angles. The angles are split (SpliteAngles, FRGate.cu:154-174,
            restated by swaptest_oracle.split_angles), then each split angle is
            a rotation of the target followed by a CNOT from the control whose
            Gray-code bit changes next (GrayCodeDifferent, FRGate.h:16-58).
            Angle index bit b belongs to local qubit n-2-b.
  FRyz      FRGate.cu:400-460. The FRy ladder without its last CNOT, then the
            Rz ladder in reverse order with negated split angles.
  FRp       FRGate.cu:460-515. X on every control, then per control value i a
            multi-controlled phase on all n qubits (CreateCnP, CnUGate.cpp:287),
            flipping the controls whose bit of i changed. Angle index bit j
            belongs to local qubit j. One qubit: a phase shift (EBO_P).
  complex AE AmplitudeEncodeCuda.cu:20-30 (|v_i| and arg v_i), :61-109
            (_kernelAE_Rotations: level l pairs left = k 2^(l+1), right =
            left + 2^l; Y = 2 atan2(|right|, |left|), running magnitude as in the
            real encoder; Z = phase(left) - phase(right), running phase = their
            mean), :452-496 (ExchangeToYZGate: Ry then Rz(-Z) on the top qubit,
            then FRyz per level), :691-724 (AmplitudeEncodeOneVector). With
            bHasPhase false, the final global phase (the mean of all phases) is
            not applied; QuantumLab's QuEST simulators have no EBO_Phase either.

Gate tuples: ("ry"|"rz"|"p", t, theta), ("cx", c, t), ("x", t),
("mcp", [qubits], theta). `bits[j]` is the register qubit of local qubit j.
`uniform_dense` is the operator's definition (matrix k on the target where the
controls read k), which pins the gate lists.
"""
from __future__ import annotations

import numpy as np

from oracle.swaptest_oracle import (DBL_EPSILON, _apply_1q, _apply_cx, _ry, _rz, gray_code_different,
                                    run_gates_dense as _run_basic, split_angles)


def _p(theta):
    return np.diag([1.0, np.exp(1j * theta)])


def fr_gates(kind: str, angles, bits) -> list:
    """FRy / FRz (kind "ry" / "rz") on register qubits bits[0..n-1]."""
    n = len(bits)
    deg = 1 << (n - 1)
    if n == 1:
        return [(kind, bits[0], float(angles[0]))]
    theta = split_angles(list(angles[:deg]), deg)
    out = []
    for i in range(deg):
        out.append((kind, bits[n - 1], float(theta[i])))
        out.append(("cx", bits[n - 2 - gray_code_different(i, deg)], bits[n - 1]))
    return out


def fryz_gates(angles_y, angles_z, bits) -> list:
    """FRyz on register qubits bits[0..n-1]."""
    n = len(bits)
    deg = 1 << (n - 1)
    t = bits[n - 1]
    if n == 1:
        return [("ry", t, float(angles_y[0])), ("rz", t, -float(angles_z[0]))]
    ty = split_angles(list(angles_y[:deg]), deg)
    tz = split_angles(list(angles_z[:deg]), deg)
    out = []
    for i in range(deg):
        out.append(("ry", t, float(ty[i])))
        if i != deg - 1:
            out.append(("cx", bits[n - 2 - gray_code_different(i, deg)], t))
    for i in range(deg):
        j = deg - 1 - i
        if i != 0:
            out.append(("cx", bits[n - 2 - gray_code_different(j, deg)], t))
        out.append(("rz", t, -float(tz[j])))
    return out


def frp_gates(angles, bits) -> list:
    """FRp on register qubits bits[0..n-1]."""
    n = len(bits)
    deg = 1 << (n - 1)
    if n == 1:
        return [("p", bits[0], float(angles[0]))]
    out = [("x", bits[j]) for j in range(n - 1)]
    out.append(("mcp", list(bits), float(angles[0])))
    for i in range(1, deg):
        diff = i ^ (i - 1)
        out += [("x", bits[j]) for j in range(n - 1) if diff & (1 << j)]
        out.append(("mcp", list(bits), float(angles[i])))
    return out


def fr_controls(kind: str, bits) -> tuple:
    """The (controls, target) of the same operator in the library's convention
    (bit b of the angle index is controls[b])."""
    n = len(bits)
    if kind == "p":
        return list(bits[:n - 1]), bits[n - 1]
    return [bits[n - 2 - b] for b in range(n - 1)], bits[n - 1]


def rotation(kind: str, theta: float) -> np.ndarray:
    return {"ry": _ry, "rz": _rz, "p": _p}[kind](theta)


def uniform_dense(state, n, controls, target, mats) -> np.ndarray:
    """Matrix k on `target` for the amplitudes whose controls read k (bit b of k
    = qubit controls[b]): the definition of a uniformly controlled gate."""
    out = state.copy()
    idx = np.arange(state.size)
    lower = idx[((idx >> target) & 1) == 0]
    k = np.zeros(lower.size, dtype=np.int64)
    for b, c in enumerate(controls):
        k |= ((lower >> c) & 1) << b
    upper = lower | (1 << target)
    m = np.asarray(mats)[k]
    out[lower] = m[:, 0, 0] * state[lower] + m[:, 0, 1] * state[upper]
    out[upper] = m[:, 1, 0] * state[lower] + m[:, 1, 1] * state[upper]
    return out


def run_gates_dense(gates, n, state=None) -> np.ndarray:
    """The basic gate list of swaptest_oracle plus "p", "x" and "mcp"."""
    if state is None:
        state = np.zeros(1 << n, dtype=complex)
        state[0] = 1.0
    state = np.asarray(state, dtype=complex)
    idx = np.arange(state.size)
    for g in gates:
        if g[0] == "p":
            state = _apply_1q(state, n, g[1], _p(g[2]))
        elif g[0] == "x":
            state = state[idx ^ (1 << g[1])]
        elif g[0] == "mcp":
            mask = sum(1 << q for q in g[1])
            state = np.where((idx & mask) == mask, state * np.exp(1j * g[2]), state)
        else:
            state = _run_basic([g], n, state)
    return state


def run_gates_quest(qe, qureg, gates) -> None:
    """The gate list through a QuEST C API binding, one QuEST call per basic
    operation as QLGate::PerformBasicOperation makes them (QLGate.cpp:652-803)."""
    import ctypes as C
    L = qe.lib
    for g in gates:
        if g[0] == "ry":
            L.rotateY(qureg, g[1], g[2])
        elif g[0] == "rz":
            L.rotateZ(qureg, g[1], g[2])
        elif g[0] == "p":
            L.phaseShift(qureg, g[1], g[2])
        elif g[0] == "cx":
            L.controlledNot(qureg, g[1], g[2])
        elif g[0] == "x":
            L.pauliX(qureg, g[1])
        elif g[0] == "mcp":
            arr = (C.c_int * len(g[1]))(*g[1])
            L.multiControlledPhaseShift(qureg, arr, len(g[1]), g[2])
        else:
            raise ValueError(g[0])


# ---- complex amplitude encoding ----------------------------------------------

def ae_angles_complex(v, eps: float = DBL_EPSILON):
    """_kernelAE_Rotations restated for one vector: Y (2^q - 1 angles, level-major)
    and Z (2^q entries, the last the mean phase)."""
    v = np.asarray(v, dtype=complex)
    D = v.size
    q = D.bit_length() - 1
    mag = np.abs(v).astype(np.float64)
    ph = np.angle(v).astype(np.float64)
    Y = np.zeros(D)
    Z = np.zeros(D)
    for level in range(q):
        pairs = D >> (level + 1)
        base = D - (D >> level)
        for k in range(pairs):
            left = k << (level + 1)
            right = left + (1 << level)
            c, s = mag[left], mag[right]
            Y[base + k] = 2 * np.arctan2(s, c)
            if pairs != 1:
                m2 = c * c + s * s
                mag[left] = np.sqrt(m2) if m2 > eps else 0.0
            Z[base + k] = ph[left] - ph[right]
            ph[left] = 0.5 * (ph[left] + ph[right])
    Z[D - 1] = ph[0]
    return Y, Z


def ae_gates_complex(v, qubits) -> list:
    """ExchangeToYZGate (bHasPhase false) with FRyz expanded; `qubits[j]` is the
    register qubit of the encoder's qubit j."""
    D = len(v)
    q = D.bit_length() - 1
    Y, Z = ae_angles_complex(v)
    gates = []
    idx = D - 2
    for i in range(q):
        bits = [qubits[q - 1 - j] for j in range(i + 1)]
        if i == 0:
            gates += [("ry", bits[0], float(Y[idx])), ("rz", bits[0], -float(Z[idx]))]
            idx -= 2
            continue
        deg = 1 << i
        gates += fryz_gates(Y[idx:idx + deg], Z[idx:idx + deg], bits)
        idx -= 1 << (i + 1)
    return gates


def ae_complex_closed_form(v) -> np.ndarray:
    """What the encoder prepares when no running magnitude is cut by the epsilon:
    v / |v| times exp(-i mean(arg v))."""
    v = np.asarray(v, dtype=complex)
    return v / np.linalg.norm(v) * np.exp(-1j * np.mean(np.angle(v)))
