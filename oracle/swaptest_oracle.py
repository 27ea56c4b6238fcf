"""Oracle for amplitude encoding, the swap test and the k-means iteration.

TEST INFRASTRUCTURE ONLY (tests/, smoke() and bench.py's cpu_baseline may use
it; the product never does). numpy FP64.

The reference builds these circuits gate by gate (QuantumLab):
  angles  AmplitudeEncodeCuda.cu:112-150 (_kernelAE_RotationsReal): level l
          pairs (left = k 2^(l+1), right = left + 2^l) of running magnitudes,
          theta = 2 atan2(right, left), stored at index 2^q - 2^(q-l) + k;
  circuit AmplitudeEncodeCuda.cu:498-533 (ExchangeToYGate): Ry on the top
          qubit, then a uniformly controlled Ry (FRy, FRGate.cu:181-252,
          angle split FRGate.cu:154-174, Gray codes FRGate.h:16-58) per level;
  swap    SwapTest.cpp:70-123 (CreateSwapTestReal): H(0); AE(v1) on 1..q;
          AE(v2) on q+1..2q; CSWAP(0, i+1, i+1+q) as CX, CCX, CX
          (SimpleGates.cpp:161-176); H(0).
`ae_circuit_dense` and `swap_test_circuit_dense` restate those gate lists on a
dense vector (small q only) and pin the closed forms the fast paths use:
the encoder prepares v/|v| with qubit l <-> bit l, and
P(ancilla = 0) = (1 + <a|b>^2) / 2.
"""
from __future__ import annotations

import numpy as np


# ---- gate-level restatement (small registers) --------------------------------

def _apply_1q(state, n, q, U):
    s = state.reshape([2] * n)  # axis 0 is qubit n-1
    ax = n - 1 - q
    s = np.moveaxis(s, ax, 0)
    s = np.tensordot(U, s, axes=([1], [0]))
    return np.moveaxis(s, 0, ax).reshape(-1)


def _apply_cx(state, n, c, t):
    idx = np.arange(state.size)
    sel = ((idx >> c) & 1) == 1
    out = state.copy()
    out[sel] = state[idx[sel] ^ (1 << t)]
    return out


def _ry(theta):
    c, s = np.cos(theta / 2), np.sin(theta / 2)
    return np.array([[c, -s], [s, c]])


def gray(n):
    return n ^ (n >> 1)


def gray_code_different(n, maxn):
    a = gray(n)
    b = gray(0) if n == maxn - 1 else gray(n + 1)
    return int(a ^ b).bit_length() - 1


def split_angles(angles, length):
    """FRGate.cu:154-174: theta'_i = (1/len) sum_j (-1)^{|j & gray(i)|} angles_j."""
    out = []
    for i in range(length):
        th = 0.0
        for j in range(length):
            th += -angles[j] if bin(j & gray(i)).count("1") & 1 else angles[j]
        out.append(th / length)
    return out


DBL_EPSILON = 2.220446049250313e-16
FLT_EPSILON = 1.1920928955078125e-07


def ae_angles(v, eps: float = DBL_EPSILON) -> np.ndarray:
    """_kernelAE_RotationsReal restated: the 2^q - 1 angles, level-major.
    Running magnitudes with m^2 <= eps are zeroed (the reference's
    _QL_FLT_EPSILON: DBL_EPSILON in a double build, FLT_EPSILON in a float one,
    QLFloat.h:120,139)."""
    v = np.asarray(v, dtype=np.float64)
    D = v.size
    q = D.bit_length() - 1
    mag = v.copy()
    Y = np.zeros(D)
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
    return Y


def ae_gates(v, qubits) -> list:
    """ExchangeToYGate (AmplitudeEncodeCuda.cu:498-533) with FRy expanded
    (FRGate.cu:181-252) as a list of QuEST calls: ("ry", t, theta) and
    ("cx", c, t). `qubits[j]` is the register qubit of the encoder's qubit j."""
    D = len(v)
    q = D.bit_length() - 1
    Y = ae_angles(v)
    gates = []
    idx = D - 2
    for i in range(q):
        bits = [qubits[q - 1 - j] for j in range(i + 1)]   # local qubit j -> register qubit bits[j]
        if i == 0:
            gates.append(("ry", bits[0], float(Y[idx])))
            idx -= 2
            continue
        nq = i + 1
        deg = 1 << (nq - 1)
        theta = split_angles(Y[idx:idx + deg], deg)
        target = bits[nq - 1]
        for k in range(deg):
            gates.append(("ry", target, float(theta[k])))
            ctrl_local = nq - 2 - gray_code_different(k, deg)
            gates.append(("cx", bits[ctrl_local], target))
        idx -= 1 << (i + 1)
    return gates


def toffoli_gates(c0, c1, t) -> list:
    """CreateToffoliGate (CnUGate.h:50-78): H, CX and RZ(-/+ pi/4) only."""
    T, Td = ("rz", -np.pi / 4), ("rz", np.pi / 4)
    return [("h", t), ("cx", c1, t), (Td[0], t, Td[1]), ("cx", c0, t), (T[0], t, T[1]), ("cx", c1, t),
            (Td[0], t, Td[1]), ("cx", c0, t), (T[0], t, T[1]), ("h", t), (T[0], c1, T[1]), ("cx", c0, c1),
            (T[0], c0, T[1]), (Td[0], c1, Td[1]), ("cx", c0, c1)]


def swap_test_gates(v1, v2) -> list:
    """CreateSwapTestReal (SwapTest.cpp:70-123) on 2q+1 qubits as QuEST calls:
    H(0); AE(v1) on 1..q; AE(v2) on q+1..2q; per i CSWAP(0, i+1, i+1+q) =
    CX(i+1+q -> i+1), Toffoli(0, i+1 -> i+1+q), CX(i+1+q -> i+1)
    (CreateControlledSwap, SimpleGates.cpp:161-176); H(0)."""
    q = len(v1).bit_length() - 1
    gates = [("h", 0)]
    gates += ae_gates(v1, [1 + j for j in range(q)])
    gates += ae_gates(v2, [1 + q + j for j in range(q)])
    for i in range(q):
        t1, t2 = i + 1, i + 1 + q
        gates += [("cx", t2, t1)] + toffoli_gates(0, t1, t2) + [("cx", t2, t1)]
    gates.append(("h", 0))
    return gates


def _rz(theta):
    return np.diag([np.exp(-0.5j * theta), np.exp(0.5j * theta)])


_H = np.array([[1, 1], [1, -1]]) / np.sqrt(2)


def run_gates_dense(gates, n, state=None) -> np.ndarray:
    """The gate list on `state` (default |0...0>) of n qubits (dense complex
    vector, small n)."""
    if state is None:
        state = np.zeros(1 << n, dtype=complex)
        state[0] = 1.0
    for g in gates:
        if g[0] == "ry":
            state = _apply_1q(state, n, g[1], _ry(g[2]))
        elif g[0] == "rz":
            state = _apply_1q(state, n, g[1], _rz(g[2]))
        elif g[0] == "h":
            state = _apply_1q(state, n, g[1], _H)
        elif g[0] == "cx":
            state = _apply_cx(state, n, g[1], g[2])
        else:
            raise ValueError(g[0])
    return state


def run_gates_quest(qe, qureg, gates) -> None:
    """The gate list through a QuEST C API binding (QLGate::PerformBasicOperation,
    QLGate.cpp:652-703, maps each basic operation to one QuEST call)."""
    L = qe.lib
    for g in gates:
        if g[0] == "ry":
            L.rotateY(qureg, g[1], g[2])
        elif g[0] == "rz":
            L.rotateZ(qureg, g[1], g[2])
        elif g[0] == "h":
            L.hadamard(qureg, g[1])
        elif g[0] == "cx":
            L.controlledNot(qureg, g[1], g[2])
        else:
            raise ValueError(g[0])


def ae_circuit_dense(v) -> np.ndarray:
    """ExchangeToYGate + FRy applied gate by gate to |0...0> of q qubits."""
    q = len(v).bit_length() - 1
    return run_gates_dense(ae_gates(v, list(range(q))), q).real


def ae_state_tree(v) -> np.ndarray:
    """Amplitude of x = product over levels of cos / sin(theta/2) along x's path
    (the state the Ry tree prepares from |0...0>)."""
    D = len(v)
    q = D.bit_length() - 1
    Y = ae_angles(v)
    amp = np.ones(D)
    x = np.arange(D)
    for level in range(q):
        idx = D - (D >> level) + (x >> (level + 1))
        bit = (x >> level) & 1
        amp *= np.where(bit == 1, np.sin(Y[idx] / 2), np.cos(Y[idx] / 2))
    return amp


def swap_test_circuit_dense(v1, v2) -> float:
    """CreateSwapTestReal's gate list on 2q+1 qubits, P(ancilla = 0)."""
    q = len(v1).bit_length() - 1
    n = 2 * q + 1
    state = run_gates_dense(swap_test_gates(v1, v2), n)
    idx = np.arange(state.size)
    return float(np.sum(np.abs(state[(idx & 1) == 0]) ** 2))


def ccx_dense(n, c0, c1, t) -> np.ndarray:
    """Exact Toffoli as a permutation of the basis (for pinning the decomposition)."""
    idx = np.arange(1 << n)
    sel = (((idx >> c0) & 1) == 1) & (((idx >> c1) & 1) == 1)
    src = idx.copy()
    src[sel] = idx[sel] ^ (1 << t)
    return src


# ---- k-means in FP64 (closed forms) ----------------------------------------------

def normalised_points(data):
    pts = []
    for p in range(len(data.qubits)):
        v = data.point(p)
        pts.append(v / np.linalg.norm(v))
    return pts


def p0_matrix(pts, cents):
    """P(ancilla = 0) = (1 + <a|b>^2)/2 with b = centroid truncated to the
    point's size and normalised."""
    C = cents.shape[0]
    out = np.empty((len(pts), C))
    for p, a in enumerate(pts):
        D = a.size
        b = cents[:, :D]
        nb = np.linalg.norm(b, axis=1)
        ov = (b @ a) / np.where(nb > 0, nb, 1.0)
        out[p] = 0.5 * (1.0 + ov ** 2)
    return out


def kmeans_step(pts, cents):
    p0 = p0_matrix(pts, cents)
    assign = np.argmax(p0, axis=1)
    new = cents.copy()
    dim = cents.shape[1]
    for c in range(cents.shape[0]):
        members = [pts[p] for p in np.nonzero(assign == c)[0]]
        if not members:
            continue
        s = np.zeros(dim)
        n = np.zeros(dim)
        for v in members:
            s[:v.size] += v
            n[:v.size] += 1
        ok = n > 0
        new[c, ok] = s[ok] / n[ok]
    return p0, assign, new
