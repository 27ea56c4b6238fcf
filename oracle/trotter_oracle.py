"""Oracle for the learned Trotter planner: features, decision rule and the MLP.

TEST INFRASTRUCTURE ONLY (tests/ and bench.py's cpu_baseline leg may use it;
the product never does). numpy.

Trotter-Suzuki itself is QuEST's (QuEST_common.c:762-844): the oracle for the
batched labelling kernel is applyTrotterCircuit of QuEST's CPU backend
(oracle/_ref) and of the C restatement (oracle/build), driven through the C
API. Here:
  * features of a Pauli sum (coefficient norms, first- and second-order
    commutator bounds), restated from their definitions with Pauli masks and
    pinned by dense-matrix brute force (`dense_commutator_sums`);
  * the 16 -> 64 -> 64 -> 2 ReLU MLP with bfloat16-rounded weights and FP32
    arithmetic, forward and backward (MSE over masked outputs), pinned by
    finite differences.
"""
from __future__ import annotations

import numpy as np

NUM_FEATURES = 16
LAYERS = ((16, 64), (64, 64), (64, 2))
NUM_PARAMS = sum(a * b + b for a, b in LAYERS)  # 5378
LOG_FLOOR = -60.0


# ---- Pauli strings as masks -----------------------------------------------------

def masks(codes: np.ndarray):
    """codes (T, n) -> x mask, z mask (Python ints), #Y per term."""
    out = []
    for row in codes:
        x = z = ny = 0
        for q, p in enumerate(row):
            if p == 1:
                x |= 1 << q
            elif p == 2:
                x |= 1 << q
                z |= 1 << q
                ny += 1
            elif p == 3:
                z |= 1 << q
        out.append((x, z, ny))
    return out


def anticommute(a, b) -> bool:
    return (bin(a[0] & b[1]).count("1") + bin(a[1] & b[0]).count("1")) & 1 == 1


def commutator_sums(codes, coeffs):
    """C1 = sum_{j<k} ||[c_j P_j, c_k P_k]|| and
    C2 = sum_{j<k} sum_l ||[c_l P_l, [c_j P_j, c_k P_k]]|| over non-identity terms."""
    m = masks(codes)
    keep = [i for i, (x, z, _) in enumerate(m) if x or z]
    C1 = C2 = 0.0
    pairs = anti = 0
    for a in range(len(keep)):
        for b in range(a + 1, len(keep)):
            j, k = keep[a], keep[b]
            pairs += 1
            if not anticommute(m[j], m[k]):
                continue
            anti += 1
            cjk = abs(coeffs[j] * coeffs[k])
            C1 += 2 * cjk
            prod = (m[j][0] ^ m[k][0], m[j][1] ^ m[k][1])
            for l in keep:
                if anticommute(m[l], prod):
                    C2 += 4 * cjk * abs(coeffs[l])
    return C1, C2, anti, pairs


def _pauli_dense(row):
    I = np.eye(2)
    X = np.array([[0, 1], [1, 0]], dtype=complex)
    Y = np.array([[0, -1j], [1j, 0]])
    Z = np.diag([1.0, -1.0]).astype(complex)
    mats = [I, X, Y, Z]
    out = np.ones((1, 1), dtype=complex)
    for q in range(len(row)):          # qubit q is bit q: kron(new, out) puts it higher
        out = np.kron(mats[row[q]], out)
    return out


def dense_commutator_sums(codes, coeffs):
    """C1, C2 from spectral norms of dense commutators (tiny n)."""
    keep = [i for i, row in enumerate(codes) if np.any(row)]
    H = [coeffs[i] * _pauli_dense(codes[i]) for i in keep]
    C1 = C2 = 0.0
    for a in range(len(H)):
        for b in range(a + 1, len(H)):
            c = H[a] @ H[b] - H[b] @ H[a]
            C1 += np.linalg.norm(c, 2)
            for l in range(len(H)):
                C2 += np.linalg.norm(H[l] @ c - c @ H[l], 2)
    return C1, C2


def _log2(v):
    return max(LOG_FLOOR, float(np.log2(v))) if v > 0 else LOG_FLOOR


def features(codes, coeffs, time: float) -> np.ndarray:
    """The planner's 16 per-qubit inputs, in the library's order:
    t, log2 t, terms/n, lambda/n, log2(lambda/n), max|c|, ||c||_2/sqrt(n),
    log2(C1/n), log2(C2/n), log2(lambda t/n), log2(C1 t^2/n), log2(C2 t^3/n),
    anticommuting partners per term, mean weight, max weight, share of lambda
    on weight >= 3 terms (identity strings dropped, logs floored at -60)."""
    codes = np.asarray(codes)
    coeffs = np.asarray(coeffs, dtype=np.float64)
    n = codes.shape[1]
    live = np.any(codes != 0, axis=1)
    c = coeffs[live]
    w = np.count_nonzero(codes[live], axis=1)
    T = int(live.sum())
    lam = float(np.sum(np.abs(c)))
    C1, C2, anti, _ = commutator_sums(codes, coeffs)
    t = time
    f = [t, _log2(t), T / n, lam / n, _log2(lam / n), float(np.max(np.abs(c))) if T else 0.0,
         float(np.sqrt(np.sum(c * c) / n)), _log2(C1 / n), _log2(C2 / n), _log2(lam * t / n),
         _log2(C1 * t * t / n), _log2(C2 * t ** 3 / n), 2.0 * anti / T if T else 0.0,
         float(np.mean(w)) if T else 0.0, float(np.max(w)) if T else 0.0,
         float(np.sum(np.abs(c[w >= 3])) / lam) if lam > 0 else 0.0]
    return np.asarray(f, dtype=np.float64)


def size_curve(kx, ky, n: int, o: int) -> float:
    """Piecewise-linear interpolation in log2 n through the knots, the end
    segments extended (restated from the planner's definition)."""
    x = np.log2(n)
    k = 0
    while k < len(kx) - 2 and x > kx[k + 1]:
        k += 1
    return ky[k][o] + (ky[k + 1][o] - ky[k][o]) * (x - kx[k]) / (kx[k + 1] - kx[k])


def choose(y, n: int, margin: float, kx=(0.0, 1.0), ky=((0.0, 0.0), (0.25, 0.125))):
    """The planner's decision from its two outputs (orders 2 and 4): repetitions
    ceil(2^(y_o + curve_o(n) + margin)) and the order with fewer Pauli
    exponentials (2 r_2 versus 10 r_4; ties go to order 2)."""
    best = None
    for c, (o, stages) in enumerate(((2, 2), (4, 10))):
        l2 = float(y[c]) + size_curve(kx, ky, n, c) + margin
        r = max(1, int(np.ceil(2.0 ** min(30.0, l2) - 1e-9)))
        if best is None or stages * r < best[2]:
            best = (o, r, stages * r)
    return best[0], best[1]


# ---- the MLP --------------------------------------------------------------------------

def bf16(a) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).reshape(a.shape)


def unpack(params: np.ndarray):
    """Flat parameter vector -> [(W (in, out), b (out,)), ...] (row-major W)."""
    out, o = [], 0
    for a, b in LAYERS:
        W = params[o:o + a * b].reshape(a, b)
        o += a * b
        bias = params[o:o + b]
        o += b
        out.append((W, bias))
    return out


def _layers(params, rounded: bool, dtype):
    p = np.asarray(params, np.float32)
    if rounded:
        return [(bf16(W).astype(dtype), bf16(b).astype(dtype)) for W, b in unpack(p)]
    return [(W.astype(dtype), b.astype(dtype)) for W, b in unpack(np.asarray(params, dtype))]


def forward(params, X, rounded: bool = True, dtype=np.float32):
    """Y = W3^T relu(W2^T relu(W1^T x + b1) + b2) + b3 with bf16-rounded weights
    and biases (rounded=True), in `dtype` arithmetic; returns Y and the
    activations."""
    (W1, b1), (W2, b2), (W3, b3) = _layers(params, rounded, dtype)
    X = np.asarray(X, dtype)
    Z1 = X @ W1 + b1
    H1 = np.maximum(Z1, 0)
    Z2 = H1 @ W2 + b2
    H2 = np.maximum(Z2, 0)
    Y = H2 @ W3 + b3
    return Y, (X, Z1, H1, Z2, H2)


def loss_and_grad(params, X, T, M, rounded: bool = True, dtype=np.float32):
    """Masked MSE: L = sum(M (Y - T)^2) / max(1, sum M); gradient with respect
    to the (bf16-rounded when rounded=True) parameters, flat like `params`."""
    (W1, b1), (W2, b2), (W3, b3) = _layers(params, rounded, dtype)
    Y, (X, Z1, H1, Z2, H2) = forward(params, X, rounded, dtype)
    M = np.asarray(M, dtype)
    denom = max(1.0, float(M.sum()))
    D = Y - np.asarray(T, dtype)
    L = float(np.sum(M * D * D) / denom)
    dY = (2.0 / denom) * M * D
    gW3 = H2.T @ dY
    gb3 = dY.sum(0)
    dZ2 = (dY @ W3.T) * (Z2 > 0)
    gW2 = H1.T @ dZ2
    gb2 = dZ2.sum(0)
    dZ1 = (dZ2 @ W2.T) * (Z1 > 0)
    gW1 = X.T @ dZ1
    gb1 = dZ1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3]).astype(dtype)
    return L, g


def init_params(seed: int) -> np.ndarray:
    """He-uniform weights, zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for a, b in LAYERS:
        lim = np.sqrt(6.0 / a)
        parts += [rng.uniform(-lim, lim, size=a * b), np.zeros(b)]
    return np.concatenate(parts).astype(np.float32)


def adam_step(params, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v
