"""Python side of the learned Trotter planner (qmxPlanner* / qmxTrotter* in
libquest_mi355x): batched small-register Trotter runs for labels, features,
and the network (training, prediction, decision) — all computed by the
library; this module only moves arrays."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .hamiltonians import HamilBatch, TrotterInstance, default_strings, stages, trotter_instance

NUM_FEATURES = 16
NUM_OUTPUTS = 2
ORDERS = (2, 4)


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None:
        a = a.reshape(shape)
    return a, a.ctypes.data_as(C.POINTER(C.c_float))


def _f64p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def features(lib, inst: TrotterInstance) -> np.ndarray:
    """qmxTrotterFeatures (host code; needs no device)."""
    hb = HamilBatch([inst])
    out = np.zeros(NUM_FEATURES, dtype=np.float32)
    lib.qmxTrotterFeatures.argtypes = [type(hb[0]), C.c_double, C.POINTER(C.c_float)]
    lib.qmxTrotterFeatures.restype = C.c_int
    if not lib.qmxTrotterFeatures(hb[0], inst.time, out.ctypes.data_as(C.POINTER(C.c_float))):
        raise ValueError("qmxTrotterFeatures rejected the Hamiltonian")
    return out


def feature_matrix(lib, insts) -> np.ndarray:
    return np.stack([features(lib, i) for i in insts])


def batch_evolve(qe, insts, orders, reps):
    """Final FP64 states of applyTrotterCircuit on each instance's product state."""
    hb = HamilBatch(insts)
    o = np.ascontiguousarray(orders, dtype=np.int32)
    r = np.ascontiguousarray(reps, dtype=np.int32)
    tot = sum(1 << i.n for i in insts)
    re, im = np.zeros(tot), np.zeros(tot)
    ok = qe.lib.qmxTrotterBatchEvolve(len(hb), hb.hamils, _f64p(hb.times), _f64p(hb.theta), _f64p(hb.phi),
                                      o.ctypes.data_as(C.POINTER(C.c_int)), r.ctypes.data_as(C.POINTER(C.c_int)),
                                      _f64p(re), _f64p(im))
    if not ok:
        raise ValueError("qmxTrotterBatchEvolve rejected the batch")
    out, off = [], 0
    for i in insts:
        d = 1 << i.n
        out.append(re[off:off + d] + 1j * im[off:off + d])
        off += d
    return out


def batch_infidelity(qe, insts, orders, reps) -> np.ndarray:
    """1 - |<psi_r|psi_4r>|^2 per instance (FP64, batched in LDS)."""
    hb = HamilBatch(insts)
    o = np.ascontiguousarray(orders, dtype=np.int32)
    r = np.ascontiguousarray(reps, dtype=np.int32)
    out = np.zeros(len(insts))
    ok = qe.lib.qmxTrotterBatchInfidelity(len(hb), hb.hamils, _f64p(hb.times), _f64p(hb.theta), _f64p(hb.phi),
                                          o.ctypes.data_as(C.POINTER(C.c_int)),
                                          r.ctypes.data_as(C.POINTER(C.c_int)), _f64p(out))
    if not ok:
        raise ValueError("qmxTrotterBatchInfidelity rejected the batch")
    return out


def cheapest_reps(qe, insts, order: int, eps: float, max_log2: int = 12):
    """The smallest r (integer) with 1 - |<psi_r|psi_4r>|^2 <= eps for each
    instance: the ladder r = 1, 2, 4, ... brackets it, bisection (batched)
    finds it; None where the ladder never passes. (The first passing r of the
    bisection, should the criterion not be monotonic inside the bracket.)"""
    inf, pas = batch_ladder(qe, insts, [order] * len(insts), max_log2, eps)
    lo = np.zeros(len(insts), dtype=np.int64)
    hi = np.zeros(len(insts), dtype=np.int64)
    for j, k in enumerate(pas):
        if k >= 0:
            hi[j] = 1 << int(k)
            lo[j] = (1 << int(k)) // 2 if k > 0 else 0
    while True:
        act = np.nonzero((pas >= 0) & (hi - lo > 1))[0]
        if act.size == 0:
            break
        mid = (lo[act] + hi[act]) // 2
        v = batch_infidelity(qe, [insts[j] for j in act], [order] * act.size, mid)
        ok = v <= eps
        hi[act[ok]] = mid[ok]
        lo[act[~ok]] = mid[~ok]
    return [int(hi[j]) if pas[j] >= 0 else None for j in range(len(insts))]


def batch_ladder(qe, insts, orders, max_log2: int, eps: float):
    """Infidelity ladders 1 - |<psi_r|psi_4r>|^2 for r = 1, 2, 4, ... (NaN past
    the first rung at or below eps) and the index of that rung (-1 if none)."""
    hb = HamilBatch(insts)
    o = np.ascontiguousarray(orders, dtype=np.int32)
    inf = np.zeros((len(insts), max_log2 + 1))
    pas = np.zeros(len(insts), dtype=np.int32)
    ok = qe.lib.qmxTrotterBatchLadder(len(hb), hb.hamils, _f64p(hb.times), _f64p(hb.theta), _f64p(hb.phi),
                                      o.ctypes.data_as(C.POINTER(C.c_int)), max_log2, eps, _f64p(inf),
                                      pas.ctypes.data_as(C.POINTER(C.c_int)))
    if not ok:
        raise ValueError("qmxTrotterBatchLadder rejected the batch")
    return inf, pas


class Planner:
    """A qmxPlanner (device-resident network)."""

    def __init__(self, qe, seed: int = 0, path: str | None = None):
        self.qe = qe
        self.lib = qe.lib
        if path is not None:
            self.h = self.lib.qmxPlannerLoad(path.encode())
            if not self.h:
                raise ValueError(f"cannot load planner from {path}")
        else:
            self.h = self.lib.qmxPlannerCreate(seed)
        self.P = self.lib.qmxPlannerNumParams()

    def params(self) -> np.ndarray:
        out = np.zeros(self.P, dtype=np.float32)
        self.lib.qmxPlannerGetParams(self.h, out.ctypes.data_as(C.POINTER(C.c_float)))
        return out

    def set_params(self, p):
        a, ptr = _f32(p)
        self.lib.qmxPlannerSetParams(self.h, ptr)

    def set_normalisation(self, mean, scale):
        m, mp = _f32(mean)
        s, sp = _f32(scale)
        self.lib.qmxPlannerSetNormalisation(self.h, mp, sp)

    def normalisation(self):
        m = np.zeros(NUM_FEATURES, dtype=np.float32)
        s = np.zeros(NUM_FEATURES, dtype=np.float32)
        self.lib.qmxPlannerGetNormalisation(self.h, m.ctypes.data_as(C.POINTER(C.c_float)),
                                            s.ctypes.data_as(C.POINTER(C.c_float)))
        return m, s

    def set_margin(self, m: float):
        self.lib.qmxPlannerSetMargin(self.h, float(m))

    def margin(self) -> float:
        return float(self.lib.qmxPlannerGetMargin(self.h))

    def set_size_curve(self, log2n, log2reps):
        x, xp = _f32(log2n)
        y, yp = _f32(log2reps, (-1, 2))
        if not self.lib.qmxPlannerSetSizeCurve(self.h, len(x), xp, yp):
            raise ValueError("invalid size curve")

    def size_curve(self):
        k = self.lib.qmxPlannerGetSizeCurve(self.h, None, None)
        x = np.zeros(k, dtype=np.float32)
        y = np.zeros((k, 2), dtype=np.float32)
        self.lib.qmxPlannerGetSizeCurve(self.h, x.ctypes.data_as(C.POINTER(C.c_float)),
                                        y.ctypes.data_as(C.POINTER(C.c_float)))
        return x, y

    def predict(self, X) -> np.ndarray:
        X, xp = _f32(X)
        Y = np.zeros((X.shape[0], NUM_OUTPUTS), dtype=np.float32)
        self.lib.qmxPlannerPredict(self.h, xp, X.shape[0], Y.ctypes.data_as(C.POINTER(C.c_float)))
        return Y

    def gradient(self, X, T, M):
        X, xp = _f32(X)
        T, tp = _f32(T)
        M, mp = _f32(M)
        g = np.zeros(self.P, dtype=np.float32)
        loss = self.lib.qmxPlannerGradient(self.h, xp, tp, mp, X.shape[0], g.ctypes.data_as(C.POINTER(C.c_float)))
        return loss, g

    def train(self, X, T, M, epochs: int, batch: int = 256, lr: float = 1e-3, seed: int = 0):
        X, xp = _f32(X)
        T, tp = _f32(T)
        M, mp = _f32(M)
        losses = np.zeros(epochs, dtype=np.float32)
        self.lib.qmxPlannerTrain(self.h, xp, tp, mp, X.shape[0], epochs, batch, lr, seed,
                                 losses.ctypes.data_as(C.POINTER(C.c_float)))
        return losses

    def choose(self, inst: TrotterInstance):
        hb = HamilBatch([inst])
        o, r = C.c_int(), C.c_int()
        if not self.lib.qmxPlannerChoose(self.h, hb[0], inst.time, C.byref(o), C.byref(r)):
            raise ValueError("qmxPlannerChoose failed")
        return o.value, r.value

    def save(self, path: str):
        if not self.lib.qmxPlannerSave(self.h, path.encode()):
            raise OSError(f"cannot write {path}")

    def close(self):
        if self.h:
            self.lib.qmxPlannerDestroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def log2_reps_from_ladders(inf2, inf4, eps: float):
    """log2 of the cheapest repetitions for orders 2 and 4 (NaN where no rung
    of the ladder met eps)."""
    L = np.full((len(inf2), 2), np.nan)
    for j in range(len(inf2)):
        for c, (o, inf) in enumerate(((2, inf2[j]), (4, inf4[j]))):
            l2 = label_log2_reps(inf, o, eps)
            if l2 is not None:
                L[j, c] = l2
    return L


def fit_size_curve(L, sizes):
    """Knots of the size curve: log2 n of each distinct size and the mean
    log2 r*_o (orders 2, 4) over its labelled instances."""
    sizes = np.asarray(sizes)
    uniq = np.unique(sizes)
    x = np.log2(uniq.astype(np.float64))
    y = np.array([[np.nanmean(L[sizes == n, c]) for c in range(2)] for n in uniq])
    return x, y


def size_curve(x, y, log2n):
    """The library's piecewise-linear curve (end slopes extended), per order."""
    log2n = np.atleast_1d(np.asarray(log2n, dtype=np.float64))
    k = np.clip(np.searchsorted(x, log2n, side="left") - 1, 0, len(x) - 2)
    x0, x1 = x[k], x[k + 1]
    return y[k] + (y[k + 1] - y[k]) * ((log2n - x0) / (x1 - x0))[:, None]


def size_free_targets(L, sizes, x, y):
    """Targets y_o = log2 r*_o - curve_o(n) and their masks."""
    T = L - size_curve(x, y, np.log2(np.asarray(sizes, dtype=np.float64)))
    M = np.isfinite(T).astype(np.float32)
    return np.nan_to_num(T).astype(np.float32), M


def label_log2_reps(infid, order: int, eps: float):
    """log2 of the cheapest repetitions reaching eps: at the first passing rung
    k (infidelity v <= eps), r* = 2^k (v / eps)^(1/(2 order)), not below the
    failing rung k - 1; None when no rung passes."""
    for k, v in enumerate(infid):
        if np.isfinite(v) and v <= eps:
            if k == 0:
                return 0.0
            r = 2.0 ** k * (max(float(v), 1e-300) / eps) ** (1.0 / (2 * order))
            return float(max(np.log2(r), k - 1))
    return None
