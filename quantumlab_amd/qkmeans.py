"""Batched quantum k-means (swap tests) on libquest_mi355x.

Dataset of the QKMeans workload: `num_clusters` centroids with N(0, 1)
coordinates in 2^max_q dimensions; each point picks a cluster uniformly and a
register size q uniformly in [q_lo, q_hi]; the point is the centroid's first
2^q coordinates plus N(0, noise^2) noise (points are L2-normalised by the
plan). Centroids are compared with a point at the point's size: truncated to
2^q entries and L2-normalised. Initial k-means centroids are the true ones
plus N(0, init_noise^2).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np


@dataclass
class Dataset:
    points: np.ndarray      # packed float64, point p at offsets[p] : offsets[p] + 2^q[p]
    offsets: np.ndarray     # int64
    qubits: np.ndarray      # int32, register size of each point
    labels: np.ndarray      # int32, generating cluster
    centroids: np.ndarray   # (C, 2^max_q) true centroids
    init: np.ndarray        # (C, 2^max_q) initial k-means centroids
    max_q: int

    def point(self, p: int) -> np.ndarray:
        return self.points[self.offsets[p]:self.offsets[p] + (1 << int(self.qubits[p]))]


def make_dataset(num_points=16384, num_clusters=16, q_lo=4, q_hi=10, noise=0.3, init_noise=0.5,
                 seed=7) -> Dataset:
    rng = np.random.default_rng(seed)
    dim = 1 << q_hi
    cents = rng.standard_normal((num_clusters, dim))
    labels = rng.integers(0, num_clusters, size=num_points).astype(np.int32)
    qs = rng.integers(q_lo, q_hi + 1, size=num_points).astype(np.int32)
    sizes = (1 << qs.astype(np.int64))
    offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    pts = np.empty(int(sizes.sum()))
    for p in range(num_points):
        D = int(sizes[p])
        pts[offsets[p]:offsets[p] + D] = cents[labels[p], :D] + noise * rng.standard_normal(D)
    init = cents + init_noise * rng.standard_normal(cents.shape)
    return Dataset(pts, offsets, qs, labels, cents, init, q_hi)


class SwapPlan:
    """A qmxSwapPlan: points and centroids resident on the GPU."""

    def __init__(self, qe, data: Dataset, precision: int, centroids: np.ndarray | None = None):
        self.qe = qe
        self.lib = qe.lib
        self.data = data
        self.C = data.init.shape[0]
        self.P = len(data.qubits)
        cents = np.ascontiguousarray(data.init if centroids is None else centroids, dtype=np.float64)
        qs = np.ascontiguousarray(data.qubits, dtype=np.int32)
        pts = np.ascontiguousarray(data.points, dtype=np.float64)
        self.plan = self.lib.qmxSwapPlanCreate(self.P, qs.ctypes.data_as(C.POINTER(C.c_int)),
                                               pts.ctypes.data_as(C.POINTER(C.c_double)), self.C, data.max_q,
                                               cents.ctypes.data_as(C.POINTER(C.c_double)), precision)
        if not self.plan:
            raise ValueError("qmxSwapPlanCreate rejected the sizes")

    def autotune(self) -> int:
        return self.lib.qmxSwapPlanAutotune(self.plan)

    def run(self, iterations: int, graph: bool = True):
        self.lib.qmxSwapPlanRun(self.plan, iterations, 1 if graph else 0)

    def p0(self) -> np.ndarray:
        out = np.empty(self.P * self.C)
        self.lib.qmxSwapPlanP0(self.plan, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out.reshape(self.P, self.C)

    def assignments(self) -> np.ndarray:
        out = np.empty(self.P, dtype=np.int32)
        self.lib.qmxSwapPlanAssignments(self.plan, out.ctypes.data_as(C.POINTER(C.c_int)))
        return out

    def centroids(self) -> np.ndarray:
        out = np.empty(self.C * (1 << self.data.max_q))
        self.lib.qmxSwapPlanCentroids(self.plan, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out.reshape(self.C, -1)

    def set_centroids(self, centroids: np.ndarray):
        cents = np.ascontiguousarray(centroids, dtype=np.float64)
        if cents.shape != (self.C, 1 << self.data.max_q):
            raise ValueError(f"centroids must be {self.C} x {1 << self.data.max_q}")
        self.lib.qmxSwapPlanSetCentroids(self.plan, cents.ctypes.data_as(C.POINTER(C.c_double)))

    def point_state(self, p: int) -> np.ndarray:
        out = np.empty(1 << int(self.data.qubits[p]))
        self.lib.qmxSwapPlanPointState(self.plan, p, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def rows_per_tile(self, q: int) -> int:
        return self.lib.qmxSwapPlanRowsPerTile(self.plan, q)

    def num_tiles(self) -> int:
        return self.lib.qmxSwapPlanNumTiles(self.plan)

    def close(self):
        if self.plan:
            self.lib.qmxSwapPlanDestroy(self.plan)
            self.plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def swap_test_amplitude_evaluations(data: Dataset) -> float:
    """Ancilla-0 amplitudes evaluated per iteration: 2^(2q) per swap test."""
    return float(np.sum((1 << (2 * data.qubits.astype(np.int64))))) * data.init.shape[0]
