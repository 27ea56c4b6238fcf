"""Oracle of QuantumLab's operation loop — test infrastructure only.

Restates, over QuEST's C API of any QuEST-ABI library (in tests: the C
restatement oracle/build/libcpuref.so or QuEST's own CPU build oracle/_ref),
the loop the circuit entry points of the MI355X library replace:

- QLGate::PerformBasicOperation (Code/QuantumLab/Circuit/QLGate.cpp:652-803):
  one QuEST call per EBasicOperation (codes QLGate.h:18-104); EBO_CC returns
  collapseToOutcome's probability; noise only on density registers; EBO_Phase
  and EBO_Composite are not dispatched;
- the density simulator's after-gate noise
  (Code/QuantumLab/Simulator/QLSimulatorDensityMatrix.cpp:88-153), applied
  after every operation that is neither noise nor a collapse;
- the simulators' probability product (QLSimulatorVector.cpp:53-59).

Only tests/ import this module; the product never does.
"""
from __future__ import annotations

from typing import Sequence

import numpy as np

# EBasicOperation codes (QLGate.h:18-104)
H, X, Y, Z, P, PHASE, RX, RY, RZ, CX, CY, CZ, CP, CRX, CRY, CRZ, CC = range(17)
(NOISE_DAMPING, NOISE_DEPHASING, NOISE_DEPOLARISING, NOISE_MIX_X, NOISE_MIX_Y, NOISE_MIX_Z, NOISE_MIX_ALL,
 NOISE_TWO_DEPHASING, NOISE_TWO_DEPOLARISING, COMPOSITE) = range(17, 27)
NOISE = range(NOISE_DAMPING, NOISE_TWO_DEPOLARISING + 1)
ONE_QUBIT = (H, X, Y, Z, P, RX, RY, RZ)
TWO_QUBIT = (CX, CY, CZ, CP, CRX, CRY, CRZ)


def perform(qe, q, op: tuple) -> float:
    """PerformBasicOperation (QLGate.cpp:652-803) for op = (code, qubits, param)."""
    code, qs, t = op
    dm = bool(q.isDensityMatrix)
    one = {H: "hadamard", X: "pauliX", Y: "pauliY", Z: "pauliZ"}
    onep = {P: "phaseShift", RX: "rotateX", RY: "rotateY", RZ: "rotateZ"}
    two = {CX: "controlledNot", CY: "controlledPauliY", CZ: "controlledPhaseFlip"}
    twop = {CP: "controlledPhaseShift", CRX: "controlledRotateX", CRY: "controlledRotateY",
            CRZ: "controlledRotateZ"}
    if code in one:
        qe.call(one[code], q, qs[0])
    elif code in onep:
        qe.call(onep[code], q, qs[0], t)
    elif code in two:
        qe.call(two[code], q, qs[0], qs[1])
    elif code in twop:
        qe.call(twop[code], q, qs[0], qs[1], t)
    elif code == CC:
        return float(qe.call("collapseToOutcome", q, qs[0], int(t)))
    elif code in NOISE:
        if not dm:
            return 1.0  # "Noise gate encountered, but the state is pure, ignored"
        if code == NOISE_DAMPING:
            qe.call("mixDamping", q, qs[0], t)
        elif code == NOISE_DEPHASING:
            qe.call("mixDephasing", q, qs[0], t)
        elif code == NOISE_DEPOLARISING:
            qe.call("mixDepolarising", q, qs[0], t)
        elif code == NOISE_MIX_X:
            qe.call("mixPauli", q, qs[0], t, 0.0, 0.0)
        elif code == NOISE_MIX_Y:
            qe.call("mixPauli", q, qs[0], 0.0, t, 0.0)
        elif code == NOISE_MIX_Z:
            qe.call("mixPauli", q, qs[0], 0.0, 0.0, t)
        elif code == NOISE_MIX_ALL:
            qe.call("mixPauli", q, qs[0], t, t, t)
        elif code == NOISE_TWO_DEPHASING:
            qe.call("mixTwoQubitDephasing", q, qs[0], qs[1], t)
        else:
            qe.call("mixTwoQubitDepolarising", q, qs[0], qs[1], t)
    return 1.0  # PHASE, COMPOSITE: "not supported type!"


def noise_after_gate(qe, q, op: tuple, noise: dict):
    """QLSimulatorDensityMatrix.cpp:88-153 for one operation."""
    code, qs, _ = op
    if code in NOISE or code == CC:
        return
    tiny = np.finfo(np.float64).tiny
    get = lambda k: float(noise.get(k, 0.0))
    if get("damping") > tiny:
        for t in qs:
            qe.call("mixDamping", q, t, get("damping"))
    if get("dephasing") > tiny:
        if get("twoQubitDephasing") < tiny:
            for t in qs:
                qe.call("mixDephasing", q, t, get("dephasing"))
        elif len(qs) == 1:
            qe.call("mixDephasing", q, qs[0], get("dephasing"))
    if get("depolarising") > tiny:
        if get("twoQubitDepolarising") < tiny:
            for t in qs:
                qe.call("mixDepolarising", q, t, get("depolarising"))
        elif len(qs) == 1:
            qe.call("mixDepolarising", q, qs[0], get("depolarising"))
    if get("mixPauli") > tiny:
        for t in qs:
            p = get("mixPauli")
            qe.call("mixPauli", q, t, p, p, p)
    if get("twoQubitDepolarising") > tiny and len(qs) == 2:
        qe.call("mixTwoQubitDepolarising", q, qs[0], qs[1], get("twoQubitDepolarising"))
    if get("twoQubitDephasing") > tiny and len(qs) == 2:
        qe.call("mixTwoQubitDephasing", q, qs[0], qs[1], get("twoQubitDephasing"))


def run_ops(qe, q, ops: Sequence[tuple], noise: dict | None = None) -> float:
    """The simulators' loop: PerformBasicOperation per op (plus after-gate noise
    on density registers); returns the product of the collapse probabilities."""
    prob = 1.0
    for op in ops:
        prob *= perform(qe, q, op)
        if noise and q.isDensityMatrix:
            noise_after_gate(qe, q, op, noise)
    return prob


def random_ops(n: int, count: int, seed: int, noise: bool = False, collapses: int = 0) -> list:
    """Seeded op list over every gate code (and the noise codes when noise=True);
    `collapses` EBO_CC operations at random positions (outcome 0 or 1 chosen
    later by the caller from the probabilities, see collapse_outcomes)."""
    rng = np.random.default_rng(seed)
    codes = list(ONE_QUBIT) + list(TWO_QUBIT)
    if noise:
        codes += list(NOISE)
    ops = []
    for _ in range(count):
        c = int(rng.choice(codes))
        if c in (NOISE_TWO_DEPHASING, NOISE_TWO_DEPOLARISING) or c in TWO_QUBIT:
            qs = [int(x) for x in rng.choice(n, size=2, replace=False)]
        else:
            qs = [int(rng.integers(0, n))]
        if c in (NOISE_DAMPING,):
            t = float(rng.uniform(0, 0.2))
        elif c in (NOISE_DEPHASING, NOISE_MIX_X, NOISE_MIX_Y, NOISE_MIX_Z):
            t = float(rng.uniform(0, 0.1))
        elif c == NOISE_MIX_ALL:
            t = float(rng.uniform(0, 0.05))
        elif c in (NOISE_DEPOLARISING, NOISE_TWO_DEPHASING, NOISE_TWO_DEPOLARISING):
            t = float(rng.uniform(0, 0.1))
        else:
            t = float(rng.uniform(-np.pi, np.pi))
        ops.append((c, qs, t))
    for _ in range(collapses):
        pos = int(rng.integers(1, len(ops)))
        ops.insert(pos, (CC, [int(rng.integers(0, n))], -1.0))  # outcome filled in later
    return ops
