"""Density registers compose every operation on at most two qubits (gates,
controlled gates, channels, Kraus maps) into pending superoperators on
disjoint qubit pairs and sweep the register once per composed block. These tests drive interleavings that
stress that deferral: runs that grow from one to two qubits, operations on a
third qubit that force a sweep, operations the composition does not take
(three-qubit gates, two-qubit dephasing), reads in the middle of a run, two
registers alternating, a register destroyed with work pending, timers around
deferred work and a rejected call in the middle of a run. Every result is
compared with the oracle (QuEST's CPU algorithms restated in C, applying each
operation as its own sweep, QuEST_cpu.c) at 1e-12 max-abs."""
import ctypes as C

import numpy as np
import pytest

from quantumlab_amd import workloads as W
from quantumlab_amd.quest import ComplexMatrix2, ComplexMatrix4, QuESTError, cplx, matrix2, matrix4
from tests.cases import ints

pytestmark = pytest.mark.gpu
TOL = 1e-12


def kraus1(rng, numOps):
    V = W.haar_unitary(rng, 2 * numOps)[:, :2]
    return [V[2 * i:2 * i + 2, :] for i in range(numOps)]


def kraus2(rng, numOps):
    V = W.haar_unitary(rng, 4 * numOps)[:, :4]
    return [V[4 * i:4 * i + 4, :] for i in range(numOps)]


def random_ops(n, count, seed, hot=3):
    """A seeded operation list over the density API. Targets are drawn mostly
    from `hot` qubits so that runs compose; a few operations span three qubits
    or read the register."""
    rng = np.random.default_rng(seed)
    ops = []

    def pick(k, spread=False):
        pool = n if spread or rng.uniform() < 0.2 else min(hot, n)
        return [int(x) for x in rng.choice(pool, size=k, replace=False)]

    kinds = ["cu", "u", "cnot", "cu1", "mscu", "x", "swap", "ps", "cps", "2qu", "deph", "depol", "damp",
             "pauli", "kraus", "kraus2", "dep2", "h", "rx", "crz", "cpf3", "mqu3", "deph2", "prob", "purity"]
    weights = np.array([3, 2, 3, 2, 1, 1, 1, 1, 1, 2, 4, 4, 4, 1, 1, 1, 1, 1, 1, 1, 0.5, 0.5, 0.5, 0.5, 0.3])
    for _ in range(count):
        kind = str(rng.choice(kinds, p=weights / weights.sum()))
        if kind == "cu":
            ops.append((kind, pick(1)[0], *W.compact_pair(rng)))
        elif kind in ("u", "h", "x"):
            ops.append((kind, pick(1)[0], W.haar_unitary(rng, 2)))
        elif kind in ("cnot", "swap"):
            ops.append((kind, *pick(2)))
        elif kind == "cu1":
            ops.append((kind, *pick(2), W.haar_unitary(rng, 2)))
        elif kind == "mscu":
            ops.append((kind, *pick(2), int(rng.integers(0, 2)), W.haar_unitary(rng, 2)))
        elif kind in ("ps", "rx"):
            ops.append((kind, pick(1)[0], float(rng.uniform(-np.pi, np.pi))))
        elif kind in ("cps", "crz"):
            ops.append((kind, *pick(2), float(rng.uniform(-np.pi, np.pi))))
        elif kind == "2qu":
            ops.append((kind, *pick(2), W.haar_unitary(rng, 4)))
        elif kind in ("deph", "damp"):
            ops.append((kind, pick(1)[0], float(rng.uniform(0, 0.4))))
        elif kind == "depol":
            ops.append((kind, pick(1)[0], float(rng.uniform(0, 0.7))))
        elif kind == "pauli":
            px, py, pz = rng.uniform(0, 0.15, size=3)
            ops.append((kind, pick(1)[0], float(px), float(py), float(pz)))
        elif kind == "kraus":
            ops.append((kind, pick(1)[0], kraus1(rng, int(rng.integers(1, 4)))))
        elif kind == "kraus2":
            ops.append((kind, *pick(2), kraus2(rng, int(rng.integers(1, 3)))))
        elif kind == "dep2":
            ops.append((kind, *pick(2), float(rng.uniform(0, 0.9))))
        elif kind == "deph2":
            ops.append((kind, *pick(2), float(rng.uniform(0, 0.7))))
        elif kind == "cpf3":
            ops.append((kind, pick(3, spread=True)))
        elif kind == "mqu3":
            ops.append((kind, pick(3, spread=True), W.haar_unitary(rng, 8)))
        elif kind == "prob":
            ops.append((kind, pick(1, spread=True)[0]))
        else:
            ops.append((kind,))
    return ops


def apply_op(lib, q, op):
    """Apply one operation; reads return their value, everything else None."""
    kind = op[0]
    if kind == "cu":
        lib.call("compactUnitary", q, op[1], cplx(op[2]), cplx(op[3]))
    elif kind == "u":
        lib.call("unitary", q, op[1], matrix2(op[2]))
    elif kind == "h":
        lib.call("hadamard", q, op[1])
    elif kind == "x":
        lib.call("pauliX", q, op[1])
    elif kind == "cnot":
        lib.call("controlledNot", q, op[1], op[2])
    elif kind == "swap":
        lib.call("swapGate", q, op[1], op[2])
    elif kind == "cu1":
        lib.call("controlledUnitary", q, op[1], op[2], matrix2(op[3]))
    elif kind == "mscu":
        lib.call("multiStateControlledUnitary", q, ints([op[1]]), ints([op[3]]), 1, op[2], matrix2(op[4]))
    elif kind == "ps":
        lib.call("phaseShift", q, op[1], op[2])
    elif kind == "rx":
        lib.call("rotateX", q, op[1], op[2])
    elif kind == "cps":
        lib.call("controlledPhaseShift", q, op[1], op[2], op[3])
    elif kind == "crz":
        lib.call("controlledRotateZ", q, op[1], op[2], op[3])
    elif kind == "2qu":
        lib.call("twoQubitUnitary", q, op[1], op[2], matrix4(op[3]))
    elif kind == "deph":
        lib.call("mixDephasing", q, op[1], op[2])
    elif kind == "depol":
        lib.call("mixDepolarising", q, op[1], op[2])
    elif kind == "damp":
        lib.call("mixDamping", q, op[1], op[2])
    elif kind == "pauli":
        lib.call("mixPauli", q, op[1], op[2], op[3], op[4])
    elif kind == "kraus":
        ks = op[2]
        lib.call("mixKrausMap", q, op[1], (ComplexMatrix2 * len(ks))(*[matrix2(k) for k in ks]), len(ks))
    elif kind == "kraus2":
        ks = op[3]
        lib.call("mixTwoQubitKrausMap", q, op[1], op[2], (ComplexMatrix4 * len(ks))(*[matrix4(k) for k in ks]),
                 len(ks))
    elif kind == "dep2":
        lib.call("mixTwoQubitDepolarising", q, op[1], op[2], op[3])
    elif kind == "deph2":
        lib.call("mixTwoQubitDephasing", q, op[1], op[2], op[3])
    elif kind == "cpf3":
        lib.call("multiControlledPhaseFlip", q, ints(op[1]), 3)
    elif kind == "mqu3":
        lib.multiQubitUnitary(q, op[1], op[2])
    elif kind == "prob":
        return lib.calcProbOfOutcome(q, op[1], 0)
    elif kind == "purity":
        return lib.calcPurity(q)
    else:
        raise ValueError(kind)
    return None


def composed_targets(op):
    """Qubits a composed operation acts on; None for operations applied on
    their own (three-qubit gates, two-qubit dephasing) and for reads."""
    kind = op[0]
    if kind in ("cu", "u", "h", "x", "ps", "rx", "deph", "depol", "damp", "pauli", "kraus"):
        return [op[1]]
    if kind in ("cnot", "swap", "cu1", "mscu", "cps", "crz", "2qu", "kraus2", "dep2"):
        return [op[1], op[2]]
    return None


def expected_sweeps(ops, n):
    """The composition rule. Pending blocks act on disjoint sets of at most two
    qubits. An operation joins the one block it touches if the union stays
    within two qubits; a one-qubit operation touching none joins the latest
    one-qubit block; otherwise the blocks it touches are swept and it opens a
    new block. An operation that is not composed, or a read, sweeps every
    block, and so does the final state read. Blocks swept together go two per
    sweep on registers of at least 6 qubits (12 state bits), where an odd
    number of bridged blocks also takes the oldest other block along."""
    def cost(m):
        return (m + 1) // 2 if n >= 6 else m
    blocks, sweeps = [], 0
    for op in ops:
        t = composed_targets(op)
        if t is None:
            sweeps += cost(len(blocks))
            blocks = []
            continue
        hit = [i for i, b in enumerate(blocks) if set(b) & set(t)]
        into = None
        if len(hit) == 1 and len(set(blocks[hit[0]]) | set(t)) <= 2:
            into = hit[0]
        elif not hit and len(t) == 1:
            into = next((i for i in range(len(blocks) - 1, -1, -1) if len(blocks[i]) == 1), None)
        if into is None:
            if n >= 6 and len(hit) % 2 == 1:   # the oldest untouched block goes along
                hit += [i for i in range(len(blocks)) if i not in hit][:1]
            sweeps += cost(len(hit))
            blocks = [b for i, b in enumerate(blocks) if i not in hit] + [list(t)]
        else:
            blocks[into] = blocks[into] + [x for x in t if x not in blocks[into]]
    return sweeps + cost(len(blocks))


def fusion_stats(gpu):
    sweeps, ops = C.c_longlong(0), C.c_longlong(0)
    gpu.lib.qmxDensityFusionStats(C.byref(sweeps), C.byref(ops))
    return sweeps.value, ops.value


@pytest.mark.parametrize("n,seed", [(3, 1), (5, 2), (6, 3), (9, 4)])
def test_random_operation_runs_match_oracle(gpu, cpuref, n, seed):
    ops = random_ops(n, 160, seed)
    s0 = fusion_stats(gpu)
    results = []
    for lib in (gpu, cpuref):
        q = lib.qureg(n, density=True)
        lib.initPlusState(q)
        reads = [apply_op(lib, q, op) for op in ops]
        results.append((lib.state(q), np.array([r for r in reads if r is not None])))
        lib.destroy(q)
    s1 = fusion_stats(gpu)
    assert np.max(np.abs(results[0][0] - results[1][0])) < TOL
    assert np.max(np.abs(results[0][1] - results[1][1]), initial=0.0) < TOL
    sweeps, absorbed = s1[0] - s0[0], s1[1] - s0[1]
    assert absorbed == sum(composed_targets(op) is not None for op in ops)
    assert sweeps == expected_sweeps(ops, n) < absorbed


def test_two_registers_interleaved(gpu, cpuref):
    n = 5
    rng = np.random.default_rng(21)
    opsA, opsB = random_ops(n, 80, 22), random_ops(n, 80, 23)
    order = rng.permutation([0] * len(opsA) + [1] * len(opsB))
    out = []
    for lib in (gpu, cpuref):
        qa, qb = lib.qureg(n, density=True), lib.qureg(n, density=True)
        lib.initPlusState(qa)
        lib.initClassicalState(qb, 5)
        ia = ib = 0
        reads = []
        for which in order:
            if which == 0:
                r = apply_op(lib, qa, opsA[ia])
                ia += 1
            else:
                r = apply_op(lib, qb, opsB[ib])
                ib += 1
            if r is not None:
                reads.append(r)
            if (ia + ib) % 37 == 0:
                reads.append(lib.calcDensityInnerProduct(qa, qb))
        out.append((lib.state(qa), lib.state(qb), np.array(reads)))
        lib.destroy(qa)
        lib.destroy(qb)
    for a, b in zip(out[0], out[1]):
        assert np.max(np.abs(a - b)) < TOL


def test_destroy_with_pending_work(gpu, cpuref):
    n = 4
    # work left pending on a register that is then destroyed is dropped, and
    # the next register (possibly at the same device address) is unaffected
    rho = gpu.qureg(n, density=True)
    gpu.call("mixDamping", rho, 1, 0.3)
    gpu.call("hadamard", rho, 2)
    gpu.destroy(rho)
    out = []
    for lib in (gpu, cpuref):
        q = lib.qureg(n, density=True)
        for op in random_ops(n, 40, 31):
            apply_op(lib, q, op)
        out.append(lib.state(q))
        lib.destroy(q)
    assert np.max(np.abs(out[0] - out[1])) < TOL


def test_clone_and_host_copies_see_pending_work(gpu, cpuref):
    n = 4
    out = []
    for lib in (gpu, cpuref):
        a, b = lib.qureg(n, density=True), lib.qureg(n, density=True)
        lib.call("hadamard", a, 0)
        lib.call("mixDepolarising", a, 0, 0.2)
        lib.call("cloneQureg", b, a)
        lib.call("controlledNot", b, 0, 3)
        lib.call("mixDamping", b, 3, 0.1)
        amp = lib.getDensityAmp(b, 1, 9)
        out.append((lib.state(a), lib.state(b), np.array([amp.real, amp.imag])))
        lib.destroy(a)
        lib.destroy(b)
    for x, y in zip(out[0], out[1]):
        assert np.max(np.abs(x - y)) < TOL


def test_rejected_call_inside_a_run(gpu, cpuref):
    n = 3
    out = []
    for lib in (gpu, cpuref):
        q = lib.qureg(n, density=True)
        lib.call("hadamard", q, 0)
        lib.call("mixDephasing", q, 0, 0.1)
        if lib is gpu:
            # prob above 1/2 is rejected and applies nothing (the oracle exits on
            # invalid input, so only the library sees this call)
            with pytest.raises(QuESTError):
                lib.call("mixDephasing", q, 1, 0.9)
        lib.call("mixDamping", q, 0, 0.2)
        out.append(lib.state(q))
        lib.destroy(q)
    assert np.max(np.abs(out[0] - out[1])) < TOL


def test_timer_encloses_deferred_work(gpu):
    n = 11
    q = gpu.qureg(n, density=True)
    gpu.qmxSync()
    gpu.qmxTimerBegin()
    for t in range(n):
        gpu.call("hadamard", q, t)
        gpu.call("mixDamping", q, t, 0.05)
    ms = gpu.qmxTimerEndMs()
    # 11 sweeps over a 22-qubit vector (64 MiB): well above launch latency
    assert ms > 0.05
    assert abs(gpu.calcTotalProb(q) - 1) < 1e-12
    gpu.destroy(q)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_noisy_brick_at_twelve_qubits(gpu, cpuref, prec):
    """The benchmark's workload shape on a 24-bit vector (non-temporal sweeps,
    XCD-ordered grid), against the oracle; FP32 storage within 2e-6 of it."""
    from quantumlab_amd.quest import QMX_PREC_FP32, QMX_PREC_FP64
    n = 12
    gates = W.noisy_brick_circuit(n, 2, seed=5)
    q = gpu.qureg(n, density=True, precision=QMX_PREC_FP64 if prec == "fp64" else QMX_PREC_FP32)
    W.apply_gates(gpu, q, gates)
    got = gpu.state(q)
    gpu.destroy(q)
    r = cpuref.qureg(n, density=True)
    W.apply_gates(cpuref, r, gates)
    want = cpuref.state(r)
    cpuref.destroy(r)
    assert np.max(np.abs(got - want)) < (TOL if prec == "fp64" else 2e-6)


@pytest.mark.parametrize("shape", ["2q+2q", "2q+1q", "1q+2q", "low+high"])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_paired_sweep_shapes(gpu, cpuref, shape, prec):
    """Two pending blocks swept in one pass: superoperators on 4 + 4 and 4 + 2
    bits (two one-qubit blocks never coexist: a lone one-qubit operation joins
    the pending one-qubit block), with blocks on the lowest and highest
    qubits. The read sweeps both in one pass; the state matches the oracle."""
    from quantumlab_amd.quest import QMX_PREC_FP32, QMX_PREC_FP64
    n = 7
    rng = np.random.default_rng(hash(shape) % 1000)
    blocks = {"2q+2q": ([1, 4], [2, 6]), "2q+1q": ([0, 5], [3]), "1q+2q": ([0], [4, 6]),
              "low+high": ([0, 1], [5, 6])}[shape]
    ops = []
    for qs in blocks:
        if len(qs) == 2:
            ops += [("2qu", qs[0], qs[1], W.haar_unitary(rng, 4)), ("dep2", qs[0], qs[1], 0.3),
                    ("damp", qs[1], 0.2)]
        else:
            ops += [("u", qs[0], W.haar_unitary(rng, 2)), ("depol", qs[0], 0.4), ("deph", qs[0], 0.1)]
    q = gpu.qureg(n, density=True, precision=QMX_PREC_FP64 if prec == "fp64" else QMX_PREC_FP32)
    r = cpuref.qureg(n, density=True)
    for lib, reg in ((gpu, q), (cpuref, r)):
        lib.initPlusState(reg)
        lib.call("hadamard", reg, 3)
        lib.calcPurity(reg)                                  # nothing pending before the blocks
    s0 = fusion_stats(gpu)
    for op in ops:
        apply_op(gpu, q, op)
        apply_op(cpuref, r, op)
    got_p = gpu.calcPurity(q)
    s1 = fusion_stats(gpu)
    assert s1[0] - s0[0] == 1                                # both blocks in one sweep
    got, want = gpu.state(q), cpuref.state(r)
    assert abs(got_p - cpuref.calcPurity(r)) < (TOL if prec == "fp64" else 2e-6)
    assert np.max(np.abs(got - want)) < (TOL if prec == "fp64" else 2e-6)
    gpu.destroy(q)
    cpuref.destroy(r)
