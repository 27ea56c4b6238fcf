"""Shot sampling (qmxSampleShots) and amplitude encoding (qmxAmplitudeEncodeReal).

Sampling is checked against QuantumLab's own shot loop (QLSimulatorMeasure.cpp:
113-140: restore the register, measure() each listed qubit in order) run on the
oracle with the same seeds: the outcome sequences and the generator's next draw
must be identical. Encoding is checked against the reference's gate-level
encoder (angle tree + Ry / Gray-code ladder) run through the oracle's QuEST API
at 1e-12."""
import ctypes as C

import numpy as np
import pytest

from oracle import swaptest_oracle as O
from quantumlab_amd import qkmeans as K
from quantumlab_amd import workloads as W
from quantumlab_amd.quest import QMX_PREC_FP32, QMX_PREC_FP64, QuESTError, qreal
from tests.cases import channel_circuit, ints

pytestmark = pytest.mark.gpu
TOL = 1e-12


def reference_shot_loop(lib, q, saved, qubits, shots, until=0):
    """The reference's loop on any QuEST-ABI library: restore, measure each qubit."""
    out, counts = [], {}
    for s in range(shots):
        if s:
            lib.set_state(q, saved)
        res = 0
        for j, t in enumerate(qubits):
            if lib.measure(q, t) == 1:
                res |= 1 << j
        out.append(res)
        counts[res] = counts.get(res, 0) + 1
        if until and counts[res] >= until:
            break
    return np.array(out)


def sample(gpu, q, qubits, shots, until=0):
    buf = (C.c_int * max(1, shots))()
    n = gpu.qmxSampleShots(q, ints(qubits), len(qubits), shots, until, buf)
    return np.array(buf[:n])


@pytest.mark.parametrize("until", [0, 12])
def test_shots_match_the_reference_loop(gpu, cpuref, until):
    n, qubits, shots = 9, [4, 0, 7, 2], 300
    gates = W.brick_circuit(n, 3, seed=17)
    res, nxt = [], []
    for lib in (gpu, cpuref):
        lib.seed([77, 88])
        q = lib.qureg(n)
        W.apply_gates(lib, q, gates)
        saved = lib.state(q)
        if lib is gpu:
            res.append(sample(gpu, q, qubits, shots, until))
            assert np.max(np.abs(gpu.state(q) - saved)) == 0.0   # the register is untouched
        else:
            res.append(reference_shot_loop(cpuref, q, saved, qubits, shots, until))
            cpuref.set_state(q, saved)
        nxt.append(lib.measure(q, 5))                            # the generator is where the loop leaves it
        lib.destroy(q)
    assert len(res[0]) == len(res[1]) and (until == 0 or len(res[0]) < shots)
    assert np.array_equal(res[0], res[1])
    assert nxt[0] == nxt[1]


def test_density_shots_match_the_reference_loop(gpu, cpuref):
    n, qubits, shots = 5, [3, 1], 200
    res = []
    for lib in (gpu, cpuref):
        lib.seed([5, 6, 7])
        rho, pure = lib.qureg(n, density=True), lib.qureg(n)
        channel_circuit(lib, rho, pure, n, 2, seed=23)
        saved = lib.state(rho)
        res.append(sample(gpu, rho, qubits, shots) if lib is gpu
                   else reference_shot_loop(cpuref, rho, saved, qubits, shots))
        lib.destroy(rho)
        lib.destroy(pure)
    assert np.array_equal(res[0], res[1])


def test_certain_outcomes_draw_nothing(gpu):
    q = gpu.qureg(6)
    gpu.call("pauliX", q, 2)
    gpu.seed([9, 9])
    assert np.all(sample(gpu, q, [2, 0], 50) == 1)              # qubit 2 is 1, qubit 0 is 0: no draws
    gpu.call("hadamard", q, 4)
    a = gpu.measure(q, 4)
    gpu.seed([9, 9])
    q2 = gpu.qureg(6)
    gpu.call("hadamard", q2, 4)
    assert gpu.measure(q2, 4) == a
    gpu.destroy(q)
    gpu.destroy(q2)


def test_shot_frequencies_follow_the_distribution(gpu):
    n, qubits, shots = 12, [1, 6, 10], 20000
    q = gpu.qureg(n)
    W.apply_gates(gpu, q, W.brick_circuit(n, 4, seed=3))
    probs = np.zeros(8)
    gpu.lib.calcProbOfAllOutcomes(probs.ctypes.data_as(C.POINTER(qreal)), q, ints(qubits), 3)
    freq = np.bincount(sample(gpu, q, qubits, shots), minlength=8) / shots
    assert np.all(np.abs(freq - probs) < 5 * np.sqrt(probs * (1 - probs) / shots) + 1e-12)
    gpu.destroy(q)


def test_sampling_rejects_invalid_input(gpu):
    q = gpu.qureg(4)
    with pytest.raises(QuESTError, match="unique"):
        sample(gpu, q, [1, 1], 5)
    with pytest.raises(QuESTError, match="shots"):
        gpu.qmxSampleShots(q, ints([1]), 1, -1, 0, (C.c_int * 1)())
    gpu.destroy(q)


# ---- amplitude encoding ------------------------------------------------------------

def encode(lib, q, v):
    v = np.ascontiguousarray(v, dtype=np.float64)
    lib.qmxAmplitudeEncodeReal(q, v.ctypes.data_as(C.POINTER(qreal)), int(v.size).bit_length() - 1)


@pytest.mark.parametrize("qb", [1, 4, 8, 12])
def test_encoding_matches_the_reference_gate_list(gpu, cpuref, qb):
    rng = np.random.default_rng(qb)
    v = rng.standard_normal(1 << qb)
    v[rng.integers(0, v.size, size=v.size // 4)] = 0.0      # zero blocks take the epsilon path
    n = qb + 1
    q = gpu.qureg(n)
    gpu.call("hadamard", q, 0)                               # any prior state is replaced
    encode(gpu, q, v)
    got = gpu.state(q)
    r = cpuref.qureg(n)
    O.run_gates_quest(cpuref, r, O.ae_gates(v, list(range(qb))))
    want = cpuref.state(r)
    assert np.max(np.abs(got - want)) < TOL
    assert np.max(np.abs(got[:v.size] - v / np.linalg.norm(v))) < 1e-13
    gpu.destroy(q)
    cpuref.destroy(r)


def test_encoding_into_fp32_and_zero_vector(gpu):
    v = np.random.default_rng(4).standard_normal(64)
    q = gpu.qureg(8, precision=QMX_PREC_FP32)
    encode(gpu, q, v)
    want = np.zeros(256)
    want[:64] = v / np.linalg.norm(v)
    assert np.max(np.abs(gpu.state(q) - want)) < 1e-7
    encode(gpu, q, np.zeros(16))
    want = np.zeros(256)
    want[0] = 1.0
    assert np.max(np.abs(gpu.state(q) - want)) < 1e-7
    gpu.destroy(q)


def test_encoding_rejects_invalid_input(gpu):
    rho = gpu.qureg(3, density=True)
    with pytest.raises(QuESTError, match="state-vector"):
        encode(gpu, rho, np.ones(4))
    gpu.destroy(rho)
    q = gpu.qureg(3)
    with pytest.raises(QuESTError, match="encode"):
        encode(gpu, q, np.ones(16))
    gpu.destroy(q)


def test_swap_plan_with_twelve_qubit_registers(gpu):
    """12-qubit points need 96 KiB of dynamic LDS for the encoder."""
    data = K.make_dataset(num_points=6, num_clusters=2, q_lo=12, q_hi=12, seed=8)
    plan = K.SwapPlan(gpu, data, QMX_PREC_FP64)
    plan.run(1, graph=False)
    exact = O.p0_matrix(O.normalised_points(data), data.init)
    assert np.max(np.abs(plan.p0() - exact)) < 1e-12
    plan.close()
