"""Uniformly controlled one-qubit gates (qmxUniformlyControlledRotation /
qmxUniformlyControlledUnitary) and the complex-vector amplitude encoding
(qmxAmplitudeEncodeComplex), through the C ABI on the GPU.

Oracle: QuantumLab's own gate lists for FRy / FRz / FRp (FRGate.cu:181-330,
460-515) and for AmplitudeEncodeOneVector (AmplitudeEncodeCuda.cu:691-724 with
ExchangeToYZGate and FRyz), restated in oracle/encoding_oracle.py and run gate
by gate through the oracle's QuEST API on the same input state; and the
operator's definition (matrix k where the controls read k) by brute force.
The restatements themselves are pinned on the CPU (tests/test_encoding_oracle.py).
FP64 registers at 1e-12, FP32 registers at 2e-6."""
import ctypes as C

import numpy as np
import pytest

from oracle import encoding_oracle as E
from quantumlab_amd.quest import (QMX_PREC_FP32, QMX_UNIFORM_PHASE, QMX_UNIFORM_RY, QMX_UNIFORM_RZ, QuESTError)
from tests.cases import ints, random_state

pytestmark = pytest.mark.gpu

KIND = {"ry": QMX_UNIFORM_RY, "rz": QMX_UNIFORM_RZ, "p": QMX_UNIFORM_PHASE}


def uniform_rotation(qe, q, controls, target, angles, kind):
    a = np.ascontiguousarray(angles, dtype=qe.np_real)
    qe.call("qmxUniformlyControlledRotation", q, ints(controls), len(controls), target,
            a.ctypes.data_as(C.POINTER(qe.t.qreal)), KIND[kind])


def fr_list(kind, angles, bits):
    return E.frp_gates(angles, bits) if kind == "p" else E.fr_gates(kind, angles, bits)


@pytest.mark.parametrize("kind", ["ry", "rz", "p"])
@pytest.mark.parametrize("width", [1, 2, 4, 7])
def test_uniform_rotation_matches_the_fr_gate_list(gpu, cpuref, kind, width):
    n = 10
    rng = np.random.default_rng(100 * width + len(kind))
    bits = [int(b) for b in rng.permutation(n)[:width]]
    angles = rng.uniform(-4, 4, 1 << (width - 1))
    psi = random_state(n, width)
    q, r = gpu.qureg(n), cpuref.qureg(n)
    gpu.set_state(q, psi)
    cpuref.set_state(r, psi)
    controls, target = E.fr_controls(kind, bits)
    uniform_rotation(gpu, q, controls, target, angles, kind)
    E.run_gates_quest(cpuref, r, fr_list(kind, angles, bits))
    assert np.max(np.abs(gpu.state(q) - cpuref.state(r))) < 1e-12
    gpu.destroy(q)
    cpuref.destroy(r)


@pytest.mark.parametrize("kind", ["ry", "rz", "p"])
def test_uniform_rotation_on_density_and_fp32_registers(gpu, cpuref, kind):
    n, width = 5, 4
    rng = np.random.default_rng(7 + len(kind))
    bits = [3, 0, 4, 1]
    angles = rng.uniform(-4, 4, 1 << (width - 1))
    controls, target = E.fr_controls(kind, bits)
    # density register: the gate on the rows, its conjugate on the columns
    psi = random_state(n, 3)
    pure, rho = gpu.qureg(n), gpu.qureg(n, density=True)
    rpure, rrho = cpuref.qureg(n), cpuref.qureg(n, density=True)
    gpu.set_state(pure, psi)
    cpuref.set_state(rpure, psi)
    gpu.initPureState(rho, pure)
    cpuref.initPureState(rrho, rpure)
    gpu.mixDephasing(rho, 2, 0.2)                      # a mixed state, pending density work first
    cpuref.mixDephasing(rrho, 2, 0.2)
    uniform_rotation(gpu, rho, controls, target, angles, kind)
    E.run_gates_quest(cpuref, rrho, fr_list(kind, angles, bits))
    assert np.max(np.abs(gpu.state(rho) - cpuref.state(rrho))) < 1e-12
    # FP32 state vector
    n = 12
    psi = random_state(n, 4)
    q, r = gpu.qureg(n, precision=QMX_PREC_FP32), cpuref.qureg(n)
    gpu.set_state(q, psi)
    cpuref.set_state(r, psi)
    bits = [11, 0, 6, 2, 9]
    angles = rng.uniform(-4, 4, 16)
    controls, target = E.fr_controls(kind, bits)
    uniform_rotation(gpu, q, controls, target, angles, kind)
    E.run_gates_quest(cpuref, r, fr_list(kind, angles, bits))
    assert np.max(np.abs(gpu.state(q) - cpuref.state(r))) < 2e-6
    for x in (pure, rho, q):
        gpu.destroy(x)
    for x in (rpure, rrho, r):
        cpuref.destroy(x)


def haar2(rng):
    z = rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))
    qm, rm = np.linalg.qr(z)
    return qm * (np.diag(rm) / np.abs(np.diag(rm)))


@pytest.mark.parametrize("controls,target", [([], 3), ([0], 5), ([7, 1, 4], 2), ([2, 9, 0, 5, 8, 1], 6),
                                             (list(range(1, 12)), 0)])
def test_uniform_unitary_matches_its_definition(gpu, controls, target):
    n = 12
    rng = np.random.default_rng(len(controls) + target)
    mats = [haar2(rng) for _ in range(1 << len(controls))]
    psi = random_state(n, 11)
    q = gpu.qureg(n)
    gpu.set_state(q, psi)
    ops = (gpu.t.ComplexMatrix2 * len(mats))(*[gpu.t.matrix2(m) for m in mats])
    gpu.call("qmxUniformlyControlledUnitary", q, ints(controls), len(controls), target, ops)
    want = E.uniform_dense(psi.astype(complex), n, controls, target, mats)
    assert np.max(np.abs(gpu.state(q) - want)) < 1e-12
    gpu.destroy(q)


def test_uniform_gates_reject_invalid_input(gpu):
    q = gpu.qureg(6)
    a = np.zeros(64)
    with pytest.raises(QuESTError, match="unique"):
        uniform_rotation(gpu, q, [1, 1], 0, a, "ry")
    with pytest.raises(QuESTError, match="[Cc]ontrol qubit cannot also be a target|target"):
        uniform_rotation(gpu, q, [1, 2], 2, a, "ry")
    with pytest.raises(QuESTError, match="Invalid target"):
        uniform_rotation(gpu, q, [], 6, a, "rz")
    with pytest.raises(QuESTError, match="kind"):
        gpu.call("qmxUniformlyControlledRotation", q, ints([1]), 1, 0,
                 a.ctypes.data_as(C.POINTER(gpu.t.qreal)), 7)
    bad = np.array([[1.0, 0], [0, 1.1]])
    ops = (gpu.t.ComplexMatrix2 * 2)(gpu.t.matrix2(np.eye(2)), gpu.t.matrix2(bad))
    with pytest.raises(QuESTError, match="unitary"):
        gpu.call("qmxUniformlyControlledUnitary", q, ints([3]), 1, 0, ops)
    big = gpu.qureg(18)
    with pytest.raises(QuESTError, match="<=16"):
        uniform_rotation(gpu, big, list(range(1, 18)), 0, np.zeros(1 << 17), "ry")
    gpu.destroy(big)
    gpu.destroy(q)


# ---- complex amplitude encoding ---------------------------------------------

def encode_complex(qe, q, v):
    re = np.ascontiguousarray(v.real, dtype=qe.np_real)
    im = np.ascontiguousarray(v.imag, dtype=qe.np_real)
    qe.call("qmxAmplitudeEncodeComplex", q, re.ctypes.data_as(C.POINTER(qe.t.qreal)),
            im.ctypes.data_as(C.POINTER(qe.t.qreal)), int(v.size).bit_length() - 1)


@pytest.mark.parametrize("qb", [1, 3, 6, 10, 12])
def test_complex_encoding_matches_the_reference_gate_list(gpu, cpuref, qb):
    rng = np.random.default_rng(qb)
    v = rng.standard_normal(1 << qb) + 1j * rng.standard_normal(1 << qb)
    if qb > 2:
        v[rng.integers(0, v.size, size=v.size // 4)] = 0.0   # zero blocks take the epsilon path
        v[: 1 << (qb - 2)] = 0.0                               # a whole zero subtree
    n = qb + 1
    q = gpu.qureg(n)
    gpu.call("hadamard", q, 0)                                 # any prior state is replaced
    encode_complex(gpu, q, v)
    r = cpuref.qureg(n)
    E.run_gates_quest(cpuref, r, E.ae_gates_complex(v, list(range(qb))))
    got, want = gpu.state(q), cpuref.state(r)
    assert np.max(np.abs(got - want)) < 1e-12
    if qb <= 6:
        assert np.max(np.abs(got[: 1 << qb] - E.ae_complex_closed_form(v))) < 1e-12
    gpu.destroy(q)
    cpuref.destroy(r)


def test_complex_encoding_on_fp32_and_rejections(gpu, cpuref):
    qb = 9
    rng = np.random.default_rng(3)
    v = rng.standard_normal(1 << qb) + 1j * rng.standard_normal(1 << qb)
    q = gpu.qureg(qb + 2, precision=QMX_PREC_FP32)
    encode_complex(gpu, q, v)
    want = np.zeros(1 << (qb + 2), dtype=complex)
    want[: 1 << qb] = E.ae_complex_closed_form(v)
    assert np.max(np.abs(gpu.state(q) - want)) < 2e-6
    rho = gpu.qureg(4, density=True)
    with pytest.raises(QuESTError, match="state-vector|statevector|state vector"):
        encode_complex(gpu, rho, v[:8])
    with pytest.raises(QuESTError, match="encode"):
        encode_complex(gpu, q, np.ones(1 << 12, dtype=complex))
    gpu.destroy(rho)
    gpu.destroy(q)
