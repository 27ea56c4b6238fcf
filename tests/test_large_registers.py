"""Registers whose one-item-per-amplitude launches pass 2^32 work-items: a
dispatch counts work-items in 32 bits per dimension, so the library splits such
index spaces into rows of a 2-D grid. 33 qubits in FP32 and 32 qubits in FP64
(64 GiB each). Every operation acts on qubit 0 and the top qubit only, so the
state stays on four amplitudes, and a two-qubit numpy model of the same
operations (QuEST's gate definitions) gives them exactly."""
import numpy as np
import pytest

from quantumlab_amd.quest import QMX_PREC_FP32, qreal, cplx
from tests.cases import ints

pytestmark = pytest.mark.gpu


def reals(xs):
    return (qreal * len(xs))(*xs)


def rz(theta):
    return np.diag([np.exp(-0.5j * theta), np.exp(0.5j * theta)])


def apply1(psi, u, k):
    """u on model qubit k (0: register qubit 0, 1: the top qubit); psi[b1, b0]"""
    return np.einsum("ab,bc->ac", u, psi) if k == 1 else np.einsum("ab,cb->ca", u, psi)


@pytest.mark.parametrize("n,prec,tol", [(33, QMX_PREC_FP32, 2e-6), (32, None, 1e-12)])
def test_operations_on_registers_past_two_to_the_32_items(gpu, n, prec, tol):
    top = n - 1
    th1, th2, phi, c = 0.7, -1.3, 0.45, 0.9
    alpha, beta = np.cos(0.4) * np.exp(0.3j), np.sin(0.4) * np.exp(-1.1j)
    q = gpu.qureg(n, precision=prec)
    h = np.array([[1, 1], [1, -1]]) / np.sqrt(2)
    psi = np.zeros((2, 2), complex)
    psi[0, 0] = 1.0
    gpu.call("hadamard", q, 0)
    psi = apply1(psi, h, 0)
    gpu.call("hadamard", q, top)
    psi = apply1(psi, h, 1)
    gpu.call("rotateZ", q, top, th1)
    psi = apply1(psi, rz(th1), 1)
    gpu.call("multiRotateZ", q, ints([0, top]), 2, th2)          # every amplitude: 2^n items
    par = np.array([[0, 1], [1, 0]])                            # parity of (b1, b0)
    psi = psi * np.exp(-0.5j * th2 * (1 - 2 * par))
    gpu.call("pauliX", q, top)                                   # 2^(n-1) pairs
    psi = apply1(psi, np.array([[0, 1], [1, 0]]), 1)
    gpu.call("phaseShift", q, 0, phi)                            # 2^(n-1) amplitudes
    psi = apply1(psi, np.diag([1, np.exp(1j * phi)]), 0)
    gpu.call("applyPhaseFunc", q, ints([top]), 1, 0, reals([c]), reals([1.0]), 1)   # 2^n items
    psi = apply1(psi, np.diag([1, np.exp(1j * c)]), 1)
    gpu.call("compactUnitary", q, top, cplx(alpha), cplx(beta))
    psi = apply1(psi, np.array([[alpha, -np.conj(beta)], [beta, np.conj(alpha)]]), 1)
    for b1 in (0, 1):
        for b0 in (0, 1):
            a = gpu.getAmp(q, b0 + (b1 << top))
            assert abs(complex(a.real, a.imag) - psi[b1, b0]) < tol, (n, b1, b0)
    assert abs(gpu.calcTotalProb(q) - 1) < tol
    assert abs(gpu.calcProbOfOutcome(q, top, 0) - np.sum(np.abs(psi[0]) ** 2)) < tol
    a = gpu.getAmp(q, 2)                                         # off the support
    assert a.real == 0.0 and a.imag == 0.0
    gpu.destroy(q)
