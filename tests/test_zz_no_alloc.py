"""The allocation-free outcome histogram (qmxCalcProbOfAllOutcomesNoAlloc)
against calcProbOfAllOutcomes, through the C ABI: two calls in a row on the
same device scratch, the second smaller, state vectors and density matrices,
both precision flavours; oversized requests rejected. Results must be equal
(same kernels, deterministic order) and within 1e-12 of the oracle."""
import ctypes as C

import numpy as np
import pytest

from tests.cases import random_state

pytestmark = pytest.mark.gpu


def _outcomes(qe, q, qubits, no_alloc):
    arr = (C.c_int * len(qubits))(*qubits)
    out = (qe.t.qreal * (1 << len(qubits)))()
    qe.call("qmxCalcProbOfAllOutcomesNoAlloc" if no_alloc else "calcProbOfAllOutcomes", out, q, arr, len(qubits))
    return np.array(out[:], dtype=np.float64)


@pytest.mark.parametrize("flavour", ["f64", "f32"])
def test_no_alloc_histogram_matches_calc_prob_of_all_outcomes(flavour, cpuref):
    import quantumlab_amd
    qe = quantumlab_amd.load(flavour=flavour)
    cases = [(20, False, [19, 3, 0, 12, 7, 15, 1, 8, 11, 2, 5, 17]),   # large first
             (9, False, [4, 0, 8]),                                     # then smaller, same scratch
             (5, True, [1, 4, 0]),
             (14, False, list(range(14))[::-1])]
    for n, density, qubits in cases:
        q = qe.qureg(n, density=density)
        ref = cpuref.qureg(n, density=density)
        psi = random_state(n, n)
        if density:
            pure, pref = qe.qureg(n), cpuref.qureg(n)
            qe.set_state(pure, psi)
            cpuref.set_state(pref, psi)
            qe.initPureState(q, pure)
            cpuref.initPureState(ref, pref)
            qe.destroy(pure)
            cpuref.destroy(pref)
        else:
            qe.set_state(q, psi)
            cpuref.set_state(ref, psi)
        a = _outcomes(qe, q, qubits, no_alloc=True)
        b = _outcomes(qe, q, qubits, no_alloc=False)
        assert np.array_equal(a, b), (n, density)
        want = cpuref.calcProbOfAllOutcomes(ref, qubits)
        tol = 1e-12 if flavour == "f64" else 1e-6
        assert np.max(np.abs(a - want)) < tol, (n, density)
        qe.destroy(q)
        cpuref.destroy(ref)


def test_no_alloc_histogram_rejects_oversized_requests():
    import quantumlab_amd
    from quantumlab_amd.quest import QuESTError
    qe = quantumlab_amd.load()
    q = qe.qureg(22)
    qubits = list(range(21))
    arr = (C.c_int * 21)(*qubits)
    out = (qe.t.qreal * (1 << 21))()
    with pytest.raises(QuESTError, match="allocation-free histogram"):
        qe.call("qmxCalcProbOfAllOutcomesNoAlloc", out, q, arr, 21)
    # a request that fits still works after the rejection
    assert abs(sum(_outcomes(qe, q, [0, 21], no_alloc=True)) - 1.0) < 1e-12
    qe.destroy(q)
