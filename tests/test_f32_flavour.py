"""The QuEST_PREC=1 flavour (libquest_mi355x_f32.so, qreal = float: QuEST's and
QuantumLab's default precision) against golden fixtures written by QuEST's own
CPU backend built at QuEST_PREC=1 (tests/golden/f32_*.npz, make_golden.py --f32).

Tolerance: QuEST's own for this precision, REAL_EPS = 1e-5 absolute
(QuEST_precision.h:35), relative to the largest entry for the unnormalised
debug state; measurement outcomes identical under the same seeds."""
import ctypes as C

import numpy as np
import pytest

from quantumlab_amd import workloads as W
from quantumlab_amd.quest import QMX_PREC_FP32, QMX_PREC_FP64
from tests.cases import channel_circuit, misc_ops, random_state
from tests.conftest import golden

pytestmark = pytest.mark.gpu
EPS = 1e-5


@pytest.fixture(scope="module")
def gpu32():
    import quantumlab_amd
    return quantumlab_amd.load(flavour="f32")


def test_float_registers_default_to_fp32_storage(gpu32):
    q = gpu32.qureg(4)
    assert gpu32.qmxQuregPrecision(q) == QMX_PREC_FP32
    assert C.sizeof(gpu32.t.qreal) == 4
    gpu32.destroy(q)
    q = gpu32.qureg(4, precision=QMX_PREC_FP64)
    assert gpu32.qmxQuregPrecision(q) == QMX_PREC_FP64
    gpu32.destroy(q)


def test_brick_extras_golden(gpu32):
    g = golden("f32_brick12_extras")
    n = int(g["n"])
    q = gpu32.qureg(n)
    W.apply_gates(gpu32, q, W.brick_circuit(n, int(g["layers"]), int(g["seed"]), extra_every=5))
    assert np.max(np.abs(gpu32.state(q) - g["state"])) < EPS
    assert abs(gpu32.calcTotalProb(q) - float(g["total"])) < EPS
    p0 = np.array([gpu32.calcProbOfOutcome(q, k, 0) for k in range(n)])
    assert np.max(np.abs(p0 - g["prob0"])) < EPS
    gpu32.destroy(q)


@pytest.mark.parametrize("density,n", [(False, 5), (True, 3)])
def test_misc_ops_golden(gpu32, density, n):
    g = golden(f"f32_misc_{'dm' if density else 'sv'}{n}")
    q = gpu32.qureg(n, density=density)
    gpu32.initDebugState(q)
    misc_ops(gpu32, q, n)
    scale = max(1.0, float(np.max(np.abs(g["state"]))))
    assert np.max(np.abs(gpu32.state(q) - g["state"])) < EPS * scale
    gpu32.destroy(q)


def test_density_channels_golden(gpu32):
    g = golden("f32_density5_channels")
    n = int(g["n"])
    rho, pure = gpu32.qureg(n, density=True), gpu32.qureg(n)
    channel_circuit(gpu32, rho, pure, n, int(g["layers"]), int(g["seed"]))
    assert np.max(np.abs(gpu32.state(rho) - g["state"])) < EPS
    assert abs(gpu32.calcPurity(rho) - float(g["purity"])) < EPS
    assert abs(gpu32.calcTotalProb(rho) - float(g["total"])) < EPS
    assert abs(gpu32.calcFidelity(rho, pure) - float(g["fidelity"])) < EPS
    p0 = np.array([gpu32.calcProbOfOutcome(rho, k, 0) for k in range(n)])
    assert np.max(np.abs(p0 - g["prob0"])) < EPS
    gpu32.destroy(rho)
    gpu32.destroy(pure)


@pytest.mark.parametrize("order", [1, 2, 4])
def test_trotter_golden(gpu32, order):
    g = golden("f32_trotter8")
    n = int(g["n"])
    h = gpu32.hamil(g["coeffs"], g["codes"])
    q = gpu32.qureg(n)
    gpu32.set_state(q, random_state(n, seed=26))
    gpu32.applyTrotterCircuit(q, h, float(g["time"]), order, int(g["reps"]))
    # QuEST's float build applies each rotation as 1 + 2w float sweeps (1800 at
    # order 4 here); its own drift from FP64 reaches 1.2e-5, so the bound is 3e-5
    assert np.max(np.abs(gpu32.state(q) - g[f"order{order}"])) < 3 * EPS
    gpu32.lib.destroyPauliHamil(h)
    gpu32.destroy(q)


def test_measurement_outcomes_golden(gpu32):
    g = golden("f32_measure10")
    n = int(g["n"])
    gpu32.seed([int(x) for x in g["seeds"]])
    q = gpu32.qureg(n)
    gpu32.set_state(q, random_state(n, seed=27))
    outcomes, probs = [], []
    for k in list(range(n)) + list(range(n)):
        o, p = gpu32.measureWithStats(q, k)
        outcomes.append(o)
        probs.append(p)
        if k == n - 1:
            gpu32.set_state(q, random_state(n, seed=28))
    assert outcomes == [int(x) for x in g["outcomes"]]
    assert np.max(np.abs(np.array(probs) - g["probs"])) < EPS
    assert np.max(np.abs(gpu32.state(q) - g["state"])) < EPS
    gpu32.destroy(q)


def test_float_host_mirror_round_trip(gpu32):
    """QuantumLab's upload protocol with a float mirror: exact."""
    n = 9
    psi = random_state(n, 29).astype(np.complex64)
    q = gpu32.qureg(n)
    N = q.numAmpsTotal
    re = np.ctypeslib.as_array(q.stateVec.real, shape=(N,))
    im = np.ctypeslib.as_array(q.stateVec.imag, shape=(N,))
    assert re.dtype == np.float32
    re[:] = psi.real
    im[:] = psi.imag
    gpu32.copyStateToGPU(q)
    re[:] = 0
    im[:] = 0
    gpu32.copyStateFromGPU(q)
    assert np.array_equal(re + 1j * im, psi)
    gpu32.destroy(q)
