"""Parity of the HIP library (through the QuEST C ABI) with the oracle.

FP64 registers: amplitudes and probabilities within 1e-12 max-abs of QuEST's
CPU backend (golden vectors) and of the oracle on identical seeded inputs;
measurement outcomes identical under the same MT19937 seeds. FP32 registers:
within the FP32 bound stated per test."""
import numpy as np
import pytest

from quantumlab_amd import workloads as W
from quantumlab_amd.quest import QMX_PREC_FP32, QuESTError, Vector, cplx, matrix2
from tests import dense as D
from tests.cases import channel_circuit, ints, misc_ops, pauli_rotation_cases, random_hamil, random_state
from tests.conftest import golden

pytestmark = pytest.mark.gpu
TOL = 1e-12


def run_both(gpu, cpuref, n, fn, density=False, init=None):
    out = []
    for lib in (gpu, cpuref):
        q = lib.qureg(n, density=density)
        if init is not None:
            init(lib, q)
        fn(lib, q)
        out.append(lib.state(q))
        lib.destroy(q)
    return out


# ---- golden vectors of the reference ----------------------------------------------

def test_c1_brick16_golden(gpu):
    g = golden("c1_brick16")
    n = int(g["n"])
    q = gpu.qureg(n)
    W.apply_gates(gpu, q, W.brick_circuit(n, int(g["layers"]), int(g["seed"])))
    assert np.max(np.abs(gpu.state(q) - g["state"])) < TOL
    assert abs(gpu.calcTotalProb(q) - float(g["total"])) < TOL
    p0 = np.array([gpu.calcProbOfOutcome(q, k, 0) for k in range(n)])
    assert np.max(np.abs(p0 - g["prob0"])) < TOL
    gpu.destroy(q)


def test_brick_extras_golden(gpu):
    g = golden("brick10_extras")
    q = gpu.qureg(int(g["n"]))
    W.apply_gates(gpu, q, W.brick_circuit(int(g["n"]), int(g["layers"]), int(g["seed"]), extra_every=5))
    assert np.max(np.abs(gpu.state(q) - g["state"])) < TOL
    gpu.destroy(q)


@pytest.mark.parametrize("density,n", [(False, 5), (True, 3)])
def test_misc_ops_golden(gpu, density, n):
    g = golden(f"misc_{'dm' if density else 'sv'}{n}")
    q = gpu.qureg(n, density=density)
    gpu.initDebugState(q)
    misc_ops(gpu, q, n)
    assert np.max(np.abs(gpu.state(q) - g["state"])) < 1e-11  # debug state norm ~ 850
    gpu.destroy(q)


def test_density_channels_golden(gpu):
    g = golden("density6_channels")
    n = int(g["n"])
    rho, pure = gpu.qureg(n, density=True), gpu.qureg(n)
    channel_circuit(gpu, rho, pure, n, int(g["layers"]), int(g["seed"]))
    assert np.max(np.abs(gpu.state(rho) - g["state"])) < TOL
    assert abs(gpu.calcPurity(rho) - float(g["purity"])) < TOL
    assert abs(gpu.calcTotalProb(rho) - float(g["total"])) < TOL
    assert abs(gpu.calcFidelity(rho, pure) - float(g["fidelity"])) < TOL
    p0 = np.array([gpu.calcProbOfOutcome(rho, k, 0) for k in range(n)])
    assert np.max(np.abs(p0 - g["prob0"])) < TOL
    gpu.destroy(rho)
    gpu.destroy(pure)


@pytest.mark.parametrize("density,n", [(False, 6), (True, 3)])
def test_pauli_rotations_golden(gpu, density, n):
    g = golden(f"pauli_{'dm' if density else 'sv'}{n}")
    q = gpu.qureg(n, density=density)
    gpu.initDebugState(q)
    for targs, paulis, angle in pauli_rotation_cases(n, seed=4):
        gpu.multiRotatePauli(q, targs, paulis, angle)
    assert np.max(np.abs(gpu.state(q) - g["state"])) < 1e-11
    gpu.destroy(q)


def test_trotter_golden(gpu):
    g = golden("trotter8")
    n = int(g["n"])
    h = gpu.hamil(g["coeffs"], g["codes"])
    for order in (1, 2, 4):
        q = gpu.qureg(n)
        gpu.set_state(q, random_state(n, seed=6))
        gpu.applyTrotterCircuit(q, h, float(g["time"]), order, int(g["reps"]))
        assert np.max(np.abs(gpu.state(q) - g[f"order{order}"])) < TOL
        gpu.destroy(q)
    rho = gpu.qureg(4, density=True)
    gpu.initDebugState(rho)
    gpu.applyTrotterCircuit(rho, gpu.hamil(g["coeffs4"], g["codes4"]), 0.5, 2, 2)
    assert np.max(np.abs(gpu.state(rho) - g["density4"])) < 1e-11
    gpu.destroy(rho)


def test_measurement_outcomes_identical_to_reference(gpu):
    g = golden("measure10")
    n = int(g["n"])
    gpu.seed([int(x) for x in g["seeds"]])
    q = gpu.qureg(n)
    gpu.set_state(q, random_state(n, seed=8))
    outs, probs = [], []
    for k in list(range(n)) + list(range(n)):
        o, p = gpu.measureWithStats(q, k)
        outs.append(o)
        probs.append(p)
        if k == n - 1:
            gpu.set_state(q, random_state(n, seed=9))
    assert outs == [int(x) for x in g["outcomes"]]
    assert np.max(np.abs(np.array(probs) - g["probs"])) < TOL
    assert np.max(np.abs(gpu.state(q) - g["state"])) < TOL
    gpu.destroy(q)


# ---- against the oracle on wider registers -------------------------------------------

@pytest.mark.parametrize("n", [1, 2, 3, 7, 14, 20])
def test_brick_circuits_match_oracle(gpu, cpuref, n):
    gates = W.brick_circuit(n, 6, seed=100 + n, extra_every=3 if n >= 5 else 0)
    a, b = run_both(gpu, cpuref, n, lambda lib, q: W.apply_gates(lib, q, gates))
    assert np.max(np.abs(a - b)) < TOL


def test_every_target_and_control_position(gpu, cpuref):
    n = 9
    rng = np.random.default_rng(5)
    ops = []
    for t in range(n):
        a, b = W.compact_pair(rng)
        ops.append(("cu", t, a, b))
        for c in range(n):
            if c != t:
                ops.append(("cnot", c, t))
    ops.append(("mcu", [0, 8, 4], 2, W.haar_unitary(rng, 2)))
    ops.append(("mcu", [1], 0, W.haar_unitary(rng, 2)))
    for k in (1, 2, 3, 5, 6):
        ops.append(("mqu", [int(x) for x in rng.choice(n, size=k, replace=False)], W.haar_unitary(rng, 1 << k)))
    a, b = run_both(gpu, cpuref, n, lambda lib, q: W.apply_gates(lib, q, ops),
                    init=lambda lib, q: lib.set_state(q, random_state(n, 6)))
    assert np.max(np.abs(a - b)) < TOL


def test_reductions_match_oracle(gpu, cpuref):
    n = 13
    psi, phi = random_state(n, 1), random_state(n, 2)
    vals = []
    for lib in (gpu, cpuref):
        a, b = lib.qureg(n), lib.qureg(n)
        lib.set_state(a, psi)
        lib.set_state(b, phi)
        ip = lib.calcInnerProduct(a, b)
        row = [lib.calcTotalProb(a), ip.real, ip.imag, lib.calcFidelity(a, b)]
        row += [lib.calcProbOfOutcome(a, k, o) for k in range(n) for o in (0, 1)]
        row += list(lib.calcProbOfAllOutcomes(a, [3, 0, 7]))
        vals.append(np.array(row))
        lib.destroy(a)
        lib.destroy(b)
    assert np.max(np.abs(vals[0] - vals[1])) < TOL


def test_density_reductions_and_channels_match_oracle(gpu, cpuref):
    n = 5
    vals, states = [], []
    for lib in (gpu, cpuref):
        rho, pure, rho2 = lib.qureg(n, density=True), lib.qureg(n), lib.qureg(n, density=True)
        channel_circuit(lib, rho, pure, n, 3, seed=77)
        lib.call("mixPauli", rho, 2, 0.05, 0.02, 0.03)
        lib.call("mixTwoQubitDephasing", rho, 0, 4, 0.2)
        lib.initPlusState(rho2)
        lib.call("mixDensityMatrix", rho2, 0.3, rho)
        vals.append(np.array([lib.calcPurity(rho), lib.calcTotalProb(rho), lib.calcFidelity(rho, pure),
                              lib.calcDensityInnerProduct(rho, rho2), lib.calcHilbertSchmidtDistance(rho, rho2)]
                             + [lib.calcProbOfOutcome(rho, k, 1) for k in range(n)]
                             + list(lib.calcProbOfAllOutcomes(rho, [4, 1]))))
        states.append(lib.state(rho2))
        for q in (rho, pure, rho2):
            lib.destroy(q)
    assert np.max(np.abs(vals[0] - vals[1])) < TOL
    assert np.max(np.abs(states[0] - states[1])) < TOL


def test_kraus_maps_match_dense(gpu):
    n = 3
    rng = np.random.default_rng(9)
    rho = gpu.qureg(n, density=True)
    psi = random_state(n, 10)
    pq = gpu.qureg(n)
    gpu.set_state(pq, psi)
    gpu.initPureState(rho, pq)
    r = np.outer(psi, psi.conj())
    # a random 2-operator CPTP map: K0 = V[:2], K1 = V[2:] from a 4x2 isometry
    V = W.haar_unitary(rng, 4)[:, :2]
    K = [V[:2, :], V[2:, :]]
    from quantumlab_amd.quest import ComplexMatrix2
    arr = (ComplexMatrix2 * 2)(matrix2(K[0]), matrix2(K[1]))
    gpu.call("mixKrausMap", rho, 1, arr, 2)
    r = D.kraus(r, [D.full_op(n, [1], k) for k in K])
    assert np.max(np.abs(gpu.density(rho) - r)) < TOL
    gpu.call("mixTwoQubitDepolarising", rho, 0, 2, 0.3)
    paulis = [D.I2, D.X, D.Y, D.Z]
    ops = []
    for a in range(4):
        for b in range(4):
            s = np.sqrt(1 - 0.3) if a == b == 0 else np.sqrt(0.3 / 15)
            ops.append(s * D.full_op(n, [0], paulis[a]) @ D.full_op(n, [2], paulis[b]))
    r = D.kraus(r, ops)
    assert np.max(np.abs(gpu.density(rho) - r)) < TOL
    gpu.destroy(rho)
    gpu.destroy(pq)


# ---- FP32 storage with FP64 reductions -------------------------------------------------

def test_fp32_register_tracks_fp64(gpu):
    n = 16
    gates = W.brick_circuit(n, 20, seed=3, extra_every=5)
    q64, q32 = gpu.qureg(n), gpu.qureg(n, precision=QMX_PREC_FP32)
    W.apply_gates(gpu, q64, gates)
    W.apply_gates(gpu, q32, gates)
    assert gpu.qmxQuregPrecision(q32) == QMX_PREC_FP32
    # ~1000 FP32 gates let the norm random-walk by ~1e-6; the FP64 reduction
    # sees it and the FP64 renormalisation removes it
    norm2 = gpu.calcTotalProb(q32)
    assert abs(norm2 - 1) < 1e-5
    assert abs(gpu.qmxRenormalise(q32, 1) - norm2) < 1e-12
    assert abs(gpu.calcTotalProb(q32) - 1) < 1e-7
    maxabs = gpu.qmxMaxAbsDiff(q64, q32)
    ip = gpu.calcInnerProduct(q64, q32)
    fid = ip.real ** 2 + ip.imag ** 2
    assert maxabs < 1e-5 * np.max(np.abs(gpu.state(q64)))
    assert 1 - fid < 1e-8
    gpu.destroy(q64)
    gpu.destroy(q32)


# ---- fusion --------------------------------------------------------------------------------

@pytest.mark.parametrize("k", [2, 3, 5])
def test_fused_blocks_equal_per_gate(gpu, k):
    n = 14
    gates = W.brick_circuit(n, 10, seed=12, extra_every=5)
    c = W.build_circuit(gpu, n, gates)
    nb = gpu.qmxCircuitFuse(c, k)
    assert 0 < nb < len(gates)
    a, b = gpu.qureg(n), gpu.qureg(n)
    gpu.qmxCircuitApply(a, c, 0)
    gpu.qmxCircuitApply(b, c, 1)
    assert gpu.qmxMaxAbsDiff(a, b) < TOL
    W.apply_gates(gpu, b, [])  # no-op
    ref = gpu.qureg(n)
    W.apply_gates(gpu, ref, gates)
    assert gpu.qmxMaxAbsDiff(a, ref) < TOL
    for q in (a, b, ref):
        gpu.destroy(q)
    gpu.qmxCircuitDestroy(c)


def _fidelity_after_renorm(gpu, q64, q32):
    """1 - |<a|b>|^2 / (|a|^2 |b|^2), norms in FP64. The normalised form is used
    because storing the renormalised vector in FP32 rounds its norm^2 again by
    ~1e-9, which is representation noise, not simulation error."""
    gpu.qmxRenormalise(q32, 1)
    ip = gpu.calcInnerProduct(q64, q32)
    f = (ip.real ** 2 + ip.imag ** 2) / (gpu.calcTotalProb(q64) * gpu.calcTotalProb(q32))
    return 1 - f, gpu.qmxMaxAbsDiff(q64, q32)


@pytest.mark.parametrize("n", [9, 12, 20])
def test_fp32_mfma_five_qubit_unitary_matches_fp64(gpu, n):
    """multiQubitUnitary with 5 targets on an FP32 register runs the MFMA kernel;
    targets unsorted and including qubit 0 and the top qubit."""
    rng = np.random.default_rng(n)
    psi = random_state(n, 40 + n)
    q64, q32 = gpu.qureg(n), gpu.qureg(n, precision=QMX_PREC_FP32)
    gpu.set_state(q64, psi)
    gpu.set_state(q32, psi)
    for targs in ([n - 1, 0, 3, 2, 5], [4, 1, 7, 0, 8], [int(x) for x in rng.choice(n, 5, replace=False)]):
        u = W.haar_unitary(rng, 32)
        gpu.multiQubitUnitary(q64, targs, u)
        gpu.multiQubitUnitary(q32, targs, u)
    one_minus_f, maxabs = _fidelity_after_renorm(gpu, q64, q32)
    scale = np.max(np.abs(gpu.state(q64)))
    assert maxabs < 2e-6 * scale
    assert one_minus_f < 1e-9
    gpu.destroy(q64)
    gpu.destroy(q32)


def test_fused_fp32_mfma_circuit_matches_fp64(gpu):
    """Brick circuit fused into 5-qubit MFMA blocks on FP32 amplitudes against
    the per-gate FP64 run: fidelity error <= 1e-6, max-abs
    amplitude error at FP32 resolution."""
    n = 18
    gates = W.brick_circuit(n, 15, seed=13, extra_every=5)
    c = W.build_circuit(gpu, n, gates)
    nb = gpu.qmxCircuitFuse(c, 5)
    assert nb < len(gates) / 4
    q64, q32 = gpu.qureg(n), gpu.qureg(n, precision=QMX_PREC_FP32)
    W.apply_gates(gpu, q64, gates)
    gpu.qmxCircuitApply(q32, c, 1)
    one_minus_f, maxabs = _fidelity_after_renorm(gpu, q64, q32)
    assert one_minus_f < 1e-6
    assert maxabs < 1e-5 * np.max(np.abs(gpu.state(q64)))
    p64 = [gpu.calcProbOfOutcome(q64, k, 0) for k in range(n)]
    p32 = [gpu.calcProbOfOutcome(q32, k, 0) for k in range(n)]
    assert np.max(np.abs(np.array(p64) - np.array(p32))) < 1e-6
    gpu.destroy(q64)
    gpu.destroy(q32)
    gpu.qmxCircuitDestroy(c)


# ---- input validation (QuEST340/tests/*: message substrings) ------------------------------------

@pytest.mark.parametrize("call,substr", [
    (lambda g, q: g.call("hadamard", q, 5), "Invalid target"),
    (lambda g, q: g.call("controlledNot", q, 2, 2), "Control qubit cannot equal target"),
    (lambda g, q: g.call("controlledNot", q, -1, 2), "Invalid control"),
    (lambda g, q: g.call("compactUnitary", q, 0, cplx(1), cplx(1)), "unitary"),
    (lambda g, q: g.multiControlledUnitary(q, [1, 2], 1, np.eye(2)), "Control qubits cannot include target"),
    (lambda g, q: g.multiQubitUnitary(q, [0, 0], np.eye(4)), "target qubits must be unique"),
    (lambda g, q: g.multiQubitUnitary(q, [0, 1], 2 * np.eye(4)), "not unitary"),
    (lambda g, q: g.multiRotatePauli(q, [0, 1], [1, 7], 0.1), "Invalid Pauli code"),
    (lambda g, q: g.call("calcProbOfOutcome", q, 0, 2), "Invalid measurement outcome"),
    (lambda g, q: g.call("mixDephasing", q, 0, 0.1), "valid only for density matrices"),
    (lambda g, q: g.call("collapseToOutcome", q, 0, 1), "zero probability"),
])
def test_validation_messages(gpu, call, substr):
    q = gpu.qureg(5)
    before = gpu.state(q)
    with pytest.raises(QuESTError, match=substr):
        call(gpu, q)
    assert np.array_equal(gpu.state(q), before)  # no state change on invalid input
    gpu.destroy(q)


def test_density_validation_messages(gpu):
    rho = gpu.qureg(3, density=True)
    with pytest.raises(QuESTError, match="cannot exceed 1/2"):
        gpu.call("mixDephasing", rho, 0, 0.6)
    with pytest.raises(QuESTError, match="cannot exceed 3/4"):
        gpu.call("mixDepolarising", rho, 0, 0.8)
    h = gpu.hamil([1.0], np.array([[1, 0, 0]]))
    with pytest.raises(QuESTError, match="order must be 1, or an even number"):
        gpu.call("applyTrotterCircuit", rho, h, 1.0, 3, 1)
    with pytest.raises(QuESTError, match="repetitions must be >=1"):
        gpu.call("applyTrotterCircuit", rho, h, 1.0, 2, 0)
    with pytest.raises(QuESTError, match="Invalid number of qubits"):
        gpu.qureg(0)
    gpu.destroy(rho)


def test_mismatched_register_messages(gpu):
    """QuEST340/tests/test_calculations.cpp:93,270-285,687-692,859-868,1298,1410
    and test_data_structures.cpp:158-165,638-644: register pairs of different
    sizes or kinds, Hamiltonians of the wrong width, and registers too large
    for a size_t amplitude count are refused with QuEST's messages."""
    v3, v4 = gpu.qureg(3), gpu.qureg(4)
    m3, m4 = gpu.qureg(3, density=True), gpu.qureg(4, density=True)
    h3 = gpu.hamil([0.5], np.array([[1, 0, 3]]))
    h4 = gpu.hamil([0.5], np.array([[1, 0, 3, 2]]))
    w3, wm3 = gpu.qureg(3), gpu.qureg(3, density=True)
    cases = [
        (lambda: gpu.calcInnerProduct(v3, v4), "Dimensions.*don't match"),
        (lambda: gpu.calcInnerProduct(v3, m3), "valid only for state-vectors"),
        (lambda: gpu.calcInnerProduct(m3, v3), "valid only for state-vectors"),
        (lambda: gpu.calcInnerProduct(m3, m3), "valid only for state-vectors"),
        (lambda: gpu.calcFidelity(v4, v3), "Dimensions.*don't match"),
        (lambda: gpu.calcFidelity(m4, v3), "Dimensions.*don't match"),
        (lambda: gpu.calcDensityInnerProduct(m3, m4), "Dimensions.*don't match"),
        (lambda: gpu.calcHilbertSchmidtDistance(m3, m4), "Dimensions.*don't match"),
        (lambda: gpu.calcExpecPauliHamil(v3, h3, v4), "Dimensions.*don't match"),
        (lambda: gpu.calcExpecPauliHamil(m3, h3, m4), "Dimensions.*don't match"),
        (lambda: gpu.calcExpecPauliHamil(v3, h4, w3), "same number of qubits"),
        (lambda: gpu.calcExpecPauliHamil(m3, h4, wm3), "same number of qubits"),
        (lambda: gpu.getAmp(m3, 0), "valid only for state-vectors"),
        (lambda: gpu.getNumAmps(m3), "valid only for state-vectors"),
        (lambda: gpu.qureg(64), "Too many qubits.*size_t type"),
        (lambda: gpu.qureg(32, density=True), "Too many qubits.*size_t type"),
    ]
    before = [gpu.state(q) for q in (v3, v4, m3, m4)]
    for fn, substr in cases:
        with pytest.raises(QuESTError, match=substr):
            fn()
    for q, b in zip((v3, v4, m3, m4), before):
        assert np.array_equal(gpu.state(q), b)
    for h in (h3, h4):
        gpu.lib.destroyPauliHamil(h)
    for q in (v3, v4, m3, m4, w3, wm3):
        gpu.destroy(q)


# ---- size-independent properties at larger sizes -----------------------------------------------

def inverse(gates):
    inv = []
    for g in reversed(gates):
        if g[0] == "cu":
            inv.append(("cu", g[1], np.conj(g[2]), -g[3]))
        elif g[0] == "cnot":
            inv.append(g)
        elif g[0] == "rz":
            inv.append(("rz", g[1], -g[2]))
        elif g[0] == "mcu":
            inv.append(("mcu", g[1], g[2], g[3].conj().T))
        elif g[0] == "mqu":
            inv.append(("mqu", g[1], g[2].conj().T))
    return inv


@pytest.mark.parametrize("n", [24, 27])
def test_large_circuit_round_trip_returns_to_zero_state(gpu, n):
    gates = W.brick_circuit(n, 4, seed=n, extra_every=2)
    q = gpu.qureg(n)
    W.apply_gates(gpu, q, gates)
    assert abs(gpu.calcTotalProb(q) - 1) < 1e-12
    W.apply_gates(gpu, q, inverse(gates))
    a0 = gpu.getAmp(q, 0)
    assert abs(a0.real - 1) < 1e-12 and abs(a0.imag) < 1e-12
    assert abs(gpu.calcProbOfOutcome(q, n - 1, 0) - 1) < 1e-12
    gpu.destroy(q)


def test_low_target_stream_forms(gpu, cpuref):
    """2x2 gates and NOTs on the lowest qubits run as in-place streams: partners
    by cross-lane shuffle (targets 0-5) or through LDS (targets 6-10), with
    controls where a control is among the lowest bits (amplitudes whose controls
    differ written back unchanged), including controls on |0>, swaps of two
    low qubits, Pauli-string rotations and phase gates on the lowest qubits.
    Every target on a 12-qubit register, against the oracle."""
    n = 12
    rng = np.random.default_rng(77)

    def ops(lib, q):
        for t in range(n):
            a, b = pairs[t]
            lib.call("compactUnitary", q, t, cplx(a), cplx(b))
            for c in (0, 1, 2, 7, 11):
                if c != t:
                    lib.call("controlledCompactUnitary", q, c, t, cplx(b), cplx(a))
                    lib.call("controlledNot", q, c, t)
            lib.call("pauliX", q, t)
            lib.call("multiStateControlledUnitary", q, ints(ctrls[t]), ints([0, 1]), 2, t, matrix2(us[t]))
        for q1, q2 in ((0, 1), (2, 5), (1, 4), (0, 9), (3, 11)):
            lib.call("swapGate", q, q1, q2)
        for targs, paulis, ang in (([0, 1], [1, 2], 0.3), ([2, 5, 9], [2, 3, 1], -0.7), ([4], [1], 1.1),
                                   ([0, 3, 7, 11], [3, 2, 2, 1], 0.5)):
            lib.multiRotatePauli(q, targs, paulis, ang)
        lib.call("phaseShift", q, 0, 0.4)
        lib.call("sGate", q, 1)
        lib.call("tGate", q, 2)
        lib.call("controlledPhaseShift", q, 0, 2, -0.8)
        lib.call("controlledPhaseFlip", q, 1, 2)
        lib.call("controlledPhaseShift", q, 1, 9, 0.6)
        lib.call("multiControlledPhaseFlip", q, ints([0, 1, 2]), 3)
        lib.call("multiQubitNot", q, ints([0, 3, 5]), 3)
        lib.call("multiControlledMultiQubitNot", q, ints([1, 9]), 2, ints([0, 4]), 2)

    pairs = [W.compact_pair(rng) for _ in range(n)]
    us = [W.haar_unitary(rng, 2) for _ in range(n)]
    ctrls = [[c for c in (0, 1, 2, 3) if c != t][:2] for t in range(n)]
    a, b = run_both(gpu, cpuref, n, ops, init=lambda lib, q: lib.set_state(q, random_state(n, 9)))
    assert np.max(np.abs(a - b)) < TOL


def test_controlled_pauli_rotations_match_dense(gpu):
    """multiControlledMultiRotatePauli with low controls (in-place stream form)
    and high ones (pair form) against the dense operator exp(-i theta/2 P) on
    the controlled subspace, 10 qubits."""
    n = 10
    psi = random_state(n, 12)
    q = gpu.qureg(n)
    gpu.set_state(q, psi)
    want = psi.copy()
    for cs, targs, paulis, ang in (([0], [2, 4], [1, 2], 0.9), ([9], [1, 3], [2, 1], -0.4),
                                   ([1, 6], [0, 5], [1, 3], 0.2), ([2], [0], [2], 1.3), ([5, 0], [3, 8], [1, 1], 0.7)):
        gpu.call("multiControlledMultiRotatePauli", q, ints(cs), len(cs), ints(targs), ints(paulis), len(targs), ang)
        P = np.ones((1, 1), dtype=complex)
        for p in paulis:                      # targets[0] is the least significant factor
            P = np.kron(D.PAULI[p], P)
        U = np.cos(ang / 2) * np.eye(len(P)) - 1j * np.sin(ang / 2) * P
        want = D.full_op(n, targs, U, controls=cs) @ want
    assert np.max(np.abs(gpu.state(q) - want)) < TOL
    gpu.destroy(q)


# ---- regressions ---------------------------------------------------------------------------

def _zz_hamil(n, terms, seed, with_x):
    rng = np.random.default_rng(seed)
    codes, coeffs = [], []
    for _ in range(terms):
        a, b = rng.choice(n, size=2, replace=False)
        row = np.zeros(n, dtype=np.int32)
        row[a] = row[b] = 3
        codes.append(row)
        coeffs.append(float(rng.uniform(0.9, 1.1)))  # rotation angle 2c: |cos(c)| >= 0.45
    if with_x:
        for k in range(n):
            row = np.zeros(n, dtype=np.int32)
            row[k] = 1
            codes.append(row)
            coeffs.append(0.3)
    return np.array(coeffs), np.stack(codes)


@pytest.mark.parametrize("density,n", [(False, 12), (False, 13), (True, 6)])
@pytest.mark.parametrize("with_x", [False, True])
def test_trotter_many_large_angle_rotations_fp32(gpu, cpuref, density, n, with_x):
    """~100 ZZ terms at time 1 in one block: every rotation has |cos(theta/2)|
    >= 1/2, so the tangent form would scale the tile by their product (2^-100
    and below, past FP32's range). FP32 registers, basis / product start, against
    the oracle."""
    coeffs, codes = _zz_hamil(n, 100, seed=n, with_x=with_x)
    outs = []
    for lib, prec in ((gpu, QMX_PREC_FP32), (cpuref, None)):
        q = lib.qureg(n, density=density, precision=prec)
        if with_x:
            lib.initPlusState(q)
        else:
            lib.initClassicalState(q, 5)
        h = lib.hamil(coeffs, codes)
        lib.call("applyTrotterCircuit", q, h, 1.0, 1, 1)
        lib.lib.destroyPauliHamil(h)
        outs.append(lib.state(q))
        lib.destroy(q)
    assert np.all(np.isfinite(outs[0]))
    assert np.max(np.abs(outs[0] - outs[1])) < 2e-5


@pytest.mark.parametrize("density,n", [(False, 3), (False, 7), (False, 13), (True, 3), (True, 6)])
def test_prob_of_all_outcomes_matches_oracle(gpu, cpuref, density, n):
    """Subsets of every size and order, low and high qubits, one pass each."""
    rng = np.random.default_rng(n + 10 * density)
    subsets = [[0], [n - 1], list(range(n)), list(reversed(range(n)))]
    subsets += [[int(x) for x in rng.choice(n, size=int(rng.integers(1, n + 1)), replace=False)] for _ in range(6)]
    regs = []
    for lib in (gpu, cpuref):
        q = lib.qureg(n, density=density)
        if density:
            pure = lib.qureg(n)
            channel_circuit(lib, q, pure, n, 2, seed=3)
            lib.destroy(pure)
        else:
            lib.initDebugState(q)
            misc_ops(lib, q, n)
        regs.append(q)
    for qs in subsets:
        a = gpu.calcProbOfAllOutcomes(regs[0], qs)
        b = cpuref.calcProbOfAllOutcomes(regs[1], qs)
        scale = max(1.0, float(np.max(np.abs(b))))
        assert np.max(np.abs(a - b)) < 1e-12 * scale, qs
    for lib, q in zip((gpu, cpuref), regs):
        lib.destroy(q)


@pytest.mark.parametrize("n,targets", [(11, [0, 1, 2, 3, 4]), (11, [6, 7, 8, 9, 10]), (12, [0, 3, 7, 10, 11]),
                                       (13, [4, 0, 12, 2, 1]), (14, [5, 6, 7, 8, 9]), (14, [13, 0, 9, 1, 12]),
                                       (15, [0, 1, 2, 5, 9]), (15, [3, 1, 14, 6, 2]), (12, [1, 2, 3, 4, 5]),
                                       (10, [9, 0, 4, 2, 7])])
def test_fp64_five_target_blocks_every_placement(gpu, cpuref, n, targets):
    """Uncontrolled 5-target multiQubitUnitary on FP64 registers (the FP64
    MFMA tiles from 11 qubits, kDense below) against the oracle's
    multiQubitUnitary on the same state, for low, high, mid, scattered and
    unsorted targets (low group bits in every order of the LDS swizzle)."""
    rng = np.random.default_rng(100 + n + sum(targets))
    psi = random_state(n, 5 + n)
    regs = []
    for lib in (gpu, cpuref):
        q = lib.qureg(n)
        lib.set_state(q, psi)
        regs.append(q)
    for _ in range(3):
        u = W.haar_unitary(rng, 32)
        for lib, q in zip((gpu, cpuref), regs):
            lib.multiQubitUnitary(q, targets, u)
    got, want = gpu.state(regs[0]), cpuref.state(regs[1])
    assert np.max(np.abs(got - want)) < 1e-12
    gpu.destroy(regs[0])
    cpuref.destroy(regs[1])


@pytest.mark.parametrize("n,targets", [(6, [0, 1, 2, 3, 4]), (6, [5, 1, 3, 0, 2]), (7, [6, 0, 3, 5, 1])])
def test_five_target_blocks_on_density_registers(gpu, cpuref, n, targets):
    """multiQubitUnitary with 5 targets on density registers of 6-7 qubits
    (12-14-qubit vectors: the tile kernels on the rows, then on the columns
    with the conjugate matrix, QuEST.c:302-317): FP64 against the oracle at
    1e-12; an FP32-storage register against the FP64 one within FP32
    resolution."""
    rng = np.random.default_rng(300 + n + sum(targets))
    regs = []
    for lib in (gpu, cpuref):
        rho, pure = lib.qureg(n, density=True), lib.qureg(n)
        channel_circuit(lib, rho, pure, n, 2, seed=n)
        lib.destroy(pure)
        regs.append(rho)
    us = [W.haar_unitary(rng, 32) for _ in range(2)]
    for lib, rho in zip((gpu, cpuref), regs):
        for u in us:
            lib.multiQubitUnitary(rho, targets, u)
    got, want = gpu.state(regs[0]), cpuref.state(regs[1])
    assert np.max(np.abs(got - want)) < 1e-12
    r32 = gpu.qureg(n, density=True, precision=QMX_PREC_FP32)
    pure = gpu.qureg(n)
    channel_circuit(gpu, r32, pure, n, 2, seed=n)
    gpu.destroy(pure)
    for u in us:
        gpu.multiQubitUnitary(r32, targets, u)
    assert np.max(np.abs(gpu.state(r32) - got)) < 1e-5
    gpu.destroy(r32)
    gpu.destroy(regs[0])
    cpuref.destroy(regs[1])


@pytest.mark.parametrize("n,targets", [(11, [0, 1, 2, 3, 4]), (11, [6, 7, 8, 9, 10]), (12, [0, 3, 7, 10, 11]),
                                       (13, [4, 0, 12, 2, 1]), (14, [5, 6, 7, 8, 9]), (14, [13, 0, 9, 1, 12]),
                                       (10, [9, 0, 4, 2, 7]), (9, [0, 1, 2, 3, 4])])
def test_fp32_five_target_blocks_every_placement(gpu, n, targets):
    """Uncontrolled 5-target multiQubitUnitary on FP32 registers (the fused
    MFMA kernels: LDS tiles from 11 qubits, the direct form below) against the
    FP64 dense product, for low, high, mid, scattered and unsorted targets."""
    rng = np.random.default_rng(n + sum(targets))
    psi = random_state(n, 3 + n)
    q = gpu.qureg(n, precision=QMX_PREC_FP32)
    gpu.set_state(q, psi)
    want = psi.astype(complex)
    idx = np.arange(1 << n)
    for _ in range(3):
        u = W.haar_unitary(rng, 32)
        gpu.multiQubitUnitary(q, targets, u)
        mask = sum(1 << t for t in targets)
        rest = idx[(idx & mask) == 0]
        offs = np.array([sum(((c >> j) & 1) << targets[j] for j in range(5)) for c in range(32)])
        X = want[rest[:, None] + offs[None, :]]
        want[rest[:, None] + offs[None, :]] = X @ u.T
    assert np.max(np.abs(gpu.state(q) - want)) < 2e-6
    gpu.destroy(q)
