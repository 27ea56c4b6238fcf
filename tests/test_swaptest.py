"""Amplitude encoding + swap tests (quantum k-means): oracle pinning on the CPU,
parity of the batched GPU plan on the GPU."""
import numpy as np
import pytest

from oracle import swaptest_oracle as O
from quantumlab_amd import qkmeans as K
from quantumlab_amd.quest import QMX_PREC_FP32, QMX_PREC_FP64


# ---- oracle pinning (CPU) ---------------------------------------------------------

@pytest.mark.parametrize("q", [1, 2, 3, 4, 5])
def test_gate_level_encoder_prepares_normalised_vector(q):
    """The reference's Ry + Gray-code FRy ladder on |0..0> gives v/|v| with
    qubit l <-> bit l, which is also the tree-product formula the kernel uses."""
    rng = np.random.default_rng(q)
    v = rng.standard_normal(1 << q)
    circ = O.ae_circuit_dense(v)
    tree = O.ae_state_tree(v)
    assert np.max(np.abs(circ - v / np.linalg.norm(v))) < 1e-13
    assert np.max(np.abs(tree - v / np.linalg.norm(v))) < 1e-13


def test_encoder_handles_zero_blocks():
    v = np.array([0.0, 0.0, 0.3, -0.4, 0.0, 0.0, 0.0, 0.5])
    assert np.max(np.abs(O.ae_circuit_dense(v) - v / np.linalg.norm(v))) < 1e-13
    assert np.max(np.abs(O.ae_state_tree(v) - v / np.linalg.norm(v))) < 1e-13


def test_toffoli_decomposition_is_a_toffoli():
    """The H/CX/RZ(pi/4) list equals the basis permutation of CCX up to one
    global phase, for every control/target placement on 4 qubits."""
    rng = np.random.default_rng(0)
    for c0, c1, t in [(0, 1, 2), (0, 2, 3), (3, 1, 0), (2, 0, 1)]:
        psi = rng.standard_normal(16) + 1j * rng.standard_normal(16)
        psi /= np.linalg.norm(psi)
        state = O.run_gates_dense(O.toffoli_gates(c0, c1, t), 4, psi.copy())
        want = psi[O.ccx_dense(4, c0, c1, t)]
        phase = np.vdot(want, state)
        assert abs(abs(phase) - 1) < 1e-13
        assert np.max(np.abs(state - phase * want)) < 1e-13


@pytest.mark.parametrize("q", [1, 2, 3])
def test_gate_level_swap_test_matches_closed_form(q):
    rng = np.random.default_rng(10 + q)
    a, b = rng.standard_normal(1 << q), rng.standard_normal(1 << q)
    p0 = O.swap_test_circuit_dense(a, b)
    ov = (a @ b) / (np.linalg.norm(a) * np.linalg.norm(b))
    assert abs(p0 - 0.5 * (1 + ov ** 2)) < 1e-13
    assert abs(O.swap_test_circuit_dense(a, a) - 1.0) < 1e-13


@pytest.mark.parametrize("lib", ["reflib", "cpuref"])
def test_swap_test_gate_list_through_quest_cpu(lib, request):
    """The same gate list through QuEST's C API on the CPU (QuEST 3.4's own
    backend, and the C restatement) gives the dense state and the closed-form
    P(ancilla = 0)."""
    qe = request.getfixturevalue(lib)
    rng = np.random.default_rng(21)
    for q in (1, 2, 3):
        a, b = rng.standard_normal(1 << q), rng.standard_normal(1 << q)
        gates = O.swap_test_gates(a, b)
        reg = qe.qureg(2 * q + 1)
        O.run_gates_quest(qe, reg, gates)
        ov = (a @ b) / (np.linalg.norm(a) * np.linalg.norm(b))
        assert abs(qe.lib.calcProbOfOutcome(reg, 0, 0) - 0.5 * (1 + ov ** 2)) < 1e-12
        assert np.max(np.abs(qe.state(reg) - O.run_gates_dense(gates, 2 * q + 1))) < 1e-12
        qe.destroy(reg)


def test_kmeans_oracle_recovers_clusters():
    data = K.make_dataset(num_points=512, num_clusters=4, q_lo=3, q_hi=6, seed=3)
    pts = O.normalised_points(data)
    cents = data.init.copy()
    for _ in range(5):
        _, assign, cents = O.kmeans_step(pts, cents)
    # each generating cluster maps to one k-means cluster for almost all points
    agree = sum(np.max(np.bincount(assign[data.labels == c], minlength=4)) for c in range(4))
    assert agree / len(assign) > 0.95


# ---- GPU plan ------------------------------------------------------------------------

def _small(seed=5, P=1024, C=8, q_lo=2, q_hi=8):
    return K.make_dataset(num_points=P, num_clusters=C, q_lo=q_lo, q_hi=q_hi, seed=seed)


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [(QMX_PREC_FP64, 1e-12), (QMX_PREC_FP32, 1e-6)])
def test_swap_test_p0_matches_exact_overlaps(gpu, prec, tol):
    data = _small()
    plan = K.SwapPlan(gpu, data, prec)
    plan.run(1, graph=False)
    p0 = plan.p0()
    exact = O.p0_matrix(O.normalised_points(data), data.init)
    # distance reported by the swap test: <a|b>^2 = 2 P0 - 1
    assert np.max(np.abs((2 * p0 - 1) - (2 * exact - 1))) < tol
    assert np.array_equal(plan.assignments(), np.argmax(exact, axis=1))
    plan.close()


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [(QMX_PREC_FP64, 1e-13), (QMX_PREC_FP32, 1e-6)])
def test_encoded_registers_are_normalised_points(gpu, prec, tol):
    data = _small(P=64)
    plan = K.SwapPlan(gpu, data, prec)
    for p in range(0, 64, 7):
        v = data.point(p)
        assert np.max(np.abs(plan.point_state(p) - v / np.linalg.norm(v))) < tol
    plan.close()


@pytest.mark.gpu
def test_kmeans_iterations_match_fp64_oracle_and_fp32_assignments_identical(gpu):
    data = _small(seed=9, P=2048, C=8, q_lo=4, q_hi=9)
    pts = O.normalised_points(data)
    cents = data.init.copy()
    p64 = K.SwapPlan(gpu, data, QMX_PREC_FP64)
    p32 = K.SwapPlan(gpu, data, QMX_PREC_FP32)
    for it in range(6):
        p0_ref, assign_ref, cents = O.kmeans_step(pts, cents)
        p64.run(1, graph=False)
        p32.run(1, graph=False)
        assert np.array_equal(p64.assignments(), assign_ref), it
        assert np.array_equal(p32.assignments(), assign_ref), it
        assert np.max(np.abs(p64.p0() - p0_ref)) < 1e-12
        assert np.max(np.abs(2 * p32.p0() - 2 * p0_ref)) < 1e-6
        assert np.max(np.abs(p64.centroids() - cents)) < 1e-12
    p64.close()
    p32.close()


@pytest.mark.gpu
def test_graph_replays_match_eager_iterations(gpu):
    """Every replay of the captured iteration equals the eager iteration, with
    eager work and synchronisations between replays."""
    data = _small(seed=11, P=768, C=8)
    eager = K.SwapPlan(gpu, data, QMX_PREC_FP32)
    graph = K.SwapPlan(gpu, data, QMX_PREC_FP32)
    for it in range(4):
        eager.run(1, graph=False)
        graph.run(1, graph=True)
        gpu.qmxSync()
        assert np.array_equal(eager.p0(), graph.p0()), it
        assert np.array_equal(eager.assignments(), graph.assignments()), it
        assert np.array_equal(eager.centroids(), graph.centroids()), it
    eager.close()
    graph.close()


@pytest.mark.gpu
def test_set_centroids_restarts_the_iteration(gpu):
    """Resetting the centroids after graph replays gives the same iterations as
    a fresh plan (the captured graph keeps reading the plan's buffers)."""
    data = _small(seed=17, P=512, C=4)
    fresh = K.SwapPlan(gpu, data, QMX_PREC_FP32)
    fresh.run(3, graph=False)
    plan = K.SwapPlan(gpu, data, QMX_PREC_FP32)
    plan.run(2, graph=True)
    plan.set_centroids(data.init)
    plan.run(3, graph=True)
    gpu.qmxSync()
    assert np.array_equal(plan.p0(), fresh.p0())
    assert np.array_equal(plan.centroids(), fresh.centroids())
    fresh.close()
    plan.close()


@pytest.mark.gpu
def test_autotune_keeps_results(gpu):
    data = _small(seed=13, P=512, C=4, q_lo=3, q_hi=10)
    plan = K.SwapPlan(gpu, data, QMX_PREC_FP32)
    plan.run(1, graph=False)
    before = plan.p0()
    plan.close()
    plan = K.SwapPlan(gpu, data, QMX_PREC_FP32)
    assert plan.autotune() > 0
    for q in range(3, 11):
        assert 1 <= plan.rows_per_tile(q) <= 1024
    plan.run(1, graph=True)
    assert np.max(np.abs(plan.p0() - before)) < 1e-6
    plan.close()


def test_closed_form_cpu_port_matches_inner_products(cpuref):
    """The like-for-like CPU baseline (oracle/cpuref/swap_closed_form.c: every
    ancilla-0 amplitude squared and summed) equals (1 + <a|b>^2)/2 for
    normalised real registers of every size in the workload, and sums the
    unnormalised form too (|a|^2 |b|^2 + <a|b>^2)/2."""
    import ctypes as C
    rng = np.random.default_rng(17)
    qs = [1, 2, 4, 5, 7, 10, 3]
    a_parts, b_parts, want = [], [], []
    for q in qs:
        a = rng.standard_normal(1 << q)
        b = rng.standard_normal(1 << q)
        if q != 3:
            a /= np.linalg.norm(a)
            b /= np.linalg.norm(b)
        a_parts.append(a)
        b_parts.append(b)
        want.append((np.dot(a, a) * np.dot(b, b) + np.dot(a, b) ** 2) / 2)
    off = np.concatenate([[0], np.cumsum([1 << q for q in qs])[:-1]]).astype(np.int64)
    A, B = np.concatenate(a_parts), np.concatenate(b_parts)
    out = np.zeros(len(qs))
    fn = cpuref.lib.cpurefSwapP0ClosedForm
    dp = C.POINTER(C.c_double)
    lp = C.POINTER(C.c_longlong)
    fn.argtypes = [C.c_int, C.POINTER(C.c_int), dp, lp, dp, lp, dp]
    qa = np.array(qs, dtype=np.int32)
    fn(len(qs), qa.ctypes.data_as(C.POINTER(C.c_int)), A.ctypes.data_as(dp), off.ctypes.data_as(lp),
       B.ctypes.data_as(dp), off.ctypes.data_as(lp), out.ctypes.data_as(dp))
    assert np.max(np.abs(out - np.array(want))) < 1e-13
