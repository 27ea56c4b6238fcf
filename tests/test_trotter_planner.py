"""Learned Trotter planner: the instance family, features, labels and the MLP
(CPU: oracle pinning and the library's host code), then on the GPU the batched
small-register Trotter kernels against applyTrotterCircuit and QuEST's CPU
backend, the network's kernels against the oracle, and what the trained
planner gains over plain step doubling."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import trotter_oracle as O
from quantumlab_amd import hamiltonians as H
from quantumlab_amd import planner as PL


@pytest.fixture(scope="module")
def hostlib():
    """The library loaded without touching the device (host-only entry points)."""
    import quantumlab_amd
    return C.CDLL(quantumlab_amd.library_path(), mode=os.RTLD_LOCAL)


# ---- CPU ---------------------------------------------------------------------------

def test_instance_family_shape():
    for idx in range(8):
        inst = H.trotter_instance(10, seed=5, index=idx)
        k = H.default_strings(10)
        bonds = (10 - 1) * (1 if inst.model == "tfim" else 3)
        assert inst.terms == bonds + 10 + k
        assert 0.5 <= inst.J <= 1.5 and 0 <= inst.h <= 2 and 0.5 <= inst.time <= 4
        w = np.count_nonzero(inst.codes[bonds + 10:], axis=1)
        assert np.all((w >= 2) & (w <= 4))
        psi = H.product_state(inst)
        assert abs(np.vdot(psi, psi) - 1) < 1e-13
    assert H.default_strings(24) == 32
    assert [H.stages(o) for o in (1, 2, 4, 6)] == [1, 2, 10, 50]


def test_commutator_sums_match_dense_brute_force():
    """The mask formulas for C1 and C2 equal spectral norms of the dense
    commutators (a wrong anticommutation test or a dropped factor 2 fails)."""
    rng = np.random.default_rng(3)
    for n, T in ((2, 5), (3, 6), (4, 5)):
        codes = rng.integers(0, 4, size=(T, n)).astype(np.int32)
        codes[0] = 0  # an identity string, dropped by both
        coeffs = rng.standard_normal(T)
        C1, C2, _, _ = O.commutator_sums(codes, coeffs)
        D1, D2 = O.dense_commutator_sums(codes, coeffs)
        assert abs(C1 - D1) < 1e-10 * max(1, D1) and abs(C2 - D2) < 1e-10 * max(1, D2)


def test_commuting_hamiltonian_has_no_commutator_features():
    codes = np.array([[3, 3, 0], [0, 3, 3], [3, 0, 0]], dtype=np.int32)
    f = O.features(codes, np.array([1.0, -0.5, 0.3]), 2.0)
    assert f[7] == O.LOG_FLOOR and f[8] == O.LOG_FLOOR and f[12] == 0.0


@pytest.mark.parametrize("n", [4, 9, 12, 24])
def test_library_features_match_oracle(hostlib, n):
    for idx in range(3):
        inst = H.trotter_instance(n, seed=11, index=idx)
        got = PL.features(hostlib, inst)
        want = O.features(inst.codes, inst.coeffs, inst.time)
        assert np.allclose(got, want.astype(np.float32), rtol=2e-6, atol=1e-6), (n, idx, got, want)


def test_bf16_rounding_matches_torch():
    import torch
    x = np.random.default_rng(0).standard_normal(4096).astype(np.float32) * 10.0 ** np.random.default_rng(1).integers(-8, 8, 4096)
    x = np.concatenate([x, np.array([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8], dtype=np.float32)])  # ties
    want = torch.from_numpy(x.astype(np.float32)).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(O.bf16(x), want)


def test_oracle_gradient_matches_finite_differences():
    """Backward of the masked MSE against central differences (FP64, no rounding)."""
    rng = np.random.default_rng(4)
    p = O.init_params(7).astype(np.float64)
    p[O.NUM_PARAMS - 2:] = [0.3, -0.2]
    X = rng.standard_normal((9, 16))
    T = rng.standard_normal((9, 2))
    M = (rng.random((9, 2)) < 0.7).astype(np.float64)
    L, g = O.loss_and_grad(p, X, T, M, rounded=False, dtype=np.float64)
    idx = np.concatenate([rng.choice(O.NUM_PARAMS, 60, replace=False), np.arange(O.NUM_PARAMS - 130, O.NUM_PARAMS)])
    for i in idx:
        e = np.zeros_like(p)
        e[i] = 1e-6
        fd = (O.loss_and_grad(p + e, X, T, M, False, np.float64)[0] - O.loss_and_grad(p - e, X, T, M, False, np.float64)[0]) / 2e-6
        assert abs(fd - g[i]) < 1e-6 + 1e-5 * abs(g[i]), (i, fd, g[i])


@pytest.mark.parametrize("order", [2, 4])
def test_labels_recover_power_law_closed_form(order):
    """For ladders that follow 1 - F = A r^(-2o) exactly, the label is log2 of
    (A / eps)^(1 / (2o)) (the exact cheapest real-valued repetitions)."""
    eps = 1e-4
    for A in (3e-3, 0.7, 40.0, 9e3):
        ladder = np.array([A * (2.0 ** k) ** (-2 * order) for k in range(12)])
        want = np.log2((A / eps) ** (1 / (2 * order)))
        got = PL.label_log2_reps(ladder, order, eps)
        assert abs(got - max(0.0, want)) < 1e-12 or (want < 0 and got == 0.0)
    assert PL.label_log2_reps(np.full(5, 0.5), order, eps) is None


# ---- GPU ---------------------------------------------------------------------------

def _ref_evolve(qe, inst, order, reps):
    q = qe.qureg(inst.n)
    H.prepare(qe, q, inst)
    hb = H.HamilBatch([inst])
    qe.lib.applyTrotterCircuit(q, hb[0], inst.time, order, reps)
    s = qe.state(q)
    qe.destroy(q)
    return s


@pytest.mark.gpu
def test_batch_evolve_matches_apply_trotter_circuit(gpu, reflib):
    cases = []
    for j, n in enumerate((3, 5, 8, 10, 12)):
        for order, reps in ((1, 2), (2, 1), (4, 3), (6, 1)):
            cases.append((H.trotter_instance(n, seed=21, index=10 * j + order), order, reps))
    states = PL.batch_evolve(gpu, [c[0] for c in cases], [c[1] for c in cases], [c[2] for c in cases])
    for (inst, order, reps), got in zip(cases, states):
        assert np.max(np.abs(got - _ref_evolve(gpu, inst, order, reps))) < 1e-12, (inst.n, order, reps)
        if inst.n <= 8:
            assert np.max(np.abs(got - _ref_evolve(reflib, inst, order, reps))) < 1e-12, (inst.n, order, reps)
    # the product state alone (one repetition of an all-zero-time evolution is exact)
    inst = H.trotter_instance(7, seed=2, index=0)
    inst.time = 0.0
    got = PL.batch_evolve(gpu, [inst], [2], [1])[0]
    assert np.max(np.abs(got - H.product_state(inst))) < 1e-14


@pytest.mark.gpu
def test_ladder_matches_reference_infidelities(gpu, reflib):
    insts = [H.trotter_instance(6, seed=31, index=i) for i in range(4)]
    for order in (2, 4):
        inf, pas = PL.batch_ladder(gpu, insts, [order] * 4, 3, 1e-300)
        assert np.all(pas == -1)
        for j, inst in enumerate(insts):
            for k in range(4):
                a = _ref_evolve(reflib, inst, order, 1 << k)
                b = _ref_evolve(reflib, inst, order, 4 << k)
                want = 1 - abs(np.vdot(a, b)) ** 2
                # |d(1 - |<a|b>|^2)| <= 2 (||da|| + ||db||) <= 4 sqrt(2^n) max|d amp|:
                # amplitudes within 1e-12 give infidelities within 3.2e-11 at n = 6
                assert abs(inf[j, k] - want) < 3.2e-11, (order, j, k, inf[j, k], want)
        # with a real target the ladder stops at the first passing rung
        inf2, pas2 = PL.batch_ladder(gpu, insts, [order] * 4, 8, 1e-4)
        for j in range(4):
            k = pas2[j]
            assert k >= 0 and inf2[j, k] <= 1e-4 and np.all(inf2[j, :k] > 1e-4) and np.all(np.isnan(inf2[j, k + 1:]))
            m = min(k, 3) + 1
            assert np.array_equal(inf2[j, :m], inf[j, :m])  # same kernel, same rungs: bitwise


@pytest.mark.gpu
def test_cheapest_reps_is_the_first_passing_integer(gpu):
    """Bisection labels: the criterion holds at r* and fails at r* - 1."""
    insts = [H.trotter_instance(7, seed=61, index=i) for i in range(6)]
    for order in (2, 4):
        rs = PL.cheapest_reps(gpu, insts, order, 1e-4)
        assert all(r is not None for r in rs)
        at = PL.batch_infidelity(gpu, insts, [order] * 6, rs)
        assert np.all(at <= 1e-4)
        below = [(i, r - 1) for i, r in zip(insts, rs) if r > 1]
        if below:
            v = PL.batch_infidelity(gpu, [b[0] for b in below], [order] * len(below), [b[1] for b in below])
            assert np.all(v > 1e-4)
        # the batched judge agrees with evolving both states
        for inst, r in zip(insts[:2], rs[:2]):
            a, b = PL.batch_evolve(gpu, [inst, inst], [order, order], [r, 4 * r])
            assert abs((1 - abs(np.vdot(a, b)) ** 2) - PL.batch_infidelity(gpu, [inst], [order], [r])[0]) < 1e-12


def _mlp_case(seed, N):
    rng = np.random.default_rng(seed)
    p = O.init_params(seed)
    p[O.NUM_PARAMS - 2:] = [0.25, -0.5]
    X = rng.standard_normal((N, 16)).astype(np.float32)
    T = rng.standard_normal((N, 2)).astype(np.float32)
    M = (rng.random((N, 2)) < 0.8).astype(np.float32)
    return p, X, T, M


@pytest.mark.gpu
@pytest.mark.parametrize("N", [1, 64, 150])
def test_network_forward_and_gradient_match_oracle(gpu, N):
    p, X, T, M = _mlp_case(N, N)
    net = PL.Planner(gpu, seed=1)
    net.set_params(p)
    Y = net.predict(X)
    Yr, _ = O.forward(p, X)
    assert np.max(np.abs(Y - Yr)) < 1e-5 * max(1.0, np.max(np.abs(Yr)))
    loss, g = net.gradient(X, T, M)
    Lr, gr = O.loss_and_grad(p, X, T, M)
    assert abs(loss - Lr) < 1e-5 * max(1.0, Lr)
    scale = np.max(np.abs(gr))
    assert np.max(np.abs(g - gr)) < 2e-5 * scale, np.max(np.abs(g - gr)) / scale
    # parameters round-trip, and the weights the layers use are the bf16-rounded ones
    assert np.array_equal(net.params(), p)
    net.close()


@pytest.mark.gpu
def test_training_learns_and_is_deterministic(gpu):
    """Adam on a smooth synthetic target: the loss falls by 20x; two runs from
    the same start give identical parameters; one step equals the oracle's
    Adam step on the oracle's gradient."""
    rng = np.random.default_rng(8)
    X = rng.standard_normal((1000, 16)).astype(np.float32)
    W = rng.standard_normal((16, 2)).astype(np.float32) / 4
    T = (np.tanh(X @ W) + 0.1 * X[:, :2] ** 2).astype(np.float32)
    M = np.ones_like(T)
    runs = []
    for _ in range(2):
        net = PL.Planner(gpu, seed=3)
        losses = net.train(X, T, M, epochs=60, batch=128, lr=3e-3, seed=5)
        runs.append(net.params())
        net.close()
    assert losses[-1] < losses[0] / 20, losses[[0, -1]]
    assert np.array_equal(runs[0], runs[1])
    # one full-batch step against the oracle
    net = PL.Planner(gpu, seed=4)
    p0 = net.params()
    net.train(X[:256], T[:256], M[:256], epochs=1, batch=256, lr=1e-3, seed=0)
    _, g = O.loss_and_grad(p0, X[:256], T[:256], M[:256])
    want, _, _ = O.adam_step(p0.astype(np.float64), g.astype(np.float64), 0, 0, 1, 1e-3)
    # Adam's first step is lr * g / (|g| + 1e-8): exact where |g| is not tiny,
    # within lr where the sign of a near-zero gradient is rounding noise
    big = np.abs(g) > 1e-3 * np.max(np.abs(g))
    d = np.abs(net.params() - want)
    assert np.max(d[big]) < 1e-6 and np.max(d) <= 1.01e-3
    net.close()


@pytest.mark.gpu
def test_choose_follows_the_decision_rule(gpu):
    net = PL.Planner(gpu, seed=9)
    net.set_margin(0.37)
    for idx in range(6):
        inst = H.trotter_instance(8 + idx, seed=41, index=idx)
        f = PL.features(gpu.lib, inst)
        y = net.predict(f[None])[0]
        assert net.choose(inst) == O.choose(y, inst.n, 0.37)
    kx = np.float32([2.5, 3.0, 3.5, 4.6])
    ky = np.float32([[4.9, 2.4], [5.0, 2.5], [5.2, 2.6], [5.7, 3.0]])
    net.set_size_curve(kx, ky)
    gx, gy = net.size_curve()
    assert np.array_equal(gx, kx) and np.array_equal(gy, ky)
    for idx, n in enumerate((4, 9, 11, 12, 24, 30)):
        inst = H.trotter_instance(n, seed=42, index=idx)
        y = net.predict(PL.features(gpu.lib, inst)[None])[0]
        assert net.choose(inst) == O.choose(y, inst.n, 0.37, kx.astype(float), ky.astype(float))
        for c, o in enumerate((2, 4)):
            assert abs(gpu.lib.qmxPlannerSizeCurve(net.h, o, n) - PL.size_curve(kx.astype(float), ky.astype(float), np.log2(n))[0, c]) < 1e-6
    net.close()


@pytest.mark.gpu
def test_step_doubling_stops_at_the_first_agreeing_pair(gpu):
    inst = H.trotter_instance(8, seed=51, index=3)
    hb = H.HamilBatch([inst])
    q = gpu.qureg(8)
    H.prepare(gpu, q, inst)
    reps = C.c_int()
    ok = gpu.lib.qmxApplyTrotterStepDoubling(q, hb[0], inst.time, 2, 1e-4, 4096, C.byref(reps))
    assert ok == 1
    r = reps.value
    got = gpu.state(q)
    a = _ref_evolve(gpu, inst, 2, r // 2)
    b = _ref_evolve(gpu, inst, 2, r)
    assert np.max(np.abs(got - b)) < 1e-12
    assert 1 - abs(np.vdot(a, b)) ** 2 <= 1e-4
    if r >= 4:
        c = _ref_evolve(gpu, inst, 2, r // 4)
        assert 1 - abs(np.vdot(c, a)) ** 2 > 1e-4
    gpu.destroy(q)


@pytest.mark.gpu
def test_trained_planner_beats_step_doubling_on_a_larger_register(gpu):
    """Train on 6..8-qubit registers (labels from the batched ladders),
    calibrate the margin on labelled 9-qubit ones, judge on fresh 10-qubit ones
    never seen: >= 95% reach 1 - F <= 1e-4 against the 4x-finer FP64 run, with
    >= 2x fewer Pauli exponentials than step doubling at either fixed order."""
    from quantumlab_amd import train_planner as TP
    insts = TP.make_instances(77, 1500, [6, 7, 8])
    calib = TP.make_instances(78, 300, [9])
    X, Xc = PL.feature_matrix(gpu.lib, insts), PL.feature_matrix(gpu.lib, calib)
    L, Lc = TP.label(gpu, insts), TP.label(gpu, calib)
    assert np.isfinite(L).mean() > 0.99
    net, losses, info = TP.fit(gpu, X, L, [i.n for i in insts], Xc, Lc, [9] * len(calib), epochs=80, seed=3,
                               quantile=0.99)
    assert losses[-1] < losses[0] / 20
    ev = TP.evaluate(gpu, net, [H.trotter_instance(10, 99, i) for i in range(60)])
    assert ev["success_rate"] >= 0.95, ev
    assert ev["gain_vs_step_doubling_o2"] >= 2 and ev["gain_vs_step_doubling_o4"] >= 2, ev
    net.close()


@pytest.mark.gpu
def test_shipped_planner_file_meets_the_target(gpu):
    """The committed planner (trained by quantumlab_amd.train_planner) on fresh
    12-qubit instances."""
    from quantumlab_amd import train_planner as TP
    path = os.path.join(os.path.dirname(PL.__file__), "data", "trotter_planner.txt")
    if not os.path.exists(path):
        pytest.skip("no trained planner file in the package yet (python -m quantumlab_amd.train_planner)")
    net = PL.Planner(gpu, path=path)
    ev = TP.evaluate(gpu, net, [H.trotter_instance(12, 4321, i) for i in range(40)])
    assert ev["success_rate"] >= 0.95 and ev["gain_vs_step_doubling_o2"] >= 2, ev
    net.close()


def _wide_instance(n, seed):
    """The family's instance plus strings with more X/Y factors than a fused
    tile holds (applied one per sweep) and an identity string."""
    inst = H.trotter_instance(n, seed=seed, index=0)
    rng = np.random.default_rng(seed)
    extra = []
    for w in (min(n, 9), min(n, n - 1)):
        row = np.zeros(n, dtype=np.int32)
        qs = rng.choice(n, size=w, replace=False)
        row[qs] = rng.integers(1, 3, size=w)  # X or Y only
        extra.append(row)
    extra.append(np.zeros(n, dtype=np.int32))
    inst.codes = np.vstack([inst.codes[:5], extra, inst.codes[5:]])
    inst.coeffs = np.concatenate([inst.coeffs[:5], [0.3, -0.2, 0.7], inst.coeffs[5:]])
    return inst


@pytest.mark.gpu
@pytest.mark.parametrize("n,density", [(9, False), (14, False), (17, False), (6, True), (7, True)])
def test_fused_trotter_sweeps_match_reference_backend(gpu, cpuref, n, density):
    """applyTrotterCircuit runs its rotations as fused LDS-tile sweeps; against
    the C restatement of QuEST's CPU backend (one rotation at a time). FP64
    tiles hold 11 qubits: 9 qubits run the pass form, the rest the coset form."""
    from tests.cases import random_state
    inst = _wide_instance(n, seed=100 + n)
    for order, reps in ((1, 2), (2, 1), (4, 1)):
        outs = []
        for lib in (gpu, cpuref):
            q = lib.qureg(n, density=density)
            if density:
                lib.initDebugState(q)
            else:
                lib.set_state(q, random_state(n, seed=n))
            hb = H.HamilBatch([inst])
            lib.lib.applyTrotterCircuit(q, hb[0], 0.37, order, reps)
            outs.append(lib.state(q))
            lib.destroy(q)
        # the debug density state is unnormalised (entries up to ~10^4): relative bound
        scale = max(1.0, float(np.max(np.abs(outs[1]))))
        assert np.max(np.abs(outs[0] - outs[1])) < 1e-12 * scale, (n, density, order)


@pytest.mark.gpu
def test_fused_trotter_fp32_against_fp64(gpu):
    from quantumlab_amd.quest import QMX_PREC_FP32
    inst = _wide_instance(18, seed=7)
    hb = H.HamilBatch([inst])
    states = []
    for prec in (None, QMX_PREC_FP32):
        q = gpu.qureg(18, precision=prec)
        H.prepare(gpu, q, inst)
        gpu.lib.applyTrotterCircuit(q, hb[0], inst.time, 2, 3)
        states.append(q)
    ip = gpu.calcInnerProduct(states[0], states[1])
    fid = (ip.real ** 2 + ip.imag ** 2) / (gpu.calcTotalProb(states[0]) * gpu.calcTotalProb(states[1]))
    assert 1 - fid < 1e-6
    for q in states:
        gpu.destroy(q)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [10, 12, 13])
def test_fused_trotter_fp32_amplitudes(gpu, n):
    """FP32 tiles hold 12 qubits: 10 qubits run the pass form, 12 and 13 the
    coset form; every amplitude against the FP64 run of the same circuit."""
    from quantumlab_amd.quest import QMX_PREC_FP32
    from tests.cases import random_state
    inst = _wide_instance(n, seed=200 + n)
    hb = H.HamilBatch([inst])
    for order, reps in ((1, 2), (2, 1), (4, 1)):
        outs = []
        for prec in (None, QMX_PREC_FP32):
            q = gpu.qureg(n, precision=prec)
            gpu.set_state(q, random_state(n, seed=n))
            gpu.lib.applyTrotterCircuit(q, hb[0], 0.37, order, reps)
            outs.append(gpu.state(q))
            gpu.destroy(q)
        assert np.max(np.abs(outs[0] - outs[1])) < 2e-6, (n, order)
