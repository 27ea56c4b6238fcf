"""Pins oracle/encoding_oracle.py (CPU only): the restated FR gate lists against
the definition of a uniformly controlled gate, brute force on dense vectors,
and the complex encoder's gate list against its closed form. A dropped CNOT, a
control index off by one, a wrong Gray-code step, a sign of Z or a swapped
Ry / Rz order fails one of these."""
import numpy as np
import pytest

from oracle import encoding_oracle as E


def random_psi(n, rng):
    psi = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    return psi / np.linalg.norm(psi)


@pytest.mark.parametrize("kind", ["ry", "rz", "p"])
@pytest.mark.parametrize("width", [1, 2, 3, 5])
def test_fr_gate_lists_are_uniformly_controlled_rotations(kind, width):
    rng = np.random.default_rng(width * 7 + len(kind))
    n = width + 2
    bits = [int(b) for b in rng.permutation(n)[:width]]
    angles = rng.uniform(-4, 4, 1 << (width - 1))
    psi = random_psi(n, rng)
    gates = E.frp_gates(angles, bits) if kind == "p" else E.fr_gates(kind, angles, bits)
    got = E.run_gates_dense(gates, n, psi)
    controls, target = E.fr_controls(kind, bits)
    want = E.uniform_dense(psi, n, controls, target, [E.rotation(kind, a) for a in angles])
    assert np.max(np.abs(got - want)) < 1e-13


def test_fr_angle_index_conventions_differ():
    """FRy reads its angle index with local qubit 0 most significant, FRp with
    local qubit 0 least significant: swapping the two mappings must fail."""
    rng = np.random.default_rng(1)
    bits, n = [0, 1, 2], 3
    angles = rng.uniform(-4, 4, 4)
    psi = random_psi(n, rng)
    fry = E.run_gates_dense(E.fr_gates("ry", angles, bits), n, psi)
    wrong = E.uniform_dense(psi, n, list(bits[:2]), 2, [E.rotation("ry", a) for a in angles])
    assert np.max(np.abs(fry - wrong)) > 1e-3
    frp = E.run_gates_dense(E.frp_gates(angles, bits), n, psi)
    wrong = E.uniform_dense(psi, n, [1, 0], 2, [E.rotation("p", a) for a in angles])
    assert np.max(np.abs(frp - wrong)) > 1e-3


@pytest.mark.parametrize("width", [1, 2, 4])
def test_fryz_is_ry_then_rz_of_minus_z(width):
    rng = np.random.default_rng(width)
    n = width + 1
    bits = [int(b) for b in rng.permutation(n)[:width]]
    ay, az = rng.uniform(-4, 4, 1 << (width - 1)), rng.uniform(-4, 4, 1 << (width - 1))
    psi = random_psi(n, rng)
    got = E.run_gates_dense(E.fryz_gates(ay, az, bits), n, psi)
    controls, target = E.fr_controls("ry", bits)
    mats = [E.rotation("rz", -z) @ E.rotation("ry", y) for y, z in zip(ay, az)]
    assert np.max(np.abs(got - E.uniform_dense(psi, n, controls, target, mats))) < 1e-13


@pytest.mark.parametrize("q", [1, 2, 4, 7])
def test_complex_encoder_prepares_v_up_to_its_mean_phase(q):
    rng = np.random.default_rng(q)
    v = rng.standard_normal(1 << q) + 1j * rng.standard_normal(1 << q)
    got = E.run_gates_dense(E.ae_gates_complex(v, list(range(q))), q)
    want = v / np.linalg.norm(v) * np.exp(-1j * np.mean(np.angle(v)))
    assert np.max(np.abs(got - want)) < 1e-13
    # the angle lists: Z's last entry is the mean phase (the skipped phase gate)
    _, Z = E.ae_angles_complex(v)
    assert abs(Z[-1] - np.mean(np.angle(v))) < 1e-13


def test_complex_encoder_of_a_real_vector_is_the_real_encoder():
    from oracle import swaptest_oracle as O
    rng = np.random.default_rng(5)
    v = np.abs(rng.standard_normal(32))
    got = E.run_gates_dense(E.ae_gates_complex(v.astype(complex), list(range(5))), 5)
    assert np.max(np.abs(got - O.ae_circuit_dense(v))) < 1e-13
