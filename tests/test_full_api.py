"""The rest of QuEST's API surface: diagonal operators, phase functions and the
QFT (QuEST.c:732-891, 1136-1244, 1557-1606).

Each case is checked two ways at 1e-12 max-abs: against a numpy model of the
reference's formulas (QuEST_cpu.c:4030-4209 diagonal operators, 4251-4620
phase functions, QuEST_common.c:846-908 QFT), and, where oracle/_ref is
built, against QuEST's own CPU backend on the same inputs. The numpy model is
itself pinned against that backend by the CPU tests below."""
import ctypes as C
import os
import tempfile

import numpy as np
import pytest

from quantumlab_amd.quest import QuESTError, qreal
from tests.cases import ints, random_state

TOL = 1e-12


def reals(xs):
    xs = [float(x) for x in xs]
    return (qreal * max(1, len(xs)))(*xs)


def longs(xs):
    xs = [int(x) for x in xs]
    return (C.c_longlong * max(1, len(xs)))(*xs)


# ---- numpy model --------------------------------------------------------------------

def reg_values(index, qubits, per_reg, encoding):
    vals, f = [], 0
    for nq in per_reg:
        v = 0
        bits = nq - 1 if encoding == 1 else nq
        for b in range(bits):
            v += ((index >> qubits[f]) & 1) << b
            f += 1
        if encoding == 1:
            if (index >> qubits[f]) & 1:
                v -= 1 << (nq - 1)
            f += 1
        vals.append(v)
    return vals


def poly_phase(vals, coeffs, exps, terms_per_reg):
    ph, t = 0.0, 0
    for r, v in enumerate(vals):
        for _ in range(terms_per_reg[r]):
            ph += coeffs[t] * float(v) ** exps[t]
            t += 1
    return ph


def named_phase(vals, code, params):
    """The named functions (QuEST_cpu.c:4530-4600): NORM family 0-4, PRODUCT
    family 5-8, DISTANCE family 9-13 over consecutive register pairs."""
    v = [float(x) for x in vals]
    eps = 1e-13
    if code <= 4:
        if code == 4:
            m = np.sqrt(sum((x - params[2 + r]) ** 2 for r, x in enumerate(v)))
        else:
            m = np.sqrt(sum(x * x for x in v))
    elif code <= 8:
        m = float(np.prod(v))
    elif code == 13:
        m = np.sqrt(sum((v[r] - v[r + 1] - params[2 + r // 2]) ** 2 for r in range(0, len(v), 2)))
    else:
        m = np.sqrt(sum((v[r + 1] - v[r]) ** 2 for r in range(0, len(v), 2)))
    kind = {0: "plain", 5: "plain", 9: "plain", 1: "scaled", 6: "scaled", 10: "scaled",
            2: "inverse", 7: "inverse", 11: "inverse"}.get(code, "scaled_inverse")
    if kind == "plain":
        return m
    if kind == "scaled":
        return params[0] * m
    if kind == "inverse":
        return params[0] if m == 0 else 1 / m
    # scaled inverse: products diverge exactly at 0, norms and distances within REAL_EPS
    if code == 8:
        return params[1] if m == 0 else params[0] / m
    return params[1] if m <= eps else params[0] / m


def phase_table(n, qubits, per_reg, encoding, fn, overrides=()):
    """Phase of each basis state of n qubits; overrides map index tuples to phases."""
    ov = {tuple(k): p for k, p in overrides}
    out = np.empty(1 << n)
    for i in range(1 << n):
        vals = reg_values(i, qubits, per_reg, encoding)
        out[i] = ov[tuple(vals)] if tuple(vals) in ov else fn(vals)
    return out


def apply_phases(flat, n, phases, density):
    if not density:
        return flat * np.exp(1j * phases)
    dim = 1 << n
    rho = flat.reshape(dim, dim).T   # rho[r, c]
    rho = np.exp(1j * phases)[:, None] * rho * np.exp(-1j * phases)[None, :]
    return rho.T.reshape(-1)


def qft_model(flat, n, qubits, density):
    """|x> -> 2^-k/2 Σ_y exp(2 pi i x y / 2^k) |y> on the sub-register (qubits[0] least significant)."""
    k = len(qubits)
    F = np.exp(2j * np.pi * np.outer(np.arange(1 << k), np.arange(1 << k)) / (1 << k)) / np.sqrt(1 << k)
    dim = 1 << n
    U = np.zeros((dim, dim), dtype=complex)
    for i in range(dim):
        x = sum(((i >> q) & 1) << j for j, q in enumerate(qubits))
        base = i
        for q in qubits:
            base &= ~(1 << q)
        for y in range(1 << k):
            o = base
            for j, q in enumerate(qubits):
                if (y >> j) & 1:
                    o |= 1 << q
            U[o, i] += F[y, x]
    if not density:
        return U @ flat
    rho = flat.reshape(dim, dim).T
    return (U @ rho @ U.conj().T).T.reshape(-1)


# ---- cases -----------------------------------------------------------------------------

def prepared(lib, n, density, seed):
    psi = random_state(n, seed)
    if not density:
        q = lib.qureg(n)
        lib.set_state(q, psi)
        return q
    p = lib.qureg(n)
    lib.set_state(p, psi)
    q = lib.qureg(n, density=True)
    lib.initPureState(q, p)
    lib.call("pauliX", q, 0)        # make it a mixed-looking generic state: mix in another pure state
    other = lib.qureg(n, density=True)
    lib.set_state(p, random_state(n, seed + 1))
    lib.initPureState(other, p)
    lib.call("mixDensityMatrix", q, 0.3, other)
    lib.destroy(p)
    lib.destroy(other)
    return q


PHASE_CASES = [
    # (name, qubits, encoding, coeffs, exponents, overrides ((index,), phase))
    ("poly_unsigned", [3, 0, 4], 0, [0.7, -1.3], [1.0, 2.0], ()),
    ("poly_twos", [1, 4, 2, 5], 1, [0.5, 0.25, -0.1], [1.0, 3.0, 0.0], ()),
    ("poly_negative_exp", [2, 0, 5], 0, [1.1, 0.4], [-1.0, 1.0], [((0,), 0.9)]),
    ("poly_fraction_twos", [0, 3], 1, [0.8], [0.5], [((-1,), 0.3), ((-2,), -0.2)]),
]


def run_single_var(lib, q, case):
    _, qubits, enc, coeffs, exps, overrides = case
    if overrides:
        lib.call("applyPhaseFuncOverrides", q, ints(qubits), len(qubits), enc, reals(coeffs), reals(exps), len(coeffs),
                 longs([k[0] for k, _ in overrides]), reals([p for _, p in overrides]), len(overrides))
    else:
        lib.call("applyPhaseFunc", q, ints(qubits), len(qubits), enc, reals(coeffs), reals(exps), len(coeffs))


def model_single_var(n, case):
    _, qubits, enc, coeffs, exps, overrides = case
    return phase_table(n, qubits, [len(qubits)], enc,
                       lambda v: poly_phase(v, coeffs, exps, [len(coeffs)]), overrides)


MULTI_CASE = ([1, 0, 3, 5, 4], [2, 3], 0, [0.3, -0.2, 0.7], [1.0, 2.0, 1.0], [2, 1], [((1, 3), 0.5), ((0, 0), -1.0)])

NAMED_CASES = [
    # (qubits, per_reg, encoding, code, params, overrides)
    ([0, 1, 2, 3], [2, 2], 0, 0, [], []),                               # NORM
    ([0, 1, 2, 3, 4], [3, 2], 1, 1, [0.6], []),                         # SCALED_NORM, two's complement
    ([5, 1, 2, 0], [2, 2], 0, 2, [2.5], []),                            # INVERSE_NORM (diverges at 0)
    ([0, 1, 2, 3], [2, 2], 1, 3, [0.7, 1.9], []),                       # SCALED_INVERSE_NORM
    ([0, 1, 2, 3], [2, 2], 0, 4, [0.7, 1.9, 0.5, -1.0], []),            # SCALED_INVERSE_SHIFTED_NORM
    ([3, 1, 0, 2], [2, 2], 0, 5, [], [((1, 1), 0.25)]),                  # PRODUCT with an override
    ([0, 1, 2, 3], [2, 2], 1, 6, [-0.4], []),                           # SCALED_PRODUCT
    ([0, 1, 2, 3], [2, 2], 0, 7, [3.0], []),                            # INVERSE_PRODUCT
    ([0, 1, 2, 3], [2, 2], 0, 8, [0.5, 1.5], []),                       # SCALED_INVERSE_PRODUCT
    ([0, 1, 2, 3, 4, 5], [2, 1, 2, 1], 0, 9, [], []),                   # DISTANCE over two pairs
    ([0, 1, 2, 3], [2, 2], 1, 10, [0.9], []),                           # SCALED_DISTANCE
    ([0, 1, 2, 3], [2, 2], 0, 11, [1.7], []),                           # INVERSE_DISTANCE
    ([0, 1, 2, 3], [2, 2], 0, 12, [0.4, 2.2], []),                      # SCALED_INVERSE_DISTANCE
    ([0, 1, 2, 3], [2, 2], 1, 13, [0.4, 2.2, 0.5], []),                 # SCALED_INVERSE_SHIFTED_DISTANCE
]


def run_named(lib, q, case):
    qubits, per, enc, code, params, overrides = case
    if overrides:
        flat = [x for k, _ in overrides for x in k]
        if params:
            lib.call("applyParamNamedPhaseFuncOverrides", q, ints(qubits), ints(per), len(per), enc, code,
                     reals(params), len(params), longs(flat), reals([p for _, p in overrides]), len(overrides))
        else:
            lib.call("applyNamedPhaseFuncOverrides", q, ints(qubits), ints(per), len(per), enc, code, longs(flat),
                     reals([p for _, p in overrides]), len(overrides))
    elif params:
        lib.call("applyParamNamedPhaseFunc", q, ints(qubits), ints(per), len(per), enc, code, reals(params),
                 len(params))
    else:
        lib.call("applyNamedPhaseFunc", q, ints(qubits), ints(per), len(per), enc, code)


def model_named(n, case):
    qubits, per, enc, code, params, overrides = case
    return phase_table(n, qubits, per, enc, lambda v: named_phase(v, code, params), overrides)


def run_multi(lib, q):
    qubits, per, enc, coeffs, exps, terms, overrides = MULTI_CASE
    flat = [x for k, _ in overrides for x in k]
    lib.call("applyMultiVarPhaseFuncOverrides", q, ints(qubits), ints(per), len(per), enc, reals(coeffs), reals(exps),
             ints(terms), longs(flat), reals([p for _, p in overrides]), len(overrides))


def model_multi(n):
    qubits, per, enc, coeffs, exps, terms, overrides = MULTI_CASE
    return phase_table(n, qubits, per, enc, lambda v: poly_phase(v, coeffs, exps, terms), overrides)


# ---- CPU: the numpy model against QuEST's own backend ----------------------------------

@pytest.mark.parametrize("density", [False, True])
def test_model_matches_reference_phase_functions(reflib, density):
    n = 6 if not density else 4
    for case in PHASE_CASES:
        if max(case[1]) >= n:
            continue
        q = prepared(reflib, n, density, 3)
        before = reflib.state(q)
        run_single_var(reflib, q, case)
        want = apply_phases(before, n, model_single_var(n, case), density)
        assert np.max(np.abs(reflib.state(q) - want)) < TOL, case[0]
        reflib.destroy(q)
    for case in NAMED_CASES:
        if max(case[0]) >= n:
            continue
        q = prepared(reflib, n, density, 4)
        before = reflib.state(q)
        run_named(reflib, q, case)
        want = apply_phases(before, n, model_named(n, case), density)
        assert np.max(np.abs(reflib.state(q) - want)) < TOL, case[3]
        reflib.destroy(q)


def test_model_matches_reference_qft_and_diagonal(reflib):
    n = 5
    for density in (False, True):
        nn = n if not density else 3
        q = prepared(reflib, nn, density, 7)
        before = reflib.state(q)
        reflib.call("applyQFT", q, ints([2, 0] if density else [3, 0, 4]), 2 if density else 3)
        want = qft_model(before, nn, [2, 0] if density else [3, 0, 4], density)
        assert np.max(np.abs(reflib.state(q) - want)) < TOL
        reflib.destroy(q)
    q = prepared(reflib, n, False, 8)
    psi = reflib.state(q)
    op = reflib.lib.createDiagonalOp(n, reflib.env)
    d = np.random.default_rng(9).standard_normal(1 << n) + 1j * np.random.default_rng(10).standard_normal(1 << n)
    reflib.lib.initDiagonalOp(op, reals(d.real), reals(d.imag))
    e = reflib.lib.calcExpecDiagonalOp(q, op)
    assert abs(complex(e.real, e.imag) - np.sum(np.abs(psi) ** 2 * d)) < TOL
    reflib.lib.applyDiagonalOp(q, op)
    assert np.max(np.abs(reflib.state(q) - d * psi)) < TOL
    reflib.lib.destroyDiagonalOp(op, reflib.env)
    reflib.destroy(q)


# ---- GPU: the library against the model and the reference ----------------------------

def both(gpu, reflib_or_none, n, density, seed, fn):
    out = []
    for lib in [gpu] + ([reflib_or_none] if reflib_or_none is not None else []):
        q = prepared(lib, n, density, seed)
        before = lib.state(q)
        fn(lib, q)
        out.append((before, lib.state(q)))
        lib.destroy(q)
    return out


@pytest.fixture
def ref_or_none(request):
    try:
        return request.getfixturevalue("reflib")
    except pytest.skip.Exception:
        return None


@pytest.mark.gpu
@pytest.mark.parametrize("density", [False, True])
def test_phase_functions(gpu, ref_or_none, density):
    n = 6 if not density else 4
    for i, case in enumerate(PHASE_CASES):
        if max(case[1]) >= n:
            continue
        res = both(gpu, ref_or_none, n, density, 20 + i, lambda lib, q: run_single_var(lib, q, case))
        want = apply_phases(res[0][0], n, model_single_var(n, case), density)
        assert np.max(np.abs(res[0][1] - want)) < TOL, case[0]
        if len(res) > 1:
            assert np.max(np.abs(res[0][1] - res[1][1])) < TOL, case[0]
    if not density:
        res = both(gpu, ref_or_none, n, density, 30, run_multi)
        want = apply_phases(res[0][0], n, model_multi(n), density)
        assert np.max(np.abs(res[0][1] - want)) < TOL
        if len(res) > 1:
            assert np.max(np.abs(res[0][1] - res[1][1])) < TOL
    for i, case in enumerate(NAMED_CASES):
        if max(case[0]) >= n:
            continue
        res = both(gpu, ref_or_none, n, density, 40 + i, lambda lib, q: run_named(lib, q, case))
        want = apply_phases(res[0][0], n, model_named(n, case), density)
        assert np.max(np.abs(res[0][1] - want)) < TOL, case[3]
        if len(res) > 1:
            assert np.max(np.abs(res[0][1] - res[1][1])) < TOL, case[3]


@pytest.mark.gpu
@pytest.mark.parametrize("density,n,qubits", [(False, 7, [4, 1, 6, 0]), (False, 6, None), (True, 4, [3, 1, 2]),
                                               (True, 3, None)])
def test_qft(gpu, ref_or_none, density, n, qubits):
    def fn(lib, q):
        if qubits is None:
            lib.call("applyFullQFT", q)
        else:
            lib.call("applyQFT", q, ints(qubits), len(qubits))
    res = both(gpu, ref_or_none, n, density, 50, fn)
    want = qft_model(res[0][0], n, list(range(n)) if qubits is None else qubits, density)
    assert np.max(np.abs(res[0][1] - want)) < TOL
    if len(res) > 1:
        assert np.max(np.abs(res[0][1] - res[1][1])) < TOL


@pytest.mark.gpu
@pytest.mark.parametrize("density", [False, True])
def test_diagonal_operator(gpu, ref_or_none, density):
    n = 6 if not density else 4
    rng = np.random.default_rng(60)
    d = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    part = rng.standard_normal(5) + 1j * rng.standard_normal(5)
    d2 = d.copy()
    d2[3:8] = part
    results = []
    for lib in [gpu] + ([ref_or_none] if ref_or_none is not None else []):
        q = prepared(lib, n, density, 61)
        before = lib.state(q)
        op = lib.lib.createDiagonalOp(n, lib.env)
        lib.lib.initDiagonalOp(op, reals(d.real), reals(d.imag))
        e1 = lib.lib.calcExpecDiagonalOp(q, op)
        lib.lib.setDiagonalOpElems(op, 3, reals(part.real), reals(part.imag), 5)
        lib.lib.applyDiagonalOp(q, op)
        e2 = lib.lib.calcExpecDiagonalOp(q, op)
        results.append((before, lib.state(q), complex(e1.real, e1.imag), complex(e2.real, e2.imag)))
        lib.lib.destroyDiagonalOp(op, lib.env)
        lib.destroy(q)
    before, after, e1, e2 = results[0]
    dim = 1 << n
    if density:
        rho = before.reshape(dim, dim).T
        assert abs(e1 - np.sum(np.diag(rho) * d)) < TOL
        want = (d2[:, None] * rho).T.reshape(-1)
        assert abs(e2 - np.sum(np.diag(d2[:, None] * rho) * d2)) < TOL
    else:
        assert abs(e1 - np.sum(np.abs(before) ** 2 * d)) < TOL
        want = d2 * before
        assert abs(e2 - np.sum(np.abs(want) ** 2 * d2)) < TOL
    assert np.max(np.abs(after - want)) < TOL
    if len(results) > 1:
        r = results[1]
        assert np.max(np.abs(after - r[1])) < TOL and abs(e1 - r[2]) < TOL and abs(e2 - r[3]) < TOL


@pytest.mark.gpu
def test_diagonal_operator_from_pauli_hamiltonian_and_file(gpu):
    n = 4
    codes = [[3, 0, 3, 0], [0, 3, 3, 3], [0, 0, 0, 0]]
    coeffs = [0.5, -1.25, 0.3]
    want = np.zeros(1 << n)
    for i in range(1 << n):
        for c, t in zip(coeffs, codes):
            par = sum((i >> q) & 1 for q in range(n) if t[q] == 3) & 1
            want[i] += -c if par else c
    h = gpu.lib.createPauliHamil(n, 3)
    for t in range(3):
        h.termCoeffs[t] = coeffs[t]
        for q in range(n):
            h.pauliCodes[t * n + q] = codes[t][q]
    op = gpu.lib.createDiagonalOp(n, gpu.env)
    gpu.call("initDiagonalOpFromPauliHamil", op, h)
    assert np.max(np.abs(np.ctypeslib.as_array(op.real, shape=(1 << n,)) - want)) < TOL
    # the device copy took the same values: apply to the all-ones-amplitude state
    q = gpu.qureg(n)
    gpu.lib.initPlusState(q)
    gpu.lib.applyDiagonalOp(q, op)
    assert np.max(np.abs(gpu.state(q) - want / 4.0)) < TOL
    # host edits reach the device after syncDiagonalOp
    op.real[0] = 7.0
    gpu.lib.syncDiagonalOp(op)
    gpu.lib.initPlusState(q)
    gpu.lib.applyDiagonalOp(q, op)
    assert abs(gpu.state(q)[0] - 7.0 / 4.0) < TOL
    gpu.lib.destroyDiagonalOp(op, gpu.env)
    gpu.lib.destroyPauliHamil(h)
    with tempfile.TemporaryDirectory() as d:
        fn = os.path.join(d, "hamil.txt")
        with open(fn, "w") as f:
            for c, t in zip(coeffs, codes):
                f.write(f"{c} " + " ".join(str(x) for x in t) + "\n")
        op = gpu.call("createDiagonalOpFromPauliHamilFile", fn.encode(), gpu.env)
        assert np.max(np.abs(np.ctypeslib.as_array(op.real, shape=(1 << n,)) - want)) < TOL
        gpu.lib.destroyDiagonalOp(op, gpu.env)
        with open(fn, "w") as f:
            f.write("0.5 1 0 0 0\n")
        with pytest.raises(QuESTError, match="PAULI_Z and PAULI_I"):
            gpu.call("createDiagonalOpFromPauliHamilFile", fn.encode(), gpu.env)
    gpu.destroy(q)


@pytest.mark.gpu
def test_full_api_validation_messages(gpu):
    q = gpu.qureg(4)
    rho = gpu.qureg(3, density=True)
    op = gpu.lib.createDiagonalOp(3, gpu.env)
    with pytest.raises(QuESTError, match="equal number of qubits"):
        gpu.call("applyDiagonalOp", q, op)
    with pytest.raises(QuESTError, match="More elements given"):
        gpu.call("setDiagonalOpElems", op, 6, reals([1, 2, 3]), reals([0, 0, 0]), 3)
    with pytest.raises(QuESTError, match="bit encoding"):
        gpu.call("applyPhaseFunc", q, ints([0, 1]), 2, 5, reals([1]), reals([1]), 1)
    with pytest.raises(QuESTError, match="TWOS_COMPLEMENT encoding. Must use >1"):
        gpu.call("applyPhaseFunc", q, ints([0]), 1, 1, reals([1]), reals([1]), 1)
    with pytest.raises(QuESTError, match="zero index was not overriden"):
        gpu.call("applyPhaseFunc", q, ints([0, 1]), 2, 0, reals([1]), reals([-1]), 1)
    with pytest.raises(QuESTError, match="number of parameters"):
        gpu.call("applyParamNamedPhaseFunc", q, ints([0, 1]), ints([1, 1]), 2, 0, 6, reals([1, 2]), 2)
    with pytest.raises(QuESTError, match="even number of sub-registers"):
        gpu.call("applyNamedPhaseFunc", q, ints([0, 1, 2]), ints([1, 1, 1]), 3, 0, 9)
    with pytest.raises(QuESTError, match="unique"):
        gpu.call("applyQFT", q, ints([1, 1]), 2)
    with pytest.raises(QuESTError, match="applyPhaseFunc"):
        gpu.call("applyPhaseFunc", q, ints([0, 1]), 2, 0, reals([1]), reals([1]), 0)
    gpu.lib.destroyDiagonalOp(op, gpu.env)
    gpu.destroy(q)
    gpu.destroy(rho)


# ---- combinations of registers, bit for bit against QuEST's own build -------------------

@pytest.mark.gpu
def test_weighted_and_mixed_registers_bit_identical_to_reference(gpu, reflib):
    """setWeightedQureg and mixDensityMatrix round every operation in QuEST's
    order (QuEST_cpu.c:3988-4027, 908-936), so the same inputs give the same
    bits as QuEST's CPU backend built without contraction. Values as in QuEST's
    suite (test_state_initialisations.cpp:540-610: entries and factors drawn
    from [-5, 5], the output aliasing each input in turn, then all three), where
    three successive weightings grow entries past 1e3 and its 1e-12 tolerance
    is below one ulp."""
    from quantumlab_amd.quest import cplx
    n = 5
    rng = np.random.default_rng(592)
    vecs = [rng.uniform(-5, 5, 1 << n) + 1j * rng.uniform(-5, 5, 1 << n) for _ in range(3)]
    regs = {}
    for name, lib in (("gpu", gpu), ("ref", reflib)):
        qs = []
        for v in vecs:
            q = lib.qureg(n)
            lib.set_state(q, v)
            qs.append(q)
        regs[name] = qs
    facs = [cplx(complex(rng.uniform(-5, 5), rng.uniform(-5, 5))) for _ in range(3)]
    fa, fb, fc = facs
    patterns = [(fa, 0, fb, 1, fc, 2), (fb, 1, fc, 2, fa, 2), (fc, 2, fb, 1, fa, 2), (fa, 2, fb, 2, fc, 2)]
    for f1, i1, f2, i2, fo, io in patterns:
        for name, lib in (("gpu", gpu), ("ref", reflib)):
            qs = regs[name]
            lib.call("setWeightedQureg", f1, qs[i1], f2, qs[i2], fo, qs[io])
        for k in range(3):
            assert np.array_equal(gpu.state(regs["gpu"][k]), reflib.state(regs["ref"][k])), (k, io)
    assert np.max(np.abs(reflib.state(regs["ref"][2]))) > 1e3
    for name, lib in (("gpu", gpu), ("ref", reflib)):
        for q in regs[name]:
            lib.destroy(q)
    # mixDensityMatrix on two density registers holding the same arbitrary bits
    m = 4
    flats = [rng.uniform(-1, 1, 4 ** m) + 1j * rng.uniform(-1, 1, 4 ** m) for _ in range(2)]
    rhos = {}
    for name, lib in (("gpu", gpu), ("ref", reflib)):
        pair = []
        for f in flats:
            q = lib.qureg(m, density=True)
            re, im = lib.reals(f.real), lib.reals(f.imag)
            lib.call("setDensityAmps", q, 0, 0, re.ctypes.data_as(lib.real_ptr), im.ctypes.data_as(lib.real_ptr),
                     4 ** m)
            pair.append(q)
        a, b = pair
        lib.call("mixDensityMatrix", a, 0.37, b)
        lib.call("mixDensityMatrix", a, 0.11, b)
        rhos[name] = (a, b)
    assert np.array_equal(gpu.state(rhos["gpu"][0]), reflib.state(rhos["ref"][0]))
    for name, lib in (("gpu", gpu), ("ref", reflib)):
        for q in rhos[name]:
            lib.destroy(q)
