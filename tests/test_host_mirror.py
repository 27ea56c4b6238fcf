"""The host mirror and text dumps at the QuEST boundary, through the C ABI.

QuantumLab uploads states by writing Qureg.stateVec.real/imag and calling
copyStateToGPU (QL/Simulator/QLSimulatorVector.cpp:41-51,119); QuEST's tests
use setAmps / setDensityAmps / copySubstateToGPU (QT/test_state_initialisations.cpp,
QT/test_data_structures.cpp). Expected states are the written values
themselves (the upload is data movement: exact in FP64, one rounding to float
in FP32 storage) and, where the operation has a CPU counterpart, QuEST's own CPU
build (oracle/_ref) fed the same inputs. reportState's CSV must equal the
reference's file byte for byte; QuantumLab's wave-function CSV
(Fermion1D.cpp:13-83) must round-trip."""
import ctypes as C
import os

import numpy as np
import pytest

from quantumlab_amd.quest import QMX_PREC_FP32, QMX_PREC_FP64, QuESTError, qreal
from tests.cases import random_state

pytestmark = pytest.mark.gpu


def mirror(q):
    n = q.numAmpsTotal
    return (np.ctypeslib.as_array(q.stateVec.real, shape=(n,)), np.ctypeslib.as_array(q.stateVec.imag, shape=(n,)))


def amps_via_getamp(gpu, q, idx):
    out = []
    for i in idx:
        z = gpu.getAmp(q, int(i))
        out.append(z.real + 1j * z.imag)
    return np.array(out)


def _round(v, prec):
    return v.astype(np.complex64).astype(np.complex128) if prec == QMX_PREC_FP32 else v


@pytest.mark.parametrize("prec", [QMX_PREC_FP64, QMX_PREC_FP32])
def test_copy_state_to_gpu_quantumlab_protocol(gpu, prec):
    """memcpy into stateVec, copyStateToGPU, then every read path sees it."""
    n = 11
    psi = random_state(n, 1)
    q = gpu.qureg(n, precision=prec)
    re, im = mirror(q)
    re[:] = psi.real
    im[:] = psi.imag
    gpu.copyStateToGPU(q)
    re[:] = 7.0  # the mirror is only read by copyStateToGPU
    im[:] = -7.0
    want = _round(psi, prec)
    idx = np.random.default_rng(2).choice(1 << n, 64, replace=False)
    assert np.array_equal(amps_via_getamp(gpu, q, idx), want[idx])
    assert np.array_equal(gpu.state(q), want)
    assert abs(gpu.calcTotalProb(q) - np.vdot(want, want).real) < 1e-12
    gpu.destroy(q)


@pytest.mark.parametrize("prec", [QMX_PREC_FP64, QMX_PREC_FP32])
@pytest.mark.parametrize("start,count", [(0, 1), (5, 17), (1000, 1048), (2047, 1)])
def test_copy_substate_to_gpu_uploads_only_its_range(gpu, prec, start, count):
    n = 11
    base, new = random_state(n, 3), random_state(n, 4)
    q = gpu.qureg(n, precision=prec)
    gpu.set_state(q, base)
    re, im = mirror(q)
    re[:] = new.real  # the whole mirror changes; only [start, start+count) is uploaded
    im[:] = new.imag
    gpu.copySubstateToGPU(q, start, count)
    want = _round(base.copy(), prec)
    want[start:start + count] = _round(new[start:start + count], prec)
    assert np.array_equal(gpu.state(q), want)
    gpu.destroy(q)


@pytest.mark.parametrize("prec", [QMX_PREC_FP64, QMX_PREC_FP32])
def test_copy_substate_from_gpu_downloads_only_its_range(gpu, prec):
    n = 9
    psi = random_state(n, 5)
    q = gpu.qureg(n, precision=prec)
    gpu.set_state(q, psi)
    re, im = mirror(q)
    re[:] = 3.0
    im[:] = 3.0
    gpu.copySubstateFromGPU(q, 100, 50)
    want = _round(psi, prec)
    assert np.array_equal(re[100:150] + 1j * im[100:150], want[100:150])
    assert np.all(re[:100] == 3.0) and np.all(re[150:] == 3.0)
    gpu.destroy(q)


@pytest.mark.parametrize("prec", [QMX_PREC_FP64, QMX_PREC_FP32])
def test_set_amps_matches_reference(gpu, reflib, prec):
    n = 8
    base = random_state(n, 6)
    vals = random_state(n, 7)[:40] * 3.0
    re = np.ascontiguousarray(vals.real)
    im = np.ascontiguousarray(vals.imag)
    ptr = lambda a: a.ctypes.data_as(C.POINTER(qreal))
    q, r = gpu.qureg(n, precision=prec), reflib.qureg(n)
    gpu.set_state(q, base)
    reflib.set_state(r, base)
    for lib, reg in ((gpu, q), (reflib, r)):
        lib.setAmps(reg, 77, ptr(re), ptr(im), 40)
        lib.setAmps(reg, 0, ptr(re), ptr(im), 3)
    want = reflib.state(r)
    assert np.array_equal(gpu.state(q), _round(want, prec))
    with pytest.raises(QuESTError, match="More amplitudes given than exist"):
        gpu.setAmps(q, 250, ptr(re), ptr(im), 10)
    gpu.destroy(q)
    reflib.destroy(r)


@pytest.mark.parametrize("prec", [QMX_PREC_FP64, QMX_PREC_FP32])
def test_set_density_amps_column_major_matches_reference(gpu, reflib, prec):
    """setDensityAmps starts at flat index startRow + startCol 2^n and fills
    consecutive flat entries (down the column, then into the next)."""
    n = 4
    dim = 1 << n
    vals = random_state(2 * n, 8)[:37]
    re = np.ascontiguousarray(vals.real)
    im = np.ascontiguousarray(vals.imag)
    ptr = lambda a: a.ctypes.data_as(C.POINTER(qreal))
    q, r = gpu.qureg(n, density=True, precision=prec), reflib.qureg(n, density=True)
    for lib, reg in ((gpu, q), (reflib, r)):
        lib.initPlusState(reg)
        lib.setDensityAmps(reg, 3, 5, ptr(re), ptr(im), 37)
        lib.setDensityAmps(reg, dim - 1, dim - 1, ptr(re), ptr(im), 1)
    want = reflib.state(r)
    assert np.array_equal(gpu.state(q), _round(want, prec))
    for row, col in ((3, 5), (4, 7), (0, 8), (dim - 1, dim - 1), (2, 2)):
        z, w = gpu.getDensityAmp(q, row, col), reflib.getDensityAmp(r, row, col)
        assert complex(z.real, z.imag) == _round(np.array([complex(w.real, w.imag)]), prec)[0]
    gpu.destroy(q)
    reflib.destroy(r)


def test_uploads_order_after_pending_density_work(gpu, reflib):
    """Density operations on <= 2 qubits are composed lazily; an upload must see
    them applied first (partial uploads keep their effect outside the range),
    and operations after the upload act on the uploaded values."""
    n = 5
    dim = 1 << n
    vals = random_state(2 * n, 9)[:100]
    re = np.ascontiguousarray(vals.real)
    im = np.ascontiguousarray(vals.imag)
    ptr = lambda a: a.ctypes.data_as(C.POINTER(qreal))
    q, r = gpu.qureg(n, density=True), reflib.qureg(n, density=True)
    for lib, reg in ((gpu, q), (reflib, r)):
        lib.initPlusState(reg)
        lib.hadamard(reg, 0)
        lib.mixDephasing(reg, 1, 0.2)
        lib.controlledNot(reg, 1, 3)
        lib.setDensityAmps(reg, 2, 9, ptr(re), ptr(im), 100)
        lib.rotateY(reg, 2, 0.4)
        lib.mixDamping(reg, 2, 0.1)
    assert np.max(np.abs(gpu.state(q) - reflib.state(r))) < 1e-12
    # full upload through the mirror with work pending: the upload wins
    full = random_state(2 * n, 10)
    gpu.hadamard(q, 4)
    gpu.mixDepolarising(q, 4, 0.3)
    mre, mim = mirror(q)
    mre[:] = full.real
    mim[:] = full.imag
    gpu.copyStateToGPU(q)
    assert np.array_equal(gpu.state(q), full)
    # partial upload through the mirror with work pending: the pending work is
    # applied first, then the range is overwritten, then later work follows
    gpu.initPlusState(q)
    reflib.initPlusState(r)
    for lib, reg in ((gpu, q), (reflib, r)):
        lib.pauliY(reg, 3)
        lib.mixDepolarising(reg, 0, 0.1)
    # (QuEST validates the start index against 2^n even for density registers,
    # QuEST_validation.c validateNumAmps -> validateAmpIndex)
    mre, mim = mirror(q)
    rre, rim = mirror(r)
    mre[20:90] = full.real[20:90]
    mim[20:90] = full.imag[20:90]
    gpu.copySubstateToGPU(q, 20, 70)
    rre[20:90] = full.real[20:90]  # QuEST's CPU state is its mirror
    rim[20:90] = full.imag[20:90]
    for lib, reg in ((gpu, q), (reflib, r)):
        lib.hadamard(reg, 2)
        lib.mixDamping(reg, 1, 0.2)
    assert np.max(np.abs(gpu.state(q) - reflib.state(r))) < 1e-12
    gpu.destroy(q)
    reflib.destroy(r)


@pytest.mark.parametrize("density", [False, True])
def test_report_state_csv_matches_reference(gpu, reflib, tmp_path, monkeypatch, density):
    """QuEST_common.c:216-232: the host mirror as 'real, imag' CSV."""
    n = 3 if density else 6
    q, r = gpu.qureg(n, density=density), reflib.qureg(n, density=density)
    # the same FP64 values in both mirrors (an upload is exact), so the files
    # must agree byte for byte: sign, padding, 14 decimals, separators
    vals = random_state(2 * n if density else n, 12) * 3.0
    vals[:4] = [0.0, -0.0, 1.0, -1e-15]
    for lib, reg in ((gpu, q), (reflib, r)):
        lib.set_state(reg, vals)
        lib.copyStateFromGPU(reg)
    files = []
    for lib, reg, sub in ((gpu, q, "gpu"), (reflib, r, "ref")):
        d = tmp_path / sub
        d.mkdir()
        monkeypatch.chdir(d)
        lib.reportState(reg)
        files.append((d / "state_rank_0.csv").read_bytes())
    assert files[0] == files[1]
    assert files[0].startswith(b"real, imag\n")
    gpu.destroy(q)
    reflib.destroy(r)


@pytest.mark.parametrize("prec", [QMX_PREC_FP64, QMX_PREC_FP32])
def test_wavefunction_csv_round_trip(gpu, tmp_path, prec):
    """SaveWaveFunction's '%f, %f' lines (Fermion1D.cpp:13-31) load back to the
    same values rounded to 6 decimals; malformed files and wrong sizes are errors."""
    n = 7
    psi = random_state(n, 11)
    q = gpu.qureg(n, precision=prec)
    gpu.set_state(q, psi)
    path = str(tmp_path / "wf.csv")
    assert gpu.qmxSaveStateCSV(q, path.encode()) == 1
    lines = open(path).read().splitlines()
    assert len(lines) == 1 << n
    st = _round(psi, prec)
    assert lines[3] == "%f, %f" % (st[3].real, st[3].imag)
    q2 = gpu.qureg(n)
    assert gpu.qmxLoadStateCSV(q2, path.encode()) == 1 << n
    want = np.array([complex(*map(float, l.split(","))) for l in lines])
    assert np.array_equal(gpu.state(q2), want)
    assert np.max(np.abs(want.real - st.real)) <= 5e-7 + 1e-12
    assert np.max(np.abs(want.imag - st.imag)) <= 5e-7 + 1e-12
    with open(path, "a") as f:
        f.write("0.5, 0.5\n")
    with pytest.raises(QuESTError, match="number of amplitudes"):
        gpu.qmxLoadStateCSV(q2, path.encode())
    with open(path, "w") as f:
        f.write("0.5; 0.5\n" * (1 << n))
    with pytest.raises(QuESTError, match="two numbers separated by a comma"):
        gpu.qmxLoadStateCSV(q2, path.encode())
    with open(path, "w") as f:
        f.write(" 0.5 ,\t-0.25 \r\n" * (1 << n))
    assert gpu.qmxLoadStateCSV(q2, path.encode()) == 1 << n
    assert np.all(gpu.state(q2) == 0.5 - 0.25j)
    gpu.destroy(q)
    gpu.destroy(q2)
