"""QuantumLab's job files (Bin/*.yaml) through quantumlab_amd.jobs, and the
swap-test application's job run at register level on a density register.

CPU part: the reader on text in the files' syntax (blocks, lists, comments,
values kept as strings as CYAMLParser keeps them), and on the reference's own
three job files where the reference tree exists (known values checked).

GPU part: APP/SwapTest/SwapTest.cpp's QuantumOverlap with noise, for a real
vector pair (ZeroTestReal, QL/Circuit/Gates/SwapTest.cpp:156-184: AE(v1), then
AE(v2) daggered, on q qubits): the gate list as QuantumLab operations through
qmxApplyOperations on a density register with the job's after-gate noise
(QLSimulatorDensityMatrix.cpp:88-153), the job's before-measure channels on
every qubit (:190-216), then MeasureRepeat shots of every qubit with
qmxSampleShots. The oracle runs the same loop on QuEST's C API (oracle/ops_oracle)
and the restore-and-measure shot loop; state to 1e-12, outcomes identical
under the same seeds, the all-zero frequency within binomial noise of
|<v1|v2>|^2 (damped by the noise)."""
import ctypes as C
import os

import numpy as np
import pytest

from quantumlab_amd import jobs as J

SWAP_TEST_JOB = """\
# ########################
# DotType : {EDT_AmplitudeEncodeComplex, EDT_AmplitudeEncodeReal, EDT_SimpleEncode}
FileName1 : v1.csv
FileName2 : v2.csv
MeasureRepeat : 2000
DotType : EDT_AmplitudeEncodeReal
EnableClassical : 0
EnableNoise : 1
SimpleEncodeQubits : 3

DampingAfterGate : 0.001
DepolarisingAfterGate : 0.002
DephasingAfterGate : 0.001
MixPauliAfterGate : 0.001
TwoQubitDephasing : 0.001
TwoQubitDepolarising : 0.003

DampingBeforeMeasure : 0.02
DepolarisingBeforeMeasure : 0.02
DephasingBeforeMeasure : 0.0
MixPauliBeforeMeasure : 0.02
"""

BLOCKS = """\
KMeansJob : EQKT_TestSimpleEncode   # trailing comment

GPUKmeans:

    CK : 50
    Energy : [1500]
    CFileName : SM

CircuitTestDim:

    DimList :      [32, 32, 32]
    # AARepeatList : [1, 2,  3]
    AARepeatList : [4, 4, 4]
    VectorCount : 16
Empty : []
Top : 7
"""


def test_reader_blocks_lists_comments():
    d = J.read_job(BLOCKS)
    assert d["KMeansJob"] == "EQKT_TestSimpleEncode"
    assert d["GPUKmeans"] == {"CK": "50", "Energy": ["1500"], "CFileName": "SM"}
    assert d["CircuitTestDim"]["DimList"] == ["32", "32", "32"]
    assert d["CircuitTestDim"]["AARepeatList"] == ["4", "4", "4"]
    assert d["Empty"] == [] and d["Top"] == "7"
    j = J.Job(d)
    assert j.block("GPUKmeans").int("CK") == 50
    assert j.block("CircuitTestDim").ints("DimList") == [32, 32, 32]
    assert J.is_enum_name(j.string("KMeansJob"))
    with pytest.raises(ValueError):
        J.read_job("Key : [1, 2\n")
    with pytest.raises(ValueError):
        J.read_job("no colon here\n")


def test_swap_test_job_mapping():
    job = J.swap_test_job(J.read_job(SWAP_TEST_JOB))
    assert job.dot_type == "EDT_AmplitudeEncodeReal" and job.measure_repeat == 2000 and job.enable_noise
    assert job.noise_after_gate == {"damping": 0.001, "depolarising": 0.002, "dephasing": 0.001,
                                    "mixPauli": 0.001, "twoQubitDephasing": 0.001, "twoQubitDepolarising": 0.003}
    assert job.before_measure == {"damping": 0.02, "depolarising": 0.02, "dephasing": 0.0, "mixPauli": 0.02}
    assert job.files["FileName1"] == "v1.csv"


BIN = "/root/reference/Bin"


@pytest.mark.skipif(not os.path.isdir(BIN), reason="the reference tree is not here")
def test_reference_job_files_parse():
    sw = J.swap_test_job(J.read_job_file(os.path.join(BIN, "SwapTest.yaml")))
    assert (sw.measure_repeat, sw.dot_type, sw.enable_noise, sw.simple_encode_qubits) == \
        (1000, "EDT_SimpleEncode", False, 3)
    assert sw.noise_after_gate["damping"] == 0.001 and sw.before_measure["damping"] == 0.02
    km = J.Job(J.read_job_file(os.path.join(BIN, "QKMeans.yaml")))
    assert km.string("KMeansJob") == "EQKT_TestSimpleEncode"
    assert km.block("GPUKmeans").int("CK") == 50
    assert km.block("CircuitTestDim").ints("DimList") == [32, 32, 32]
    fs = J.read_job_file(os.path.join(BIN, "FermionSimulation.yaml"))
    assert isinstance(fs, dict) and fs


# ---- the swap-test job at register level (GPU) ----------------------------------

def zero_test_real_ops(v1, v2):
    """ZeroTestReal's gate list as QuantumLab operations (code, qubits, param):
    AE(v1) on qubits 0..q-1, then AE(v2) daggered (reverse order, negated
    angles)."""
    from oracle import ops_oracle as O
    from oracle import swaptest_oracle as S
    q = len(v1).bit_length() - 1
    code = {"ry": O.RY, "rz": O.RZ}
    out = []
    for g in S.ae_gates(v1, list(range(q))):
        out.append((O.CX, [g[1], g[2]], 0.0) if g[0] == "cx" else (code[g[0]], [g[1]], g[2]))
    for g in reversed(S.ae_gates(v2, list(range(q)))):
        out.append((O.CX, [g[1], g[2]], 0.0) if g[0] == "cx" else (code[g[0]], [g[1]], -g[2]))
    return out


def before_measure_ops(n, bm):
    from oracle import ops_oracle as O
    out = []
    for code, key in ((O.NOISE_DAMPING, "damping"), (O.NOISE_DEPOLARISING, "depolarising"),
                      (O.NOISE_DEPHASING, "dephasing"), (O.NOISE_MIX_ALL, "mixPauli")):
        if bm[key] > np.finfo(np.float64).tiny:
            out += [(code, [t], bm[key]) for t in range(n)]
    return out


@pytest.mark.gpu
def test_swap_test_job_with_noise_at_register_level(gpu, cpuref):
    from oracle import ops_oracle as O
    from quantumlab_amd.quest import BasicOperation
    job = J.swap_test_job(J.read_job(SWAP_TEST_JOB))
    rng = np.random.default_rng(11)
    q = 3
    v1, v2 = rng.standard_normal(1 << q), rng.standard_normal(1 << q)
    ops = zero_test_real_ops(v1, v2)
    pre = before_measure_ops(q, job.before_measure)

    def op_array(lst):
        arr = (BasicOperation * len(lst))()
        for i, (code, qs, t) in enumerate(lst):
            arr[i].operation, arr[i].numQubits, arr[i].param = code, len(qs), t
            for k, x in enumerate(qs):
                arr[i].qubits[k] = x
        return arr

    noise = J.noise_struct(gpu, job)
    rho = gpu.qureg(q, density=True)
    gpu.qmxApplyOperations(rho, op_array(ops), len(ops), C.byref(noise), 0)   # circuit, after-gate noise
    gpu.qmxApplyOperations(rho, op_array(pre), len(pre), None, 0)             # before-measure channels
    ref = cpuref.qureg(q, density=True)
    O.run_ops(cpuref, ref, ops, job.noise_after_gate)
    O.run_ops(cpuref, ref, pre)
    assert np.max(np.abs(gpu.state(rho) - cpuref.state(ref))) < 1e-12
    # MeasureRepeat shots of every qubit, QuantumLab's loop on the oracle
    gpu.seed([2024])
    cpuref.seed([2024])
    shots = job.measure_repeat
    got = (C.c_int * shots)()
    assert gpu.qmxSampleShots(rho, (C.c_int * q)(*range(q)), q, shots, 0, got) == shots
    saved = cpuref.state(ref)
    want = []
    for s in range(shots):
        if s:
            cpuref.set_state(ref, saved)
        want.append(sum(int(cpuref.call("measure", ref, t)) << t for t in range(q)))
    assert list(got) == want
    p0 = np.mean(np.asarray(want) == 0)
    overlap = float(np.dot(v1, v2) / np.linalg.norm(v1) / np.linalg.norm(v2)) ** 2
    p_noisy = float(np.real(saved[0]))
    assert abs(p0 - p_noisy) < 5 * np.sqrt(p_noisy * (1 - p_noisy) / shots) + 1e-12
    assert abs(p_noisy - overlap) < 0.1      # the noise damps, it does not scramble
    gpu.destroy(rho)
    cpuref.destroy(ref)
