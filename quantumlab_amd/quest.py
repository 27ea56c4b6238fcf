"""ctypes binding of the QuEST 3.4 C API, for either qreal (double or float).

The same binding drives any shared library that exports QuEST's C ABI with the
struct layouts of include/QuEST.h (Code/QuEST340/QuEST/include/QuEST.h:94-400):
the MI355X library ``libquest_mi355x.so`` (the product), and, in tests only,
CPU libraries used as checkers. Method names are the QuEST API names so tests
read like QuEST's own tests (QuEST340/tests/test_*.cpp).

The struct layouts and signatures depend on qreal (QuEST_PREC 2: double,
QuEST_PREC 1: float); `abi(qreal)` builds them once per type, and a QuEST
instance uses the set matching its library (libquest_mi355x.so: double,
libquest_mi355x_f32.so: float). The module-level names are the double set.

Nothing here imports torch: the product library owns its HIP runtime, and
loading a second HIP runtime into the process is avoided on purpose.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Iterable, Optional, Sequence

import numpy as np

import types

PAULI_I, PAULI_X, PAULI_Y, PAULI_Z = 0, 1, 2, 3
QMX_PREC_FP64, QMX_PREC_FP32 = 0, 1
QMX_UNIFORM_RY, QMX_UNIFORM_RZ, QMX_UNIFORM_PHASE = 2, 3, 4


def _ints(xs: Iterable[int]):
    xs = list(xs)
    return (C.c_int * max(1, len(xs)))(*xs), len(xs)


def _make_abi(qreal) -> types.SimpleNamespace:
    """Struct types and signature tables of the C ABI for one qreal type."""
    class ComplexArray(C.Structure):
        _fields_ = [("real", C.POINTER(qreal)), ("imag", C.POINTER(qreal))]

    class QASMLogger(C.Structure):
        _fields_ = [("buffer", C.c_char_p), ("bufferSize", C.c_int), ("bufferFill", C.c_int),
                    ("isLogging", C.c_int)]

    class DiagonalOp(C.Structure):
        _fields_ = [("numQubits", C.c_int), ("numElemsPerChunk", C.c_longlong), ("numChunks", C.c_int),
                    ("chunkId", C.c_int), ("real", C.POINTER(qreal)), ("imag", C.POINTER(qreal)),
                    ("deviceOperator", ComplexArray)]

    class Qureg(C.Structure):
        _fields_ = [
            ("isDensityMatrix", C.c_int),
            ("numQubitsRepresented", C.c_int),
            ("numQubitsInStateVec", C.c_int),
            ("numAmpsPerChunk", C.c_longlong),
            ("numAmpsTotal", C.c_longlong),
            ("chunkId", C.c_int),
            ("numChunks", C.c_int),
            ("stateVec", ComplexArray),
            ("pairStateVec", ComplexArray),
            ("deviceStateVec", ComplexArray),
            ("firstLevelReduction", C.POINTER(qreal)),
            ("secondLevelReduction", C.POINTER(qreal)),
            ("qasmLog", C.POINTER(QASMLogger)),
        ]

    class QuESTEnv(C.Structure):
        _fields_ = [("rank", C.c_int), ("numRanks", C.c_int), ("seeds", C.POINTER(C.c_ulong)),
                    ("numSeeds", C.c_int)]

    class Complex(C.Structure):
        _fields_ = [("real", qreal), ("imag", qreal)]

    class ComplexMatrix2(C.Structure):
        _fields_ = [("real", (qreal * 2) * 2), ("imag", (qreal * 2) * 2)]

    class ComplexMatrix4(C.Structure):
        _fields_ = [("real", (qreal * 4) * 4), ("imag", (qreal * 4) * 4)]

    class ComplexMatrixN(C.Structure):
        _fields_ = [("numQubits", C.c_int), ("real", C.POINTER(C.POINTER(qreal))),
                    ("imag", C.POINTER(C.POINTER(qreal)))]

    class Vector(C.Structure):
        _fields_ = [("x", qreal), ("y", qreal), ("z", qreal)]

    class PauliHamil(C.Structure):
        _fields_ = [("pauliCodes", C.POINTER(C.c_int)), ("termCoeffs", C.POINTER(qreal)),
                    ("numSumTerms", C.c_int), ("numQubits", C.c_int)]

    class BasicOperation(C.Structure):
        """qmxBasicOperation (include/quest_mi355x.h): one QuantumLab SBasicOperation."""
        _fields_ = [("operation", C.c_int), ("numQubits", C.c_int), ("qubits", C.c_int * 2), ("param", qreal)]

    class NoiseAfterGate(C.Structure):
        """qmxNoiseAfterGate (include/quest_mi355x.h)."""
        _fields_ = [("damping", qreal), ("dephasing", qreal), ("depolarising", qreal), ("mixPauli", qreal),
                    ("twoQubitDephasing", qreal), ("twoQubitDepolarising", qreal)]

    def cplx(z) -> Complex:
        z = complex(z)
        return Complex(z.real, z.imag)

    def matrix2(u) -> ComplexMatrix2:
        u = np.asarray(u, dtype=np.complex128)
        m = ComplexMatrix2()
        for r in range(2):
            for c in range(2):
                m.real[r][c] = u[r, c].real
                m.imag[r][c] = u[r, c].imag
        return m

    def matrix4(u) -> ComplexMatrix4:
        u = np.asarray(u, dtype=np.complex128)
        m = ComplexMatrix4()
        for r in range(4):
            for c in range(4):
                m.real[r][c] = u[r, c].real
                m.imag[r][c] = u[r, c].imag
        return m

    SIGS = {
        "createQuESTEnv": (QuESTEnv, []),
        "destroyQuESTEnv": (None, [QuESTEnv]),
        "syncQuESTEnv": (None, [QuESTEnv]),
        "seedQuEST": (None, [C.POINTER(QuESTEnv), C.POINTER(C.c_ulong), C.c_int]),
        "createQureg": (Qureg, [C.c_int, QuESTEnv]),
        "createDensityQureg": (Qureg, [C.c_int, QuESTEnv]),
        "createCloneQureg": (Qureg, [Qureg, QuESTEnv]),
        "destroyQureg": (None, [Qureg, QuESTEnv]),
        "cloneQureg": (None, [Qureg, Qureg]),
        "createComplexMatrixN": (ComplexMatrixN, [C.c_int]),
        "destroyComplexMatrixN": (None, [ComplexMatrixN]),
        "createPauliHamil": (PauliHamil, [C.c_int, C.c_int]),
        "destroyPauliHamil": (None, [PauliHamil]),
        "initBlankState": (None, [Qureg]),
        "initZeroState": (None, [Qureg]),
        "initPlusState": (None, [Qureg]),
        "initClassicalState": (None, [Qureg, C.c_longlong]),
        "initPureState": (None, [Qureg, Qureg]),
        "initDebugState": (None, [Qureg]),
        "initStateFromAmps": (None, [Qureg, C.POINTER(qreal), C.POINTER(qreal)]),
        "copyStateToGPU": (None, [Qureg]),
        "copyStateFromGPU": (None, [Qureg]),
        "copySubstateToGPU": (None, [Qureg, C.c_longlong, C.c_longlong]),
        "copySubstateFromGPU": (None, [Qureg, C.c_longlong, C.c_longlong]),
        "getAmp": (Complex, [Qureg, C.c_longlong]),
        "getRealAmp": (qreal, [Qureg, C.c_longlong]),
        "getImagAmp": (qreal, [Qureg, C.c_longlong]),
        "compactUnitary": (None, [Qureg, C.c_int, Complex, Complex]),
        "unitary": (None, [Qureg, C.c_int, ComplexMatrix2]),
        "controlledUnitary": (None, [Qureg, C.c_int, C.c_int, ComplexMatrix2]),
        "multiControlledUnitary": (None, [Qureg, C.POINTER(C.c_int), C.c_int, C.c_int, ComplexMatrix2]),
        "multiStateControlledUnitary": (None, [Qureg, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int, C.c_int,
                                               ComplexMatrix2]),
        "controlledCompactUnitary": (None, [Qureg, C.c_int, C.c_int, Complex, Complex]),
        "rotateX": (None, [Qureg, C.c_int, qreal]),
        "rotateY": (None, [Qureg, C.c_int, qreal]),
        "rotateZ": (None, [Qureg, C.c_int, qreal]),
        "rotateAroundAxis": (None, [Qureg, C.c_int, qreal, Vector]),
        "controlledRotateX": (None, [Qureg, C.c_int, C.c_int, qreal]),
        "controlledRotateY": (None, [Qureg, C.c_int, C.c_int, qreal]),
        "controlledRotateZ": (None, [Qureg, C.c_int, C.c_int, qreal]),
        "hadamard": (None, [Qureg, C.c_int]),
        "pauliX": (None, [Qureg, C.c_int]),
        "pauliY": (None, [Qureg, C.c_int]),
        "pauliZ": (None, [Qureg, C.c_int]),
        "sGate": (None, [Qureg, C.c_int]),
        "tGate": (None, [Qureg, C.c_int]),
        "phaseShift": (None, [Qureg, C.c_int, qreal]),
        "controlledPhaseShift": (None, [Qureg, C.c_int, C.c_int, qreal]),
        "controlledPhaseFlip": (None, [Qureg, C.c_int, C.c_int]),
        "multiControlledPhaseFlip": (None, [Qureg, C.POINTER(C.c_int), C.c_int]),
        "multiControlledPhaseShift": (None, [Qureg, C.POINTER(C.c_int), C.c_int, qreal]),
        "controlledNot": (None, [Qureg, C.c_int, C.c_int]),
        "multiControlledMultiQubitNot": (None, [Qureg, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int), C.c_int]),
        "multiQubitNot": (None, [Qureg, C.POINTER(C.c_int), C.c_int]),
        "swapGate": (None, [Qureg, C.c_int, C.c_int]),
        "twoQubitUnitary": (None, [Qureg, C.c_int, C.c_int, ComplexMatrix4]),
        "multiQubitUnitary": (None, [Qureg, C.POINTER(C.c_int), C.c_int, ComplexMatrixN]),
        "controlledMultiQubitUnitary": (None, [Qureg, C.c_int, C.POINTER(C.c_int), C.c_int, ComplexMatrixN]),
        "multiControlledMultiQubitUnitary": (None, [Qureg, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int), C.c_int,
                                                    ComplexMatrixN]),
        "multiRotateZ": (None, [Qureg, C.POINTER(C.c_int), C.c_int, qreal]),
        "multiRotatePauli": (None, [Qureg, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int, qreal]),
        "applyTrotterCircuit": (None, [Qureg, PauliHamil, qreal, C.c_int, C.c_int]),
        "createDiagonalOp": (DiagonalOp, [C.c_int, QuESTEnv]),
        "destroyDiagonalOp": (None, [DiagonalOp, QuESTEnv]),
        "syncDiagonalOp": (None, [DiagonalOp]),
        "initDiagonalOp": (None, [DiagonalOp, C.POINTER(qreal), C.POINTER(qreal)]),
        "setDiagonalOpElems": (None, [DiagonalOp, C.c_longlong, C.POINTER(qreal), C.POINTER(qreal), C.c_longlong]),
        "initDiagonalOpFromPauliHamil": (None, [DiagonalOp, PauliHamil]),
        "createDiagonalOpFromPauliHamilFile": (DiagonalOp, [C.c_char_p, QuESTEnv]),
        "applyDiagonalOp": (None, [Qureg, DiagonalOp]),
        "calcExpecDiagonalOp": (Complex, [Qureg, DiagonalOp]),
        "applyPhaseFunc": (None, [Qureg, C.POINTER(C.c_int), C.c_int, C.c_int, C.POINTER(qreal), C.POINTER(qreal),
                                  C.c_int]),
        "applyPhaseFuncOverrides": (None, [Qureg, C.POINTER(C.c_int), C.c_int, C.c_int, C.POINTER(qreal),
                                           C.POINTER(qreal), C.c_int, C.POINTER(C.c_longlong), C.POINTER(qreal),
                                           C.c_int]),
        "applyMultiVarPhaseFunc": (None, [Qureg, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int, C.c_int,
                                          C.POINTER(qreal), C.POINTER(qreal), C.POINTER(C.c_int)]),
        "applyMultiVarPhaseFuncOverrides": (None, [Qureg, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int, C.c_int,
                                                   C.POINTER(qreal), C.POINTER(qreal), C.POINTER(C.c_int),
                                                   C.POINTER(C.c_longlong), C.POINTER(qreal), C.c_int]),
        "applyNamedPhaseFunc": (None, [Qureg, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int, C.c_int, C.c_int]),
        "applyNamedPhaseFuncOverrides": (None, [Qureg, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int, C.c_int,
                                                C.c_int, C.POINTER(C.c_longlong), C.POINTER(qreal), C.c_int]),
        "applyParamNamedPhaseFunc": (None, [Qureg, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int, C.c_int, C.c_int,
                                            C.POINTER(qreal), C.c_int]),
        "applyParamNamedPhaseFuncOverrides": (None, [Qureg, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int, C.c_int,
                                                     C.c_int, C.POINTER(qreal), C.c_int, C.POINTER(C.c_longlong),
                                                     C.POINTER(qreal), C.c_int]),
        "applyQFT": (None, [Qureg, C.POINTER(C.c_int), C.c_int]),
        "applyFullQFT": (None, [Qureg]),
        "calcTotalProb": (qreal, [Qureg]),
        "calcProbOfOutcome": (qreal, [Qureg, C.c_int, C.c_int]),
        "calcProbOfAllOutcomes": (None, [C.POINTER(qreal), Qureg, C.POINTER(C.c_int), C.c_int]),
        "calcInnerProduct": (Complex, [Qureg, Qureg]),
        "calcDensityInnerProduct": (qreal, [Qureg, Qureg]),
        "calcPurity": (qreal, [Qureg]),
        "calcFidelity": (qreal, [Qureg, Qureg]),
        "calcHilbertSchmidtDistance": (qreal, [Qureg, Qureg]),
        "collapseToOutcome": (qreal, [Qureg, C.c_int, C.c_int]),
        "measure": (C.c_int, [Qureg, C.c_int]),
        "measureWithStats": (C.c_int, [Qureg, C.c_int, C.POINTER(qreal)]),
        "mixDephasing": (None, [Qureg, C.c_int, qreal]),
        "mixTwoQubitDephasing": (None, [Qureg, C.c_int, C.c_int, qreal]),
        "mixDepolarising": (None, [Qureg, C.c_int, qreal]),
        "mixDamping": (None, [Qureg, C.c_int, qreal]),
        "mixPauli": (None, [Qureg, C.c_int, qreal, qreal, qreal]),
        "mixKrausMap": (None, [Qureg, C.c_int, C.POINTER(ComplexMatrix2), C.c_int]),
        "mixDensityMatrix": (None, [Qureg, qreal, Qureg]),
        "mixTwoQubitDepolarising": (None, [Qureg, C.c_int, C.c_int, qreal]),
        "mixTwoQubitKrausMap": (None, [Qureg, C.c_int, C.c_int, C.POINTER(ComplexMatrix4), C.c_int]),
        "mixMultiQubitKrausMap": (None, [Qureg, C.POINTER(C.c_int), C.c_int, C.POINTER(ComplexMatrixN), C.c_int]),
        "mixNonTPKrausMap": (None, [Qureg, C.c_int, C.POINTER(ComplexMatrix2), C.c_int]),
        "calcExpecPauliProd": (qreal, [Qureg, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int, Qureg]),
        "calcExpecPauliSum": (qreal, [Qureg, C.POINTER(C.c_int), C.POINTER(qreal), C.c_int, Qureg]),
        "calcExpecPauliHamil": (qreal, [Qureg, PauliHamil, Qureg]),
        "applyPauliSum": (None, [Qureg, C.POINTER(C.c_int), C.POINTER(qreal), C.c_int, Qureg]),
        "applyPauliHamil": (None, [Qureg, PauliHamil, Qureg]),
        "setWeightedQureg": (None, [Complex, Qureg, Complex, Qureg, Complex, Qureg]),
        "sqrtSwapGate": (None, [Qureg, C.c_int, C.c_int]),
        "controlledTwoQubitUnitary": (None, [Qureg, C.c_int, C.c_int, C.c_int, ComplexMatrix4]),
        "multiControlledTwoQubitUnitary": (None, [Qureg, C.POINTER(C.c_int), C.c_int, C.c_int, C.c_int,
                                                  ComplexMatrix4]),
        "applyMatrix2": (None, [Qureg, C.c_int, ComplexMatrix2]),
        "applyMatrix4": (None, [Qureg, C.c_int, C.c_int, ComplexMatrix4]),
        "applyMatrixN": (None, [Qureg, C.POINTER(C.c_int), C.c_int, ComplexMatrixN]),
        "applyProjector": (None, [Qureg, C.c_int, C.c_int]),
        "getProbAmp": (qreal, [Qureg, C.c_longlong]),
        "getDensityAmp": (Complex, [Qureg, C.c_longlong, C.c_longlong]),
        "getNumQubits": (C.c_int, [Qureg]),
        "getNumAmps": (C.c_longlong, [Qureg]),
        "setAmps": (None, [Qureg, C.c_longlong, C.POINTER(qreal), C.POINTER(qreal), C.c_longlong]),
        "setDensityAmps": (None, [Qureg, C.c_longlong, C.c_longlong, C.POINTER(qreal), C.POINTER(qreal),
                                  C.c_longlong]),
        "multiControlledMultiRotateZ": (None, [Qureg, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int), C.c_int,
                                               qreal]),
        "multiControlledMultiRotatePauli": (None, [Qureg, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int),
                                                   C.POINTER(C.c_int), C.c_int, qreal]),
        "controlledRotateAroundAxis": (None, [Qureg, C.c_int, C.c_int, qreal, Vector]),
        "controlledPauliY": (None, [Qureg, C.c_int, C.c_int]),
    }

    # extension entry points of libquest_mi355x (include/quest_mi355x.h)
    EXT_SIGS = {
        "createQuregEx": (Qureg, [C.c_int, QuESTEnv, C.c_int]),
        "createDensityQuregEx": (Qureg, [C.c_int, QuESTEnv, C.c_int]),
        "qmxQuregPrecision": (C.c_int, [Qureg]),
        "qmxSetErrorMode": (None, [C.c_int]),
        "qmxLastError": (C.c_int, [C.c_char_p, C.c_int]),
        "qmxClearError": (None, []),
        "qmxDeviceInfo": (C.c_int, [C.c_char_p, C.c_int]),
        "qmxSync": (None, []),
        "qmxDensityFusionStats": (None, [C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
        "qmxSampleShots": (C.c_int, [Qureg, C.POINTER(C.c_int), C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)]),
        "qmxAmplitudeEncodeReal": (None, [Qureg, C.POINTER(qreal), C.c_int]),
        "qmxAmplitudeEncodeComplex": (None, [Qureg, C.POINTER(qreal), C.POINTER(qreal), C.c_int]),
        "qmxUniformlyControlledUnitary": (None, [Qureg, C.POINTER(C.c_int), C.c_int, C.c_int,
                                                 C.POINTER(ComplexMatrix2)]),
        "qmxUniformlyControlledRotation": (None, [Qureg, C.POINTER(C.c_int), C.c_int, C.c_int, C.POINTER(qreal),
                                                  C.c_int]),
        "qmxTimerBegin": (None, []),
        "qmxTimerEndMs": (C.c_double, []),
        "qmxTimerSlotBegin": (None, [C.c_int]),
        "qmxTimerSlotEnd": (None, [C.c_int]),
        "qmxTimerSlotMs": (C.c_double, [C.c_int]),
        "qmxMaxAbsDiff": (qreal, [Qureg, Qureg]),
        "qmxRenormalise": (qreal, [Qureg, C.c_int]),
        "qmxCalcProbOfAllOutcomesNoAlloc": (None, [C.POINTER(qreal), Qureg, C.POINTER(C.c_int), C.c_int]),
        "qmxSaveStateCSV": (C.c_int, [Qureg, C.c_char_p]),
        "qmxLoadStateCSV": (C.c_longlong, [Qureg, C.c_char_p]),
        "reportState": (None, [Qureg]),
        "qmxCircuitCreate": (C.c_void_p, [C.c_int]),
        "qmxCircuitDestroy": (None, [C.c_void_p]),
        "qmxCircuitNumGates": (C.c_int, [C.c_void_p]),
        "qmxCircuitCompactUnitary": (None, [C.c_void_p, C.c_int, Complex, Complex]),
        "qmxCircuitUnitary": (None, [C.c_void_p, C.c_int, ComplexMatrix2]),
        "qmxCircuitRotateZ": (None, [C.c_void_p, C.c_int, qreal]),
        "qmxCircuitControlledNot": (None, [C.c_void_p, C.c_int, C.c_int]),
        "qmxCircuitMultiControlledUnitary": (None, [C.c_void_p, C.POINTER(C.c_int), C.c_int, C.c_int, ComplexMatrix2]),
        "qmxCircuitMultiQubitUnitary": (None, [C.c_void_p, C.POINTER(C.c_int), C.c_int, C.POINTER(qreal),
                                               C.POINTER(qreal)]),
        "qmxCircuitFuse": (C.c_int, [C.c_void_p, C.c_int]),
        "qmxCircuitNumBlocks": (C.c_int, [C.c_void_p]),
        "qmxCircuitStep": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(qreal), C.POINTER(qreal)]),
        "qmxCircuitApply": (qreal, [Qureg, C.c_void_p, C.c_int]),
        "qmxCircuitAddOperation": (None, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.c_int, qreal]),
        "qmxCircuitSetNoiseAfterGate": (None, [C.c_void_p, NoiseAfterGate]),
        "qmxApplyOperations": (qreal, [Qureg, C.POINTER(BasicOperation), C.c_int, C.POINTER(NoiseAfterGate), C.c_int]),
        "qmxSwapPlanCreate": (C.c_void_p, [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double), C.c_int, C.c_int,
                                           C.POINTER(C.c_double), C.c_int]),
        "qmxSwapPlanDestroy": (None, [C.c_void_p]),
        "qmxSwapPlanAutotune": (C.c_int, [C.c_void_p]),
        "qmxSwapPlanRowsPerTile": (C.c_int, [C.c_void_p, C.c_int]),
        "qmxSwapPlanNumTiles": (C.c_int, [C.c_void_p]),
        "qmxSwapPlanMfmaBlocks": (C.c_longlong, [C.c_void_p]),
        "qmxSwapPlanTimeTiles": (C.c_double, [C.c_void_p, C.c_int]),
        "qmxSwapPlanRun": (None, [C.c_void_p, C.c_int, C.c_int]),
        "qmxSwapPlanP0": (None, [C.c_void_p, C.POINTER(C.c_double)]),
        "qmxSwapPlanAssignments": (None, [C.c_void_p, C.POINTER(C.c_int)]),
        "qmxSwapPlanCentroids": (None, [C.c_void_p, C.POINTER(C.c_double)]),
        "qmxSwapPlanPointState": (None, [C.c_void_p, C.c_int, C.POINTER(C.c_double)]),
        "qmxSwapPlanSetCentroids": (None, [C.c_void_p, C.POINTER(C.c_double)]),
        "qmxTrotterBatchEvolve": (C.c_int, [C.c_int, C.POINTER(PauliHamil), C.POINTER(qreal), C.POINTER(qreal),
                                            C.POINTER(qreal), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(qreal),
                                            C.POINTER(qreal)]),
        "qmxTrotterBatchInfidelity": (C.c_int, [C.c_int, C.POINTER(PauliHamil), C.POINTER(qreal), C.POINTER(qreal),
                                                C.POINTER(qreal), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                                C.POINTER(qreal)]),
        "qmxTrotterBatchLadder": (C.c_int, [C.c_int, C.POINTER(PauliHamil), C.POINTER(qreal), C.POINTER(qreal),
                                            C.POINTER(qreal), C.POINTER(C.c_int), C.c_int, qreal, C.POINTER(qreal),
                                            C.POINTER(C.c_int)]),
        "qmxTrotterFeatures": (C.c_int, [PauliHamil, qreal, C.POINTER(C.c_float)]),
        "qmxTrotterStages": (C.c_int, [C.c_int]),
        "qmxTrotterSweepsPerRep": (C.c_int, [PauliHamil, C.c_int, C.c_int, C.c_int]),
        "qmxPlannerCreate": (C.c_void_p, [C.c_uint]),
        "qmxPlannerDestroy": (None, [C.c_void_p]),
        "qmxPlannerNumParams": (C.c_int, []),
        "qmxPlannerGetParams": (None, [C.c_void_p, C.POINTER(C.c_float)]),
        "qmxPlannerSetParams": (None, [C.c_void_p, C.POINTER(C.c_float)]),
        "qmxPlannerSetNormalisation": (None, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
        "qmxPlannerGetNormalisation": (None, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
        "qmxPlannerSetMargin": (None, [C.c_void_p, C.c_float]),
        "qmxPlannerGetMargin": (C.c_float, [C.c_void_p]),
        "qmxPlannerSetSizeCurve": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
        "qmxPlannerGetSizeCurve": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
        "qmxPlannerSizeCurve": (C.c_double, [C.c_void_p, C.c_int, C.c_int]),
        "qmxPlannerPredict": (None, [C.c_void_p, C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_float)]),
        "qmxPlannerGradient": (C.c_double, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                            C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_float)]),
        "qmxPlannerTrain": (C.c_double, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                         C.c_int, C.c_int, C.c_int, C.c_float, C.c_uint, C.POINTER(C.c_float)]),
        "qmxPlannerChoose": (C.c_int, [C.c_void_p, PauliHamil, qreal, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "qmxApplyTrotterLearned": (C.c_int, [Qureg, PauliHamil, qreal, C.c_void_p, C.POINTER(C.c_int),
                                             C.POINTER(C.c_int)]),
        "qmxApplyTrotterStepDoubling": (C.c_int, [Qureg, PauliHamil, qreal, C.c_int, qreal, C.c_int,
                                                  C.POINTER(C.c_int)]),
        "qmxPlannerSave": (C.c_int, [C.c_void_p, C.c_char_p]),
        "qmxPlannerLoad": (C.c_void_p, [C.c_char_p]),
        "qmxMtInitByArray": (None, [C.POINTER(C.c_ulong), C.c_int]),
        "qmxMtNextU32": (C.c_uint, []),
        "qmxMtNextReal1": (C.c_double, []),
    }

    return types.SimpleNamespace(**{k: v for k, v in locals().items() if not k.startswith("_")})


_ABIS = {}


def abi(qreal=C.c_double) -> types.SimpleNamespace:
    """The ABI namespace for qreal = ctypes.c_double or ctypes.c_float."""
    key = C.sizeof(qreal)
    if key not in _ABIS:
        _ABIS[key] = _make_abi(qreal)
    return _ABIS[key]


_A = abi(C.c_double)
qreal = _A.qreal
ComplexArray, QASMLogger, DiagonalOp, Qureg, QuESTEnv = _A.ComplexArray, _A.QASMLogger, _A.DiagonalOp, _A.Qureg, _A.QuESTEnv
Complex, ComplexMatrix2, ComplexMatrix4, ComplexMatrixN = _A.Complex, _A.ComplexMatrix2, _A.ComplexMatrix4, _A.ComplexMatrixN
Vector, PauliHamil, BasicOperation, NoiseAfterGate = _A.Vector, _A.PauliHamil, _A.BasicOperation, _A.NoiseAfterGate
cplx, matrix2, matrix4 = _A.cplx, _A.matrix2, _A.matrix4
_SIGS, _EXT_SIGS = _A.SIGS, _A.EXT_SIGS


class QuESTError(RuntimeError):
    """An input error reported through invalidQuESTInputError."""


class QuEST:
    """One loaded QuEST-ABI library plus its environment."""

    def __init__(self, path: str, extensions: bool = False, seeds: Sequence[int] = (12345, 6789),
                 qreal=C.c_double):
        self.path = path
        self.lib = C.CDLL(path, mode=os.RTLD_LOCAL)
        self.ext = extensions
        self.t = abi(qreal)  # struct types and signatures of this library's qreal
        self.np_real = np.float32 if C.sizeof(qreal) == 4 else np.float64
        for name, (res, args) in self.t.SIGS.items():
            fn = getattr(self.lib, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        if extensions:
            for name, (res, args) in self.t.EXT_SIGS.items():
                fn = getattr(self.lib, name)
                fn.restype = res
                fn.argtypes = args
            self.lib.qmxSetErrorMode(1)
        self.env = self.lib.createQuESTEnv()
        self.seed(seeds)

    # -- environment ------------------------------------------------------
    def seed(self, seeds: Sequence[int]):
        arr = (C.c_ulong * len(seeds))(*seeds)
        self.lib.seedQuEST(C.byref(self.env), arr, len(seeds))

    def check(self):
        """Raise QuESTError if the extension library recorded an input error."""
        if not self.ext:
            return
        buf = C.create_string_buffer(1024)
        if self.lib.qmxLastError(buf, 1024):
            self.lib.qmxClearError()
            raise QuESTError(buf.value.decode())

    def call(self, name: str, *args):
        r = getattr(self.lib, name)(*args)
        self.check()
        return r

    def __getattr__(self, name):
        fn = getattr(self.lib, name)

        def wrapped(*args):
            r = fn(*args)
            self.check()
            return r

        return wrapped

    # -- registers ---------------------------------------------------------
    def qureg(self, n: int, density: bool = False, precision: Optional[int] = None) -> Qureg:
        if precision is not None:
            q = (self.lib.createDensityQuregEx if density else self.lib.createQuregEx)(n, self.env, precision)
        else:
            q = (self.lib.createDensityQureg if density else self.lib.createQureg)(n, self.env)
        self.check()
        return q

    def destroy(self, q: Qureg):
        self.lib.destroyQureg(q, self.env)

    def state(self, q: Qureg) -> np.ndarray:
        """Full amplitude vector (density matrices: flat, element (r,c) at r + c*2^n)."""
        self.lib.copyStateFromGPU(q)
        n = q.numAmpsTotal
        re = np.ctypeslib.as_array(q.stateVec.real, shape=(n,)).astype(np.float64)
        im = np.ctypeslib.as_array(q.stateVec.imag, shape=(n,)).astype(np.float64)
        return re + 1j * im

    def set_state(self, q: Qureg, amps: np.ndarray):
        amps = np.ascontiguousarray(amps, dtype=np.complex128)
        re = np.ascontiguousarray(amps.real, dtype=self.np_real)
        im = np.ascontiguousarray(amps.imag, dtype=self.np_real)
        ptr = C.POINTER(self.t.qreal)
        self.lib.initStateFromAmps(q, re.ctypes.data_as(ptr), im.ctypes.data_as(ptr))

    def reals(self, x) -> np.ndarray:
        """A contiguous array of this library's qreal (pass .ctypes.data_as(self.real_ptr))."""
        return np.ascontiguousarray(x, dtype=self.np_real)

    @property
    def real_ptr(self):
        return C.POINTER(self.t.qreal)

    def density(self, q: Qureg) -> np.ndarray:
        dim = 1 << q.numQubitsRepresented
        return self.state(q).reshape(dim, dim).T  # rho[r, c] = flat[r + c*dim]

    # -- matrices ------------------------------------------------------------
    def matrixN(self, u: np.ndarray) -> ComplexMatrixN:
        u = np.asarray(u, dtype=np.complex128)
        k = int(round(np.log2(u.shape[0])))
        m = self.lib.createComplexMatrixN(k)
        for r in range(1 << k):
            for c in range(1 << k):
                m.real[r][c] = u[r, c].real
                m.imag[r][c] = u[r, c].imag
        return m

    def hamil(self, coeffs: Sequence[float], codes: np.ndarray) -> PauliHamil:
        codes = np.asarray(codes, dtype=np.int32)
        nt, nq = codes.shape
        h = self.lib.createPauliHamil(nq, nt)
        for t in range(nt):
            h.termCoeffs[t] = float(coeffs[t])
            for q in range(nq):
                h.pauliCodes[t * nq + q] = int(codes[t, q])
        return h

    # -- convenience wrappers taking Python lists ----------------------------
    def multiControlledUnitary(self, q, ctrls, target, u):
        arr, n = _ints(ctrls)
        self.call("multiControlledUnitary", q, arr, n, target, self.t.matrix2(u))

    def multiQubitUnitary(self, q, targs, u):
        arr, n = _ints(targs)
        m = self.matrixN(u)
        self.call("multiQubitUnitary", q, arr, n, m)
        self.lib.destroyComplexMatrixN(m)

    def multiRotateZ(self, q, qubits, angle):
        arr, n = _ints(qubits)
        self.call("multiRotateZ", q, arr, n, angle)

    def multiRotatePauli(self, q, targs, paulis, angle):
        t, n = _ints(targs)
        p, _ = _ints(paulis)
        self.call("multiRotatePauli", q, t, p, n, angle)

    def calcProbOfAllOutcomes(self, q, qubits):
        arr, n = _ints(qubits)
        out = (self.t.qreal * (1 << n))()
        self.call("calcProbOfAllOutcomes", out, q, arr, n)
        return np.array(out[:])

    def measureWithStats(self, q, qubit):
        p = self.t.qreal()
        o = self.call("measureWithStats", q, qubit, C.byref(p))
        return o, p.value


def library_path(flavour: str = "f64") -> str:
    """Path of the in-tree library (flavour "f64": libquest_mi355x.so, qreal =
    double; "f32": libquest_mi355x_f32.so, qreal = float); fails loudly if it is
    not built."""
    here = os.path.dirname(os.path.abspath(__file__))
    name = {"f64": "libquest_mi355x.so", "f32": "libquest_mi355x_f32.so"}[flavour]
    p = os.path.join(here, "lib", name)
    if not os.path.exists(p):
        raise FileNotFoundError(
            f"{p} is not built; run __graft_entry__.build() (there is no CPU fallback)")
    return p


def load(seeds: Sequence[int] = (12345, 6789), flavour: str = "f64") -> QuEST:
    """Load the MI355X library (GPU required)."""
    qr = C.c_float if flavour == "f32" else C.c_double
    return QuEST(library_path(flavour), extensions=True, seeds=seeds, qreal=qr)
