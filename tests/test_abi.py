"""The C-ABI library loads and exports every function include/*.h declares.
CPU only: loading does not touch the GPU (the HIP runtime initialises on the
first QuEST call), and no compute entry point is called here."""
import ctypes as C
import os
import re

import pytest

from tests.conftest import ROOT

HEADERS = [os.path.join(ROOT, "include", h) for h in ("QuEST.h", "quest_mi355x.h")]


def declared_functions(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    text = re.sub(r"^\s*#[^\n]*", " ", text, flags=re.M)
    names = set()
    for stmt in text.split(";"):
        if "(" not in stmt or "typedef" in stmt or "{" in stmt and "}" not in stmt.split("(")[0]:
            continue
        head = stmt.split("(")[0].strip()
        head = head.split("}")[-1].strip()
        m = re.search(r"(\w+)\s*$", head)
        if m and m.group(1) not in ("if", "for", "while", "return", "sizeof"):
            names.add(m.group(1))
    return names


def test_headers_parse_to_many_functions():
    names = set().union(*(declared_functions(h) for h in HEADERS))
    assert {"createQureg", "compactUnitary", "controlledNot", "multiControlledUnitary", "multiQubitUnitary",
            "multiRotatePauli", "mixDephasing", "mixDepolarising", "mixDamping", "calcProbOfOutcome",
            "calcInnerProduct", "calcTotalProb", "calcPurity", "applyTrotterCircuit", "createDensityQureg",
            "createQuregEx", "qmxCircuitFuse"} <= names
    assert len(names) > 120


def test_library_exports_every_declared_function():
    import quantumlab_amd
    lib = C.CDLL(quantumlab_amd.library_path(), mode=os.RTLD_LOCAL)
    missing = [n for h in HEADERS for n in sorted(declared_functions(h)) if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_weak_error_hook_is_overridable():
    out = os.popen(f"nm -D {os.path.join(ROOT, 'quantumlab_amd', 'lib', 'libquest_mi355x.so')}").read()
    line = [l for l in out.splitlines() if l.endswith(" invalidQuESTInputError")]
    assert line and line[0].split()[-2] == "W", line


def test_mt19937_host_stream_matches_reference_golden():
    """The measurement RNG is host code in the library: pin its stream against
    the reference's golden draws without touching the GPU."""
    import numpy as np

    import quantumlab_amd
    from tests.conftest import golden
    g = golden("mt19937")
    lib = C.CDLL(quantumlab_amd.library_path(), mode=os.RTLD_LOCAL)
    lib.qmxMtInitByArray.argtypes = [C.POINTER(C.c_ulong), C.c_int]
    lib.qmxMtNextU32.restype = C.c_uint
    lib.qmxMtNextReal1.restype = C.c_double
    key = (C.c_ulong * 2)(*[int(x) for x in g["key"]])
    lib.qmxMtInitByArray(key, 2)
    got = np.array([lib.qmxMtNextU32() for _ in range(len(g["u32"]))], dtype=np.uint64)
    assert np.array_equal(got, g["u32"])
    r1 = np.array([lib.qmxMtNextReal1() for _ in range(len(g["real1"]))])
    assert np.array_equal(r1, g["real1"])


@pytest.mark.parametrize("path", ["oracle/build/libcpuref.so"])
def test_oracle_library_is_separate_from_product(path):
    """The product library must not depend on the oracle."""
    out = os.popen(f"ldd {os.path.join(ROOT, 'quantumlab_amd', 'lib', 'libquest_mi355x.so')}").read()
    assert "cpuref" not in out and "quest_ref" not in out


_LAYOUT_PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include "QuEST.h"
#define S(T) printf(#T " %zu\n", sizeof(T))
#define O(T, f) printf(#T "." #f " %zu\n", offsetof(T, f))
int main(void) {
    S(Qureg); O(Qureg, isDensityMatrix); O(Qureg, numQubitsRepresented); O(Qureg, numQubitsInStateVec);
    O(Qureg, numAmpsPerChunk); O(Qureg, numAmpsTotal); O(Qureg, chunkId); O(Qureg, numChunks);
    O(Qureg, stateVec); O(Qureg, pairStateVec); O(Qureg, deviceStateVec); O(Qureg, firstLevelReduction);
    O(Qureg, secondLevelReduction); O(Qureg, qasmLog);
    S(QuESTEnv); O(QuESTEnv, rank); O(QuESTEnv, numRanks); O(QuESTEnv, seeds); O(QuESTEnv, numSeeds);
    S(Complex); S(ComplexArray); S(ComplexMatrix2); S(ComplexMatrix4); S(Vector); S(QASMLogger);
    S(ComplexMatrixN); O(ComplexMatrixN, numQubits); O(ComplexMatrixN, real); O(ComplexMatrixN, imag);
    S(PauliHamil); O(PauliHamil, pauliCodes); O(PauliHamil, termCoeffs); O(PauliHamil, numSumTerms);
    O(PauliHamil, numQubits);
    S(DiagonalOp); O(DiagonalOp, numQubits); O(DiagonalOp, numElemsPerChunk); O(DiagonalOp, numChunks);
    O(DiagonalOp, chunkId); O(DiagonalOp, real); O(DiagonalOp, imag); O(DiagonalOp, deviceOperator);
    S(qreal); printf("REAL_EPS %g\n", (double)REAL_EPS);
    printf("pauli %d %d %d %d\n", PAULI_I, PAULI_X, PAULI_Y, PAULI_Z);
    printf("gate %d %d %d\n", SIGMA_Z, S_GATE, T_GATE);
    printf("enc %d %d\n", UNSIGNED, TWOS_COMPLEMENT);
    printf("func %d %d %d %d %d %d %d %d %d %d %d %d %d %d\n", NORM, SCALED_NORM, INVERSE_NORM,
           SCALED_INVERSE_NORM, SCALED_INVERSE_SHIFTED_NORM, PRODUCT, SCALED_PRODUCT, INVERSE_PRODUCT,
           SCALED_INVERSE_PRODUCT, DISTANCE, SCALED_DISTANCE, INVERSE_DISTANCE, SCALED_INVERSE_DISTANCE,
           SCALED_INVERSE_SHIFTED_DISTANCE);
    return 0;
}
"""


@pytest.mark.parametrize("prec", [1, 2])
def test_struct_layouts_and_enums_equal_reference_header(tmp_path, prec):
    """One translation unit compiled against include/ and against the reference's
    QuEST.h prints the same sizes, field offsets and enum values."""
    import subprocess
    ref_inc = "/root/reference/Code/QuEST340/QuEST/include"
    if not os.path.isdir(ref_inc):
        pytest.skip("reference header not present")
    src = tmp_path / "probe.c"
    src.write_text(_LAYOUT_PROBE)
    outs = []
    for inc in (os.path.join(ROOT, "include"), ref_inc):
        exe = tmp_path / f"probe_{len(outs)}"
        subprocess.run(["gcc", "-std=c99", f"-DQuEST_PREC={prec}", f"-I{inc}", str(src), "-o", str(exe), "-lm"],
                       check=True)
        outs.append(subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout)
    assert outs[0] == outs[1]
    assert f"qreal {4 if prec == 1 else 8}" in outs[0]


def test_both_flavours_export_every_declared_function():
    import quantumlab_amd
    for flavour in ("f64", "f32"):
        lib = C.CDLL(quantumlab_amd.library_path(flavour), mode=os.RTLD_LOCAL)
        missing = [n for h in HEADERS for n in sorted(declared_functions(h)) if not hasattr(lib, n)]
        assert not missing, (flavour, missing)


def test_error_hook_binds_through_the_dynamic_linker():
    """Every other symbol binds inside the library, but calls to the error hook
    go through the PLT, so a caller's own invalidQuESTInputError (QuEST's test
    main, QT/main.cpp:40-42) replaces the weak default."""
    import quantumlab_amd
    for flavour in ("f64", "f32"):
        relocs = os.popen(f"readelf -rW {quantumlab_amd.library_path(flavour)}").read()
        assert any(l.split()[-3:-2] == ["invalidQuESTInputError"] or "invalidQuESTInputError" in l
                   for l in relocs.splitlines() if "JUMP_SLOT" in l)
        assert not any("createQureg" in l for l in relocs.splitlines())
