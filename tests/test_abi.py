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
