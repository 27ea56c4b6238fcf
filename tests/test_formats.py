"""Text formats at the QuEST boundary, against the reference's own readers and
writers.

PauliHamil files (QuEST.c:1459-1541): createPauliHamilFromFile of the library
and of QuEST's CPU build (oracle/_ref) read the same edge-case files to the
same width, terms, coefficients and codes, or fail with the same QuEST message.
The reference reader exits on invalid input, so each of its calls runs in a
child process. Host code only: no GPU is touched."""
import ctypes as C
import json
import os
import subprocess
import sys

import pytest

from tests.conftest import REFLIB, ROOT

# (name, file text)
HAMIL_FILES = [
    ("plain", "0.5 1 2 3\n-1.25 0 3 0\n"),
    ("no_final_newline", "0.5 1 2 3\n-1.25 0 3 0"),
    ("trailing_newlines", "0.5 1 2 3\n-1.25 0 3 0\n\n\n"),
    ("blank_lines_between", "0.5 1 2\n\n\n0.25 3 3\n"),
    ("multiple_spaces", "0.5   1  2 3\n-1.25 0    3 0\n"),
    ("trailing_space_first_line", "0.5 1 2 3 \n-1.25 0 3 0\n"),
    ("leading_space_first_line", " 0.5 1 2\n0.25 3 3 1\n0.125 2 2 2\n"),
    ("tab_separated", "0.5\t1 2 3\n-1.25 0 3 0\n"),
    ("crlf", "0.5 1 2 3\r\n-1.25 0 3 0\r\n"),
    ("scientific", "1e-3 1 2\n-2.5E+1 3 0\n+0.75 0 0\n"),
    ("terms_across_lines", "0.5 1 2\n0.25\n3 3\n"),
    ("bad_code", "0.5 1 4 3\n"),
    ("negative_code", "0.5 1 -1 3\n"),
    ("bad_coefficient", "abc 1 2\n"),
    ("missing_codes", "0.5 1 2 3\n0.25 1 2\n"),
    ("blank_last_line", "0.5 1 2\n   \n"),
    ("empty", ""),
    ("only_coefficient", "0.5\n"),
    ("code_with_suffix", "0.5 1x 2\n"),
]

_CHILD = r"""
import ctypes as C, json, sys
lib = C.CDLL(sys.argv[1])
class H(C.Structure):
    _fields_ = [("pauliCodes", C.POINTER(C.c_int)), ("termCoeffs", C.POINTER(C.c_double)),
                ("numSumTerms", C.c_int), ("numQubits", C.c_int)]
lib.createPauliHamilFromFile.restype = H
lib.createPauliHamilFromFile.argtypes = [C.c_char_p]
h = lib.createPauliHamilFromFile(sys.argv[2].encode())
print(json.dumps({"n": h.numQubits, "terms": h.numSumTerms,
                  "coeffs": [h.termCoeffs[t] for t in range(h.numSumTerms)],
                  "codes": [h.pauliCodes[i] for i in range(h.numSumTerms * h.numQubits)]}))
"""


def reference_read(path):
    r = subprocess.run([sys.executable, "-c", _CHILD, REFLIB, path], capture_output=True, text=True, timeout=60)
    if r.returncode == 0:
        return json.loads(r.stdout)
    line = [l for l in r.stderr.splitlines() if "QuEST reference error" in l]
    assert line, r.stderr
    return {"error": line[0].split(": ", 1)[1]}


def library_read(path):
    import quantumlab_amd
    from quantumlab_amd.quest import PauliHamil
    lib = C.CDLL(quantumlab_amd.library_path(), mode=os.RTLD_LOCAL)
    lib.qmxSetErrorMode(1)
    lib.createPauliHamilFromFile.restype = PauliHamil
    lib.createPauliHamilFromFile.argtypes = [C.c_char_p]
    lib.qmxLastError.argtypes = [C.c_char_p, C.c_int]
    lib.qmxClearError()
    h = lib.createPauliHamilFromFile(path.encode())
    buf = C.create_string_buffer(1024)
    if lib.qmxLastError(buf, 1024):
        lib.qmxClearError()
        return {"error": buf.value.decode().split(": ", 1)[1]}
    out = {"n": h.numQubits, "terms": h.numSumTerms,
           "coeffs": [h.termCoeffs[t] for t in range(h.numSumTerms)],
           "codes": [h.pauliCodes[i] for i in range(h.numSumTerms * h.numQubits)]}
    lib.destroyPauliHamil(h)
    return out


@pytest.mark.parametrize("name,text", HAMIL_FILES, ids=[n for n, _ in HAMIL_FILES])
def test_pauli_hamil_file_matches_reference_reader(tmp_path, name, text):
    if not os.path.exists(REFLIB):
        pytest.skip("reference QuEST build (oracle/_ref) not present")
    path = str(tmp_path / f"{name}.txt")
    with open(path, "w", newline="") as f:
        f.write(text)
    ref = reference_read(path)
    got = library_read(path)
    assert got == ref


def test_pauli_hamil_file_known_answer(tmp_path):
    """Independent of the reference: the plain file reads to its literal values."""
    path = str(tmp_path / "h.txt")
    with open(path, "w") as f:
        f.write("0.5 1 2 3\n-1.25 0 3 0\n")
    assert library_read(path) == {"n": 3, "terms": 2, "coeffs": [0.5, -1.25], "codes": [1, 2, 3, 0, 3, 0]}
    missing = str(tmp_path / "absent.txt")
    err = library_read(missing)["error"]
    assert "Could not open file" in err or "open" in err.lower()
