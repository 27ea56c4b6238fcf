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
    ("dangling_exponent", "1e 1 2\n2.5E+ 3 0\n-1.e- 0 1\n"),
    ("hex_coefficients", "0x1p-2 1 2\n-0X1.8P1 3 0\n0x1p 0 1\n"),
    ("inf_and_nan_names", "inf 1\n-Infinity 2\nNAN 3\n"),
    ("nan_with_chars", "nan(7) 0\n"),
    ("hex_without_digits", "0x 1 2\n"),
    ("lone_sign", "- 1 2\n"),
    ("lone_point", ". 1 2\n"),
    ("exponent_then_letter", "1e+x 1 2\n"),
]


def _same(a, b):
    """equal reads, a NaN coefficient matching a NaN"""
    def norm(r):
        return {k: ([("nan" if x != x else x) for x in v] if isinstance(v, list) else v) for k, v in r.items()}
    return norm(a) == norm(b)


def _random_file(rng):
    """A PauliHamil file from the format's tokens and near-misses: numbers in
    every fscanf syntax, out-of-range and malformed codes, ragged terms, tabs,
    CRLF, blank lines, leading blanks, no final newline."""
    coef = ["0.5", "-1.25", "1e-3", "+2", "3.", ".5", "-0", "abc", "1e", "0x1p3", "inf", "nan", "1x", "7", "2.5E+"]
    code = ["4", "-1", "01", "3.0", "x", "+1", "2e0"]
    width, lines = rng.integers(1, 5), []
    for _ in range(rng.integers(1, 5)):
        toks = [str(rng.choice(coef)) if rng.random() < 0.3 else str(rng.choice(["0.5", "-1.25", "0.125"]))]
        k = width + (int(rng.choice([-1, 1])) if rng.random() < 0.1 else 0)
        toks += [str(rng.choice(code)) if rng.random() < 0.1 else str(rng.integers(0, 4)) for _ in range(max(0, k))]
        line = toks[0] + "".join(str(rng.choice([" ", " ", "  ", "\t"])) + t for t in toks[1:])
        lines.append((" " if rng.random() < 0.1 else "") + line)
    text = "".join(l + str(rng.choice(["\n", "\n", "\r\n", "\n\n", " \n"])) for l in lines)
    return text.rstrip("\n") if rng.random() < 0.2 else text

_CHILD = r"""
import ctypes as C, json, sys
lib = C.CDLL(sys.argv[1])
qreal = C.c_float if sys.argv[3] == "f32" else C.c_double
class H(C.Structure):
    _fields_ = [("pauliCodes", C.POINTER(C.c_int)), ("termCoeffs", C.POINTER(qreal)),
                ("numSumTerms", C.c_int), ("numQubits", C.c_int)]
lib.createPauliHamilFromFile.restype = H
lib.createPauliHamilFromFile.argtypes = [C.c_char_p]
h = lib.createPauliHamilFromFile(sys.argv[2].encode())
print(json.dumps({"n": h.numQubits, "terms": h.numSumTerms,
                  "coeffs": [h.termCoeffs[t] for t in range(h.numSumTerms)],
                  "codes": [h.pauliCodes[i] for i in range(h.numSumTerms * h.numQubits)]}))
"""


REFLIB_F32 = os.path.join(ROOT, "oracle", "_ref", "libquest_ref_f32.so")


def reference_read(path, flavour="f64"):
    ref = REFLIB_F32 if flavour == "f32" else REFLIB
    r = subprocess.run([sys.executable, "-c", _CHILD, ref, path, flavour], capture_output=True, text=True, timeout=60)
    if r.returncode == 0:
        return json.loads(r.stdout)
    line = [l for l in r.stderr.splitlines() if "QuEST reference error" in l]
    assert line, r.stderr
    return {"error": line[0].split(": ", 1)[1]}


def library_read(path, flavour="f64"):
    from quantumlab_amd.quest import library_path

    class PauliHamil(C.Structure):
        _fields_ = [("pauliCodes", C.POINTER(C.c_int)),
                    ("termCoeffs", C.POINTER(C.c_float if flavour == "f32" else C.c_double)),
                    ("numSumTerms", C.c_int), ("numQubits", C.c_int)]

    lib = C.CDLL(library_path(flavour), mode=os.RTLD_LOCAL)
    lib.qmxSetErrorMode(1)
    lib.createPauliHamilFromFile.restype = PauliHamil
    lib.destroyPauliHamil.argtypes = [PauliHamil]
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
    assert _same(got, ref), (ref, got)


@pytest.mark.parametrize("flavour", ["f64", "f32"])
def test_random_pauli_hamil_files_match_reference_reader(tmp_path, flavour):
    """80 seeded random files: the library reads each as QuEST's CPU build at
    the same precision does (same values, or the same message)."""
    if not os.path.exists(REFLIB_F32 if flavour == "f32" else REFLIB):
        pytest.skip("reference QuEST build (oracle/_ref) not present")
    import numpy as np

    rng = np.random.default_rng(2024)
    for i in range(80):
        text = _random_file(rng)
        path = str(tmp_path / f"h{i}.txt")
        with open(path, "w", newline="") as f:
            f.write(text)
        ref, got = reference_read(path, flavour), library_read(path, flavour)
        assert _same(got, ref), (repr(text), ref, got)


def test_pauli_hamil_file_known_answer(tmp_path):
    """Independent of the reference: the plain file reads to its literal values."""
    path = str(tmp_path / "h.txt")
    with open(path, "w") as f:
        f.write("0.5 1 2 3\n-1.25 0 3 0\n")
    assert library_read(path) == {"n": 3, "terms": 2, "coeffs": [0.5, -1.25], "codes": [1, 2, 3, 0, 3, 0]}
    missing = str(tmp_path / "absent.txt")
    err = library_read(missing)["error"]
    assert "Could not open file" in err or "open" in err.lower()
