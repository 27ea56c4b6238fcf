"""QuEST's own interface tests (Code/QuEST340/tests: 126 Catch2 TEST_CASEs and
their dense-matrix utilities), compiled from where they lie against this
project's include/ and linked to the product library, run on the GPU, in both
precision flavours: QuEST_PREC=2 against libquest_mi355x.so and QuEST_PREC=1
against libquest_mi355x_f32.so (tolerance REAL_EPS = 1e-13 / 1e-5, as QuEST
sets it). The suite's main() replaces the weak invalidQuESTInputError with one
that throws, so its input-validation sections check the library's messages.

The binaries are built by oracle/Makefile (target questtests) where the
reference tree exists and travel with the repository; one pytest case runs one
tag of the suite."""
import os
import subprocess

import pytest

from tests.conftest import ROOT

TAGS = ["calculations", "data_structures", "decoherence", "gates", "operators", "state_initialisations",
        "unitaries"]


@pytest.mark.gpu
@pytest.mark.timeout(900)
@pytest.mark.parametrize("flavour", ["f64", "f32"])
@pytest.mark.parametrize("tag", TAGS)
def test_quest_catch2_suite(flavour, tag):
    exe = os.path.join(ROOT, "oracle", "_ref", f"quest_tests_{flavour}")
    if not os.path.exists(exe):
        pytest.skip("QuEST's test binaries are built only where the reference tree exists (oracle/Makefile)")
    r = subprocess.run([exe, f"[{tag}]"], capture_output=True, text=True, timeout=850)
    out = r.stdout + r.stderr
    summary = [l for l in out.splitlines() if "All tests passed" in l or l.startswith("test cases:")]
    print(f"QuEST suite {flavour} [{tag}]: {summary[-1] if summary else 'no summary'}")
    assert r.returncode == 0, "\n".join(out.splitlines()[-120:])
    assert summary and "All tests passed" in summary[-1], summary
