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

# Sections that fail on QuEST's own CPU build as well, in both precisions
# (test_reference_build_fails_the_same_sections below checks it): they set
# env.numRanks = (int) pow(2, 2 minQb) or pow(2, minQb) for minQb up to 63,
# which overflows int, so no register is "too small to distribute" and the
# expected "Too few qubits" error cannot occur. These are the suite's
# distribution-limit cases (QT/test_data_structures.cpp:158-173, 225-238,
# 357-384, 637-650).
KNOWN_SUITE_DEFECTS = {"test_data_structures.cpp:173", "test_data_structures.cpp:238",
                       "test_data_structures.cpp:384", "test_data_structures.cpp:650"}


def failed_sections(out):
    import re
    return {m.group(1) for m in re.finditer(r"(test_\w+\.cpp:\d+): FAILED:", out)}


@pytest.mark.gpu
@pytest.mark.timeout(900)
@pytest.mark.parametrize("flavour", ["f64", "f32"])
@pytest.mark.parametrize("tag", TAGS)
def test_quest_catch2_suite(flavour, tag, tmp_path):
    exe = os.path.join(ROOT, "oracle", "_ref", f"quest_tests_{flavour}")
    if not os.path.exists(exe):
        pytest.skip("QuEST's test binaries are built only where the reference tree exists (oracle/Makefile)")
    # the suite writes scratch files (PauliHamil files) into its working directory
    r = subprocess.run([exe, f"[{tag}]"], capture_output=True, text=True, timeout=850, cwd=str(tmp_path))
    out = r.stdout + r.stderr
    summary = [l for l in out.splitlines() if "All tests passed" in l or l.startswith("test cases:")]
    failed = failed_sections(out)
    print(f"QuEST suite {flavour} [{tag}]: {summary[-1] if summary else 'no summary'}; failed {sorted(failed)}")
    assert summary, "\n".join(out.splitlines()[-120:])
    assert failed <= KNOWN_SUITE_DEFECTS, "\n".join(out.splitlines()[-200:])
    assert (r.returncode == 0) == (not failed), (r.returncode, sorted(failed))


@pytest.mark.parametrize("flavour", ["f64", "f32"])
def test_reference_build_fails_the_same_sections(flavour, tmp_path):
    """The suite's [data_structures] tag on QuEST's own CPU build (oracle/_ref,
    the reference sources with their weak error hook): exactly the sections
    listed in KNOWN_SUITE_DEFECTS fail, so accepting them above excuses nothing
    of this library. CPU only."""
    exe = os.path.join(ROOT, "oracle", "_ref", f"quest_tests_ref_{flavour}")
    if not os.path.exists(exe):
        pytest.skip("built only where the reference tree exists (oracle/Makefile questtests-ref)")
    env = dict(os.environ, OMP_NUM_THREADS="2")
    r = subprocess.run([exe, "[data_structures]"], capture_output=True, text=True, timeout=600, env=env,
                       cwd=str(tmp_path))
    assert failed_sections(r.stdout + r.stderr) == KNOWN_SUITE_DEFECTS
