"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer, on the CPU.

tests/sanitize/Makefile builds one C driver (tests/sanitize/host_api.c) three
ways: against the oracle's C restatement (every oracle entry point on small
state-vector and density registers) and against libquest_mi355x's host sources
in both precision flavours (circuit recording and fusion at every block width,
PauliHamil files, matrices, seeding, and 23 inputs the C API must reject through
a returning error hook). The library's .hip objects are linked as the library
build made them; no device code is built with a sanitizer and nothing here needs
a GPU. Any sanitizer report ends the driver with a non-zero status
(-fno-sanitize-recover=all). The reference builds its CPU backend with a memory
checker option for the same purpose (QUEST_MEMCHECK, QuEST/CMakeLists.txt:357-366)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OUT = os.path.join(ROOT, "_build", "sanitize")

pytestmark = pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("make") is None,
                                reason="needs gcc and make")


@pytest.fixture(scope="module")
def drivers():
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "tests", "sanitize"), "-j4"],
                       capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return OUT


@pytest.mark.parametrize("name,expect", [
    ("host_api_cpuref", "every oracle entry point ran clean"),
    ("host_api_lib", "qreal of 8 bytes): 23 rejections reported"),
    ("host_api_lib_f32", "qreal of 4 bytes): 23 rejections reported"),
])
def test_host_code_runs_clean_under_sanitizers(drivers, tmp_path, name, expect):
    env = dict(os.environ, OMP_NUM_THREADS="2", ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1",
               TMPDIR=str(tmp_path))
    r = subprocess.run([os.path.join(drivers, name)], capture_output=True, text=True, timeout=300, cwd=tmp_path,
                       env=env)
    out = r.stdout + r.stderr
    if "does not come first in initial library list" in out:
        pytest.skip("AddressSanitizer cannot start in this process environment (another preloaded library)")
    assert r.returncode == 0, out[-4000:]
    assert "AddressSanitizer" not in out and "runtime error" not in out and "LeakSanitizer" not in out, out[-4000:]
    assert expect in out
