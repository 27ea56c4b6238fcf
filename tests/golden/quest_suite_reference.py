"""Record which sections of QuEST's own Catch2 suite fail on QuEST's own CPU
build, per tag and precision (tests/golden/quest_suite_reference.json).

The suite (Code/QuEST340/tests) draws its random states and matrices from
rand() without seeding, so a given invocation is deterministic; a few of its
sections fail on the reference build itself (a random 5-qubit matrix that
misses the 1e-13 unitarity check, a 1e-13 comparison, distribution-limit
checks whose rank count overflows int). tests/test_yy_quest_suite.py runs the
same tags the same way against this library and accepts these.

Flavours: f64 (QuEST_PREC=2) and f32 (QuEST_PREC=1).

Run in the development container, where oracle/Makefile builds the suite
against the reference sources (target questtests-ref); named flavours are
re-run and merged into the existing file:
    python tests/golden/quest_suite_reference.py [f64 f32]
"""
import json
import os
import re
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
TAGS = ["calculations", "data_structures", "decoherence", "gates", "operators", "state_initialisations",
        "unitaries"]


def main():
    path = os.path.join(ROOT, "tests", "golden", "quest_suite_reference.json")
    flavours = sys.argv[1:] or ["f64", "f32"]
    out = json.load(open(path))["results"] if os.path.exists(path) else {}
    for flavour in flavours:
        exe = os.path.join(ROOT, "oracle", "_ref", f"quest_tests_ref_{flavour}")
        if not os.path.exists(exe):
            sys.exit(f"{exe} missing: make -C oracle questtests-ref")
        out[flavour] = {}
        for tag in TAGS:
            with tempfile.TemporaryDirectory() as d:
                r = subprocess.run([exe, f"[{tag}]"], capture_output=True, text=True, cwd=d,
                                   env=dict(os.environ, OMP_NUM_THREADS="8"))
            text = r.stdout + r.stderr
            failed = sorted({m.group(1) for m in re.finditer(r"(test_\w+\.cpp:\d+): FAILED:", text)})
            summary = [l for l in text.splitlines() if l.startswith("test cases:") or "All tests passed" in l]
            out[flavour][tag] = {"failed_sections": failed, "summary": summary[-1] if summary else None}
            print(flavour, tag, out[flavour][tag], flush=True)
    with open(path, "w") as f:
        json.dump({"generator": "tests/golden/quest_suite_reference.py",
                   "reference": "QuEST 3.4 CPU backend (Code/QuEST340/QuEST) with QuEST's own Catch2 suite "
                                "(Code/QuEST340/tests), one process per tag",
                   "results": out}, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
