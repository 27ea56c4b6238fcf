import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

CPUREF = os.path.join(ROOT, "oracle", "build", "libcpuref.so")
REFLIB = os.path.join(ROOT, "oracle", "_ref", "libquest_ref.so")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs the HIP library")


@pytest.fixture(scope="session")
def cpuref():
    from quantumlab_amd.quest import QuEST
    if not os.path.exists(CPUREF):
        pytest.fail(f"{CPUREF} not built: run __graft_entry__.build()")
    return QuEST(CPUREF)


@pytest.fixture(scope="session")
def reflib():
    from quantumlab_amd.quest import QuEST
    if not os.path.exists(REFLIB):
        pytest.skip("reference QuEST build (oracle/_ref) not present")
    return QuEST(REFLIB)


@pytest.fixture(scope="session")
def gpu():
    import quantumlab_amd
    return quantumlab_amd.load()


def golden(name):
    import numpy as np
    return np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
