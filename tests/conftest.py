"""Shared fixtures of the test suite.

Tests marked ``gpu`` need an MI355X (gfx950) and the built HIP library; the
rest run on the CPU: the oracle against the reference's golden records, the
host code of the library, and the C-ABI exports.
"""
import json
import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X and the built HIP library")
    config.addinivalue_line("markers", "slow: takes more than a few seconds")


@pytest.fixture(scope="session")
def golden():
    return json.loads((GOLDEN / "expected.json").read_text())


@pytest.fixture(scope="session")
def structures():
    z = np.load(GOLDEN / "structures.npz")

    def get(name):
        xyz = z[name + "_xyz"].astype(np.float64) / 1000.0
        return np.column_stack([xyz, z[name + "_r"].astype(np.float64)])

    return get


@pytest.fixture(scope="module", autouse=True)
def free_cycles_after_module():
    """Device objects a module left in reference cycles (through tracebacks,
    closures) are freed when the module ends, not by a later collection
    inside another module's measurement."""
    yield
    import gc
    gc.collect()


@pytest.fixture(scope="session")
def device():
    import sesmesh_amd
    return sesmesh_amd.Device(0)


def cpu_threads():
    return max(1, min(16, os.cpu_count() or 1))
