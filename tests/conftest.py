import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from __graft_entry__ import load_package, load_oracle  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built HIP library")


@pytest.fixture(scope="session")
def swm():
    return load_package()


@pytest.fixture(scope="session")
def orc():
    o = load_oracle()
    o.build()
    return o


def gpu_available() -> bool:
    try:
        import ctypes
        swm = load_package()
        from importlib import import_module
        abi = import_module("shallowwaters_amd._abi")
        L = abi.load()
        n = ctypes.c_int32(0)
        L.swm_device_count(ctypes.byref(n))
        return n.value > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not gpu_available():
        pytest.fail("GPU test requested but no GPU / HIP library is available")
    return True


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    n = np.linalg.norm(a)
    return np.linalg.norm(a - b) / (n if n > 0 else 1.0)


def first_mismatch(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    bad = np.argwhere(~((a == b) | (np.isnan(a) & np.isnan(b))))
    if len(bad) == 0:
        return None
    i, j = bad[0]
    return f"{len(bad)} mismatches; first at Julia [{i + 1},{j + 1}]: expected {a[i, j]!r} got {b[i, j]!r}"
