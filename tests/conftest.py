import os
import sys
from pathlib import Path

import numpy as np
import pytest

# One HIP runtime per process: torch ships its own libamdhip64 (same soname as ROCm's). Loaded
# first, it is the runtime libqoraholo_hip.so binds to; loaded after the library, it would be a
# second runtime that cannot open the device. The rocFFT cross-checks use torch, so it goes first.
try:
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


@pytest.fixture(scope="session")
def oracle():
    import oracle as o

    o.lib()
    return o


@pytest.fixture(scope="session")
def qh():
    import qoraholo_amd as q

    return q


@pytest.fixture(scope="session")
def device(qh):
    """The HIP path must be there on a GPU run: a missing library or device fails the test."""
    from qoraholo_amd import _native

    _native.lib()
    assert qh.hip_available(), "no HIP device visible"
    return True


def rel_l2(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return float(np.sqrt(np.sum(np.abs(a - b) ** 2) / np.sum(np.abs(b) ** 2)))


def contrast_error(a, b, nm):
    """Relative L2 on the index contrast ri - n_m (the reference's tests/test_cuda.cpp:92-97)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum((b - nm) ** 2)))
