import os
import sys

import numpy as np
import pytest

# torch and the library both link libamdhip64.so.7, and the process keeps the copy
# loaded first: torch's bundled runtime must be that copy, or torch cannot see the
# GPU once the library (linked against /opt/rocm) has initialised HIP.  Importing
# torch does not initialise HIP.
try:
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden", "golden.npz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def golden():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


def gpu_available():
    try:
        import torch  # noqa: F401  (device count only; does not initialise HIP)
        return torch.cuda.device_count() > 0
    except Exception:
        return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def nyx():
    """The HIP library; GPU tests fail loudly if it is not built."""
    import nyx_amd
    nyx_amd.load()
    return nyx_amd
