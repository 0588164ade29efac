import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and libhiprite.so")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


def gpu_available() -> bool:
    try:
        import torch

        return bool(getattr(torch.version, "hip", None)) and torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(autouse=True)
def _collect_pipelines():
    """Collect each test's garbage when it ends, so the slot-row arenas of its
    pipelines return to the device-wide arena budget before the next test."""
    yield
    import gc

    gc.collect()


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o

    o.build()
    return o


@pytest.fixture(scope="session")
def device():
    if not gpu_available():
        pytest.fail("GPU test collected on a machine without a HIP device")
    import torch

    from pyrite_amd.backend import require_device

    return require_device().torch_device


os.environ.setdefault("OMP_NUM_THREADS", str(min(8, os.cpu_count() or 1)))
