import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and libeplace.so")


def _has_gpu() -> bool:
    try:
        import torch  # noqa: F401  (device discovery only)
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def device():
    """One placer context for the whole GPU session."""
    import eisenjoch_amd as ep
    if not _has_gpu():
        pytest.fail("GPU test run without a HIP device")
    dev = ep.Device(0)
    yield dev
    dev.close()


def log_case(msg: str):
    """Print what names the case before a GPU step, flushed."""
    print(f"[case] {msg}", flush=True)
