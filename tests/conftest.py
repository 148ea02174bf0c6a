import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built HIP library")
    config.addinivalue_line("markers", "slow: over ~10 s")


@pytest.fixture
def gpu():
    """The HIP engine, or a skip where no GPU is visible. A visible GPU without the built library is a failure."""
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    from blochparty_amd.engines import hip

    if not hip.LIB_PATH.exists():
        pytest.fail(f"GPU present but {hip.LIB_PATH} is not built")
    hip.load()
    return hip.HIP
