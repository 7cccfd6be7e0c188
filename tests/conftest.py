import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def hip():
    """The HIP library, loaded; the GPU tests fail loudly without it."""
    import torch

    import daxtoolkit_amd
    assert torch.cuda.is_available(), "GPU test run without a HIP device"
    return daxtoolkit_amd.load_library()
