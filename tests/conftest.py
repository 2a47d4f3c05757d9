import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the gfx950 library)")


@pytest.fixture(scope="session")
def gpu_available():
    import torch  # noqa: F401  (only to count devices without initialising HIP)

    return torch.cuda.device_count() > 0
