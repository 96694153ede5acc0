import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP libraries)")


@pytest.fixture(scope="session")
def device():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


@pytest.fixture(scope="session")
def oracle_libs():
    """Builds the CPU oracle libraries that are missing (about 10 s for all)."""
    from oracle import build

    missing = [a for a in build.APPS if not build.library_path(a).exists()]
    if missing:
        build.build(missing)
    return True
