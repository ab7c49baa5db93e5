import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

MODELS = ("plain_2d", "fiducial_3d", "wedge_2d")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library through the C ABI)")


@pytest.fixture(scope="session")
def oracles():
    from oracle.pyoracle import Oracle
    return {m: Oracle(m) for m in MODELS}


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test collected on a host without a GPU")
    return torch.device("cuda:0")
