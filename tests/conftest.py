import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libocx.so")


@pytest.fixture(scope="session")
def oracle_c():
    """The C loop oracle (built on demand)."""
    from oracle import cref
    if not cref.LIB.exists():
        cref.build()
    return cref


@pytest.fixture(scope="session")
def gpu():
    """Skip-free GPU gate: a gpu-marked test fails loudly if the GPU or the library is missing."""
    import torch
    from pyocean_amd import _native
    assert torch.cuda.is_available(), "no GPU visible"
    _native.lib()
    return torch.device("cuda", 0)


def pytest_collection_modifyitems(config, items):
    # print which case runs before each GPU test (a crashed box returns only the tail of stdout)
    os.environ.setdefault("PYTHONUNBUFFERED", "1")
