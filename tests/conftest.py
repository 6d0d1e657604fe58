import os
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parents[1]
if str(REPO) not in sys.path:
    sys.path.insert(0, str(REPO))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libzombie_amd.so")
    config.addinivalue_line("markers", "slow: long-running test")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    if not oracle.LIB.exists():
        from zombie_amd import _build
        _build.build_oracle()
    return oracle


@pytest.fixture(scope="session")
def gpu():
    """Loads the HIP library and checks that a GPU is present; fails loudly otherwise."""
    import torch
    from zombie_amd import _abi
    if not torch.cuda.is_available():
        pytest.fail("GPU test selected but torch.cuda.is_available() is False")
    _abi.lib()
    return torch
