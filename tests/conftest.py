import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def gpu():
    """Loads the HIP library and checks that a GPU is present; never falls back."""
    import torch
    import cuspfft_amd as sp
    sp.load()
    assert torch.cuda.is_available(), "GPU test run without a visible GPU"
    return sp
