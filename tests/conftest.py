import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (ROCm GPU) and the built libtraceon_hip.so")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle():
    from oracle.oracle import restatement
    return restatement()


@pytest.fixture(scope="session")
def ref_lib():
    from oracle.oracle import reference
    r = reference()
    if r is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    return r
