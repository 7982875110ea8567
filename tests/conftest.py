"""pytest configuration: the `gpu` marker and device fixtures."""
import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _has_gpu() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def dev():
    """cuda:0 with the library initialised (fails loudly if the library is missing)."""
    if not _has_gpu():
        pytest.skip("no GPU visible")
    import torch
    import helpers
    p = helpers.pkg()
    p.check(p.lib().omb_init())
    return torch.device("cuda:0")


@pytest.fixture(scope="session")
def ompblas():
    import helpers
    return helpers.pkg()


@pytest.fixture(scope="session")
def orc():
    import helpers
    return helpers.oracle()
