import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); calls the HIP library")
    config.addinivalue_line("markers", "slow: longer CPU oracle runs")


def _has_gpu() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if _has_gpu():
        return
    skip = pytest.mark.skip(reason="no HIP device visible")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    if not os.path.exists(oracle._LIB_PATH):
        oracle.build()
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def gpvi_lib():
    from geopvi_amd import _lib
    return _lib.load()
