import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the HIP library")


@pytest.fixture(scope="session")
def native():
    """The HIP library binding; a GPU test fails (never silently falls back) if it is missing."""
    from raytracergpu_mastersproject_amd import native as N
    assert N.device_count() > 0, "no HIP device visible"
    return N


@pytest.fixture(scope="session")
def ctx(native):
    c = native.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def O():
    import oracle
    return oracle
