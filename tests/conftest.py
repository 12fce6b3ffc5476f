import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    oracle.lib()  # builds oracle/_build on first use
    return oracle


@pytest.fixture(scope="session")
def gx():
    import gpuexplore_amd
    gpuexplore_amd.lib()
    return gpuexplore_amd


@pytest.fixture(scope="session")
def gpu(gx):
    if gx.device_count() == 0:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    return gx


@pytest.fixture
def model_dir(tmp_path):
    return str(tmp_path)
