import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

MESHES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "meshes")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def gpu_ctx():
    import hyteg_amd as H
    return H.context(0)


@pytest.fixture(scope="session")
def host_ctx():
    import hyteg_amd as H
    return H.Context(-1)
