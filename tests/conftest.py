import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (REPO, os.path.join(REPO, "umap-rs_amd"), os.path.join(REPO, "oracle"), os.path.join(REPO, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libumapgpu on the device)")
    config.addinivalue_line("markers", "slow: longer CPU checks")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    if not os.path.exists(os.path.join(REPO, "oracle", "liboracle.so")):
        O.build()
    return O


@pytest.fixture(scope="session")
def U():
    import umap_amd
    umap_amd.lib()  # raises if the library is not built
    return umap_amd


@pytest.fixture(scope="session")
def gpu(U):
    if U.device_count() < 1:
        pytest.fail("no GPU visible: the gpu-marked tests need an MI355X")
    return U
