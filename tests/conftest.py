import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "rle-based-voxel-raycasting_amd")
for p in (PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def c1_scene():
    import workloads
    return workloads.gen("c1", seed=1)


@pytest.fixture(scope="session")
def tiny_scene():
    import workloads
    return workloads.gen("tiny", seed=5)


@pytest.fixture(scope="session")
def ctx():
    import rlevox
    c = rlevox.Context(0)
    yield c
    c.close()
