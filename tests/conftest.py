import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "tests"))
sys.path.insert(0, os.path.join(REPO, "intel-phi-visiontrack_amd"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libvtgpu.so on the device)")


@pytest.fixture(scope="session")
def oracle():
    from oracle_ffi import Oracle, ORACLE_SO
    if not os.path.exists(ORACLE_SO):
        pytest.fail("oracle/liboracle.so is not built (make -C oracle)")
    return Oracle()


@pytest.fixture(scope="session")
def ctx():
    import vtgpu
    c = vtgpu.Context(0)
    yield c
    c.close()
