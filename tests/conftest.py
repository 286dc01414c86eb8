import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def port():
    from oracle import pyoracle
    if not pyoracle.available("port"):
        pytest.skip("oracle/libpmoracle.so not built (make -C oracle)")
    return pyoracle.Oracle("port")


@pytest.fixture(scope="session")
def ref():
    from oracle import pyoracle
    if not pyoracle.available("reference"):
        pytest.skip("oracle/_ref/libpmref.so not built (needs the reference tree; make -C oracle ref)")
    return pyoracle.Oracle("reference")


@pytest.fixture(scope="session")
def gpu_ctx_factory():
    from protomol_amd import gpu
    made = []

    def make(precision="mixed"):
        c = gpu.Context(0, precision)
        made.append(c)
        return c

    yield make
    for c in made:
        c.close()
