import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def orc():
    import oracle
    if not oracle.available("oracle"):
        oracle.build(with_ref=False)
    return oracle.Oracle()


@pytest.fixture(scope="session")
def gplib():
    import goodphotons_amd as gp
    lib = gp.load()
    info = lib.device_info()   # raises GpError if no gfx950 device
    print(f"device: {info}", flush=True)
    return lib


def golden(name):
    import numpy as np
    return np.load(os.path.join(ROOT, "tests", "golden", name))
