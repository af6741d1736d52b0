import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP backend)")


def pytest_collection_modifyitems(config, items):
    # GPU tests: fail loudly if the device is absent rather than pass on a CPU path
    pass


@pytest.fixture(scope="session")
def oracle_built():
    import pyoracle
    if not pyoracle.available("oracle"):
        import subprocess
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "oracle"], check=True)
    return True
