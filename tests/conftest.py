import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def ctx():
    import gbm3d_amd

    c = gbm3d_amd.Context(0)
    yield c
    c.close()


def pytest_collection_modifyitems(config, items):
    # the quick checks first, the long ones last, test_zz_* files at the very end
    items.sort(key=lambda it: (os.path.basename(str(it.fspath)).startswith("test_zz"),
                               1 if it.get_closest_marker("slow") else 0))
