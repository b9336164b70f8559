import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "reza-seti_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on cuda:0)")


@pytest.fixture(scope="session")
def c0_config(tmp_path_factory):
    import workloads

    d = tmp_path_factory.mktemp("c0")
    return workloads.c0_tube(str(d))


@pytest.fixture(scope="session")
def c0_geometry(c0_config):
    import lbm

    return lbm.Geometry.from_config(c0_config)
