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
def tube_config(tmp_path_factory):
    import workloads

    d = tmp_path_factory.mktemp("c0")
    return workloads.straight_tube(str(d))


@pytest.fixture(scope="session")
def tube_geometry(tube_config):
    import lbm

    return lbm.Geometry.from_config(tube_config)
