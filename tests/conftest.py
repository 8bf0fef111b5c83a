import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs the HIP library)")


@pytest.fixture(scope="session")
def fixture78():
    """The reference's only mesh fixture (inputFiles/radius015/*78.txt)."""
    from rcs_amd import mesh as M
    return M.read_reference_files(os.path.join(GOLDEN, "nodeList78.txt"), os.path.join(GOLDEN, "patchList78.txt"))


@pytest.fixture(scope="session")
def rng():
    return np.random.default_rng(20261203)
