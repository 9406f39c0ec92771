import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libosqp_amd.so on the GPU)")


@pytest.fixture(scope="session")
def algebra():
    """Initialises the algebra/hip backend once for the boundary tests."""
    from osqp_amd import algebra as alg
    assert alg.init(0) == 0, "osqp_algebra_init_libs(0) failed: no usable GPU"
    yield alg
    alg.shutdown()
