import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def ctx():
    import katric_amd as K

    c = K.Context(0)
    yield c
    c.close()


def golden(name):
    return os.path.join(GOLDEN, name)


# Known answers from the reference's examples/examples.yaml:1-13.
KNOWN_TOTALS = {"triangle.metis": 1, "metis-sample.metis": 4, "hiv.metis": 45}
