import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (HIP device) and the built libldl.so")


@pytest.fixture(scope="session")
def golden_cic():
    import numpy as np
    return dict(np.load(os.path.join(ROOT, "tests", "golden", "cic_ref_golden.npz")))
