import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def golden_driver():
    import numpy as np
    return np.load(os.path.join(ROOT, "tests", "golden", "driver_cases.npz"), allow_pickle=False)


@pytest.fixture(scope="session")
def golden_knn():
    import numpy as np
    return np.load(os.path.join(ROOT, "tests", "golden", "knn_cases.npz"), allow_pickle=False)
