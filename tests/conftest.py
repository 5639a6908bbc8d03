import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, HERE)
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: longer CPU checks")


@pytest.fixture(scope="session")
def golden_ref():
    import numpy as np
    return np.load(os.path.join(HERE, "golden", "ref_U2.npz"))


@pytest.fixture(scope="session")
def hmc_ref():
    import json
    with open(os.path.join(HERE, "golden", "hmc_ref.json")) as f:
        return json.load(f)
