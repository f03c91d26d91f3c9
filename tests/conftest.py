import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running (full-size) case")


@pytest.fixture(scope="session")
def ref_fixtures():
    with open(os.path.join(GOLDEN, "reference_tests.json")) as f:
        fx = json.load(f)
    imgs = {}
    for k, v in fx["images"].items():
        n = int(round(len(v) ** 0.5))
        imgs[k] = np.array(v, np.float32).reshape(n, n)
    fx["arrays"] = imgs
    return fx


@pytest.fixture(scope="session")
def ctx():
    import cpvs_amd
    c = cpvs_amd.default_context()
    yield c
