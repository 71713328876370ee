import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tensor-episdet_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(ROOT, "tests", "golden", "tiny.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def golden_planes(golden):
    import numpy as np
    cases = np.array(golden["cases"], dtype=np.uint32)
    controls = np.array(golden["controls"], dtype=np.uint32)
    return cases, controls, golden["num_cases"], golden["num_controls"]
