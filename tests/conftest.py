import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def golden():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "tinytest_vectors.json")) as fh:
        return json.load(fh)


def mean_rel(target, current):
    """R's all.equal() tolerance measure: mean(|t - c|) / mean(|t|)."""
    import numpy as np
    t = np.asarray(target, dtype=float)
    c = np.asarray(current, dtype=float)
    return float(np.mean(np.abs(t - c)) / np.mean(np.abs(t)))
