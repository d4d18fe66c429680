import glob
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")


def golden(name: str):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def golden_files(pattern: str):
    return sorted(os.path.basename(p) for p in glob.glob(os.path.join(GOLDEN, pattern)))


def golden_input(name: str) -> str:
    return os.path.join(GOLDEN, "inputs", name)


def diameter(xy) -> float:
    xy = np.asarray(xy)
    return float(np.ptp(xy, axis=0).max())


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle

    pyoracle.build()
    return pyoracle


@pytest.fixture(scope="session")
def bl():
    from batchlayoutcode_amd import api

    api.lib()
    return api
