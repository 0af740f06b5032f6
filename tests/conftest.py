"""Shared pytest set-up: the `gpu` marker, repository paths, fixture loading."""
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "hsi-detectors-cuda_amd")
GOLDEN = os.path.join(REPO, "tests", "golden")
for path in (REPO, PKG):
    if path not in sys.path:
        sys.path.insert(0, path)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def golden(name):
    """Load a golden fixture written by tests/golden/make_golden.py."""
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as data:
        return {k: data[k] for k in data.files}


SCENES = ["scene_24x32x16_t3", "scene_32x40x113_t4", "scene_20x20x224_t16", "scene_13x37x113_t8"]


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    oracle.lib()
    return oracle
