import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def hier_small():
    """3-level hierarchy from 20k Gaussian points, depth 7 (about 60k elements)."""
    from tests.meshes import hierarchy
    return hierarchy("gauss", 20000, 7, 3, seed=1, min_coarse=50)


@pytest.fixture(scope="session")
def hier_tiny():
    """Small 3-level hierarchy whose coarsest level a dense solve handles instantly."""
    from tests.meshes import hierarchy
    return hierarchy("gauss", 3000, 6, 3, seed=5, min_coarse=20)
