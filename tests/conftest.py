import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oxiblas.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

# The reference test seed (test/testsetup.jl:5).
SEED = 0x0B1A5


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture
def rng():
    import numpy as np
    return np.random.default_rng(SEED)


@pytest.fixture(scope="session")
def gpu():
    """The HIP library on cuda:0; fails loudly when either is missing."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test collected without a GPU")
    import oxiblas_mi355x as ox
    ox.load()
    return ox
