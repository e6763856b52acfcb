import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: longer CPU test")


def random_state(n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    s = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return (s / np.linalg.norm(s)).astype(np.complex128)


@pytest.fixture
def rstate():
    return random_state
