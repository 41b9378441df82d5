import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


_MODELS = {}


def hex_model(mesh):
    """Seeded 22-orbital hexagonal model with its gap bisected on ``mesh``."""
    from wtb_amd import tbmodel as T
    key = ("hex", tuple(mesh))
    if key not in _MODELS:
        _MODELS[key] = T.random_hexagonal(seed=2026, gap_mesh=tuple(mesh))
    return _MODELS[key].copy()


def cubic_model(mesh):
    from wtb_amd import tbmodel as T
    key = ("cubic", tuple(mesh))
    if key not in _MODELS:
        _MODELS[key] = T.random_cubic(seed=2027, gap_mesh=tuple(mesh))
    return _MODELS[key].copy()


@pytest.fixture
def rng():
    return np.random.default_rng(12345)
