import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: longer CPU test")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU visible")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)


_systems = {}


def system(cfg, n_cells, sample=0):
    """Cached generator output (felspa_amd.StokesSystem)."""
    import felspa_amd as fa
    key = (cfg, n_cells, sample)
    if key not in _systems:
        _systems[key] = fa.StokesSystem(fa.make_problem(cfg, n_cells=n_cells, sample=sample))
    return _systems[key]


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    d = np.linalg.norm(a - b)
    n = np.linalg.norm(b)
    return d / n if n > 0 else d


@pytest.fixture(scope="session")
def oracle():
    import oracle_py
    oracle_py.build()
    return oracle_py
