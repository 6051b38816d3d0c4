"""Shared test setup: import paths, the gpu marker, common fixtures."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "mg-cfd-app-op2_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def box8():
    import mgcfd
    return mgcfd.mesh_box(8, 3, seed=7)


@pytest.fixture(scope="session")
def box40():
    import mgcfd
    return mgcfd.mesh_box(40, 4, seed=1)


@pytest.fixture(scope="session")
def knn_small():
    import mgcfd
    return mgcfd.mesh_knn(3000, 9300, 4, seed=3)


def perturbed_state(n, ffv, seed, amp=0.01):
    """Freestream plus a seeded +-amp relative perturbation (no special values)."""
    rng = np.random.default_rng(seed)
    return np.asarray(ffv)[None, :] * (1.0 + amp * rng.uniform(-1, 1, size=(n, 5))) + \
        amp * rng.uniform(-1, 1, size=(n, 5)) * np.array([0, 1, 1, 1, 0])[None, :]
