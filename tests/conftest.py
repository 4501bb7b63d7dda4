"""Test configuration: the `gpu` marker, import paths and shared builders."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "multigrid-ldg_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def gctx():
    """GPU context; only GPU tests use it."""
    import mgldg
    return mgldg.Context(0)


def sine_f(X):
    """-Laplace of sin(pi x) sin(pi y) (the C1 manufactured solution)"""
    return 2 * np.pi ** 2 * np.sin(np.pi * X[:, 0]) * np.sin(np.pi * X[:, 1])
