"""Shared pytest configuration.

Native libraries are built by __graft_entry__.build() (or
``python -m lpmcxx_amd.build``) before the tests run; tests only load them.
"""
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built HIP library")


@pytest.fixture(scope="session")
def ctx():
    """One GPU context shared by the GPU tests of a session."""
    import lpmcxx_amd as L
    c = L.Context(0)
    yield c
    c.close()


_mesh_cache = {}


@pytest.fixture(scope="session")
def mesh():
    """mesh(seed, level) -> lpmcxx_amd.Mesh, cached per session."""
    import lpmcxx_amd as L

    def get(seed, level):
        key = (seed, level)
        if key not in _mesh_cache:
            _mesh_cache[key] = L.Mesh(seed, level)
        return _mesh_cache[key]
    return get
