"""Shared pytest configuration: the `gpu` marker and common fixtures."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built library")
    config.addinivalue_line("markers", "slow: minutes of GPU time; runs when PETGEM_AMD_SLOW=1")


def pytest_collection_modifyitems(config, items):
    if os.environ.get("PETGEM_AMD_SLOW") == "1":
        return
    skip = pytest.mark.skip(reason="slow test: set PETGEM_AMD_SLOW=1 to run it")
    for item in items:
        if "slow" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def test_mesh():
    """The reference's own test mesh (tests/data/test_mesh.msh): 9453 tets, 4 materials."""
    from petgem_amd.mesh import read_gmsh_v2
    return read_gmsh_v2(os.path.join(GOLDEN, "test_mesh.msh"))


@pytest.fixture(scope="session")
def golden():
    return np.load(os.path.join(GOLDEN, "hvfem_vectors.npz"))


@pytest.fixture(scope="session")
def topology_golden():
    """Outputs of the reference's integer topology routines (make_topology_vectors.py)."""
    return np.load(os.path.join(GOLDEN, "topology_vectors.npz"))
