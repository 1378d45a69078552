import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
C_AIR, RHO_AIR = 343.0, 1.21


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on cuda:0)")


def wavenumber(f):
    return 2.0 * np.pi * f / C_AIR


@pytest.fixture(scope="session")
def sphere_l3():
    """Config-1 geometry: icosphere level 3, r = 0.1 m, outward winding."""
    import fmbem_gpu_c_amd as F
    m = F.Mesh.icosphere(3, 0.1)
    V, Fc = m.arrays()
    return m, V, Fc
