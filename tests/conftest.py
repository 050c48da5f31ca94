import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


OMEGA_M2 = 2.0 * np.pi / 44714.0


@pytest.fixture(scope="session")
def case_tiny():
    from actfemtide_amd import mesh
    return mesh.tide_case(4, 4, 6, seed=11)


@pytest.fixture(scope="session")
def case_small():
    from actfemtide_amd import mesh
    return mesh.tide_case(8, 8, 6, seed=1001)


@pytest.fixture(scope="session")
def oracle_small(case_small):
    from oracle import oracle as O
    return O.forward_system(case_small, OMEGA_M2)


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
