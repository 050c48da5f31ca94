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


def crust_params(case, ncell_max=None, seed=0):
    """Parameters = crust cells (6 tets each): CSR (par_ptr, par_tet 0-based), cell ids."""
    crust = np.nonzero(case.group >= 3)[0]
    cells, inv = np.unique(case.cell_of_tet[crust], return_inverse=True)
    order = np.argsort(inv, kind="stable")
    tets = crust[order]
    counts = np.bincount(inv)
    ptr = np.concatenate([[0], np.cumsum(counts)])
    if ncell_max is not None and ncell_max < len(cells):
        rng = np.random.default_rng(seed)
        pick = np.sort(rng.choice(len(cells), ncell_max, replace=False))
        sel = [tets[ptr[m]:ptr[m + 1]] for m in pick]
        ptr = np.concatenate([[0], np.cumsum([len(s) for s in sel])])
        return ptr.astype(np.int32), np.concatenate(sel).astype(np.int32), cells[pick]
    return ptr.astype(np.int32), tets.astype(np.int32), cells


def seafloor_sites(case, n, seed):
    """Seeded sites 1 m below the sea-bottom plane, strictly inside a tet."""
    rng = np.random.default_rng(seed)
    h = case.h
    xout, yout = case.nx * h / 2, case.ny * h / 2
    pts = []
    for _ in range(n):
        i = rng.integers(1, case.nx - 1)
        j = rng.integers(1, case.ny - 1)
        fx, fy = rng.uniform(0.13, 0.87, size=2)
        if abs(fx - fy) < 0.05:
            fy = min(0.87, fy + 0.1)
        pts.append([-xout + (i + fx) * h, -yout + (j + fy) * h, -h - 0.001])
    return np.array(pts)


def near_site_params(case, sites, nsel, seed, radius=1.0, layers=1):
    """Parameters = crust cells (6 tets each) of the top `layers` crust layers whose centre lies
    within `radius` cells (horizontally) of a site, nsel of them seeded: the cells whose
    conductivity the seafloor fields are most sensitive to. Returns (par_ptr, par_tet, cell ids)."""
    crust = np.nonzero(case.group >= 3)[0]
    cells, inv = np.unique(case.cell_of_tet[crust], return_inverse=True)
    ci, cj, ck = cells % case.nx, (cells // case.nx) % case.ny, cells // (case.nx * case.ny)
    xc = -case.nx * case.h / 2 + (ci + 0.5) * case.h
    yc = -case.ny * case.h / 2 + (cj + 0.5) * case.h
    d = np.min(np.hypot(xc[:, None] - sites[None, :, 0], yc[:, None] - sites[None, :, 1]), axis=1)
    ok = np.nonzero((d <= radius * case.h) & (ck >= ck.max() - layers + 1))[0]
    rng = np.random.default_rng(seed)
    pick = np.sort(rng.choice(ok, min(nsel, len(ok)), replace=False))
    order = np.argsort(inv, kind="stable")
    tets = crust[order]
    ptr = np.concatenate([[0], np.cumsum(np.bincount(inv))])
    sel = [tets[ptr[m]:ptr[m + 1]] for m in pick]
    return (np.concatenate([[0], np.cumsum([len(x) for x in sel])]).astype(np.int32),
            np.concatenate(sel).astype(np.int32), cells[pick])
