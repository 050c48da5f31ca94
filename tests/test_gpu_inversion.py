"""Gauss-Newton updates of the tidal inversion on the device (actfemtide_amd.inversion: the body of
n_inv_ap.f90:258-430 per update), with the forward and adjoint solutions of each update reused
as the next update's initial guesses.

A conductive block (log10 sigma -0.7 in a background of -2, 5x5x3 cells under the top of the
crust) is recovered from Bx, By, Bz amplitude and phase at 8 sites 1 m above the sea bottom for
two constituents (M2, K1) on the 16x16x12 mesh, from the uniform reference model, in 3 updates
(minimum-support stabiliser with beta = 10, alpha = 0.1 x the spectral estimate, cooled):
- every forward and adjoint solve of every update reaches the 1e-8 FP64 true residual;
- the misfit falls with every update;
- the same updates run with cold starts give the same models (the solves differ only within
  the tolerance);
- the warm-started solves of updates 2 and 3 need fewer iterations than the cold ones.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OMEGAS = [2 * math.pi / 44714.0, 2 * math.pi / 86164.0]


def _central_sites(case, n, seed):
    """Seeded sites 1 m above the sea bottom over the central 6x6 cells, strictly inside a tet."""
    rng = np.random.default_rng(seed)
    h = case.h
    pts = []
    for _ in range(n):
        i = case.nx // 2 + rng.integers(-3, 3)
        j = case.ny // 2 + rng.integers(-3, 3)
        fx, fy = rng.uniform(0.13, 0.87, size=2)
        if abs(fx - fy) < 0.05:
            fy = min(0.87, fy + 0.1)
        pts.append([-case.nx * h / 2 + (i + fx) * h, -case.ny * h / 2 + (j + fy) * h, -h + 0.001])
    return np.array(pts)


def _block_model(case, par_cells, value, background):
    nx, ny = case.nx, case.ny
    i, j, k = par_cells % nx, (par_cells // nx) % ny, par_cells // (nx * ny)
    ktop = k.max()
    m = np.full(par_cells.size, background)
    inside = (abs(i - nx // 2) <= 2) & (abs(j - ny // 2) <= 2) & (k >= ktop - 2)
    m[inside] = value
    return m, inside


@pytest.fixture(scope="module")
def runs():
    from actfemtide_amd import inversion, mesh
    case = mesh.tide_case(16, 16, 12, seed=1001)
    sites = _central_sites(case, 8, 77)
    out = {}
    for warm in (True, False):
        gn = inversion.GaussNewton(case, sites, OMEGAS, np.full(len(inversion.crust_parameters(case)[2]), -2.0),
                                   beta=10.0, gamma=0.1, warm=warm)
        m_true, inside = _block_model(case, gn.cells, -0.7, -2.0)
        F, _, _ = gn.forward(m_true)
        gn.d_obs = gn.data(F)
        gn.dX.zero_()
        ups, m_fin, nrms_fin = gn.run(gn.m_ref, updates=3)
        out[warm] = (ups, m_fin, nrms_fin, inside)
        for k, u in enumerate(ups):
            print(f"{'warm' if warm else 'cold'} update {k + 1}: nRMS {u.nrms:.3f}, alpha {u.alpha:.3e}, forward "
                  f"iterations {u.forward_iters}, adjoint iterations {[sum(a) for a in u.adjoint_iters]}, worst true "
                  f"residual {u.worst_true_resid:.2e}, solves {u.solve_seconds:.3f}s of {u.seconds:.3f}s", flush=True)
        print(f"{'warm' if warm else 'cold'}: final nRMS {nrms_fin:.3f}; block mean log10 sigma "
              f"{m_fin[inside].mean():.3f} (true -0.7), outside {m_fin[~inside].mean():.3f}", flush=True)
        gn.close()
    return out


def test_updates_reduce_the_misfit(runs):
    ups, m_fin, nrms_fin, inside = runs[True]
    nr = [u.nrms for u in ups] + [nrms_fin]
    assert all(u.worst_true_resid <= 1.05e-8 for u in ups)
    assert all(b < a for a, b in zip(nr, nr[1:])), nr
    assert nr[-1] < 0.5 * nr[0]
    # the block is where the model moved towards the truth
    assert m_fin[inside].mean() > m_fin[~inside].mean()


def test_warm_and_cold_updates_agree(runs):
    (uw, mw, rw, _), (uc, mc, rc, _) = runs[True], runs[False]
    for a, b in zip(uw, uc):
        assert abs(a.nrms - b.nrms) <= 1e-5 * b.nrms
        assert np.abs(a.model - b.model).max() <= 1e-5 * max(1.0, np.abs(b.model).max())
    assert abs(rw - rc) <= 1e-5 * rc


def test_warm_starts_save_iterations(runs):
    uw, uc = runs[True][0], runs[False][0]
    for k in (1, 2):
        fw, fc = sum(uw[k].forward_iters), sum(uc[k].forward_iters)
        aw, ac = sum(map(sum, uw[k].adjoint_iters)), sum(map(sum, uc[k].adjoint_iters))
        print(f"update {k + 1}: forward {fw} warm / {fc} cold, adjoint {aw} warm / {ac} cold", flush=True)
        assert fw < fc and aw < ac
