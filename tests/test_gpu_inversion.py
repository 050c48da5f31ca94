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
The Jacobian rows the updates use (log10 amplitude and phase in degrees of the B components,
genjacobian1 :338-345, from the adjoint columns) are checked against central differences of
tight forward solves.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OMEGAS = [2 * math.pi / 44714.0, 2 * math.pi / 86164.0]


@pytest.fixture(scope="module")
def runs():
    from actfemtide_amd import inversion, mesh
    case = mesh.tide_case(16, 16, 12, seed=1001)
    sites = inversion.central_sites(case, 8, 77)
    out = {}
    for warm in (True, False):
        gn = inversion.GaussNewton(case, sites, OMEGAS, np.full(len(inversion.crust_parameters(case)[2]), -2.0),
                                   beta=10.0, gamma=0.1, warm=warm)
        m_true, inside = inversion.block_model(case, gn.cells, -0.7, -2.0)
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


def test_ampphase_jacobian_rows_against_central_differences():
    """The rows of the first update's Jacobian for the 4 most sensitive crust cells and one weak
    one, against central differences (step 1e-4 in log10 sigma) of forward solves at a 2e-11 true
    residual: within 2e-4 relative for the sensitive cells (the step's truncation error is about
    5e-5 there), 1e-2 for the weak one (whose difference quotient carries the solves' noise).
    FP64 leaves the true residual of these solves near 1e-11, so they are asked for 2e-11."""
    from actfemtide_amd import afx, inversion, mesh
    case = mesh.tide_case(16, 16, 12, seed=1001)
    sites = inversion.central_sites(case, 8, 77)
    gn = inversion.GaussNewton(case, sites, OMEGAS, np.full(len(inversion.crust_parameters(case)[2]), -2.0),
                               tol=2e-11)
    m0 = gn.m_ref.copy()
    F0, _, _ = gn.forward(m0)
    gn.d_obs = gn.data(F0)
    captured = {}
    real_update = afx.gn_update

    def capture(dJ, *a, **k):
        captured["J"] = dJ.cpu().numpy().copy()
        return real_update(dJ, *a, **k)

    afx.gn_update = capture
    try:
        gn.update(m0)
    finally:
        afx.gn_update = real_update
    J = captured["J"]
    order = np.argsort(-np.abs(J).sum(axis=0))
    h, nob = 1e-4, gn.nobs
    for rank, j in enumerate(list(order[:4]) + [order[200]]):
        dp, dm = m0.copy(), m0.copy()
        dp[j] += h
        dm[j] -= h
        Fp, _, _ = gn.forward(dp)
        Fm, _, _ = gn.forward(dm)
        diff = gn.data(Fp) - gn.data(Fm)
        for f in range(gn.nf):
            k0 = 2 * nob * f + nob
            diff[k0:k0 + nob] = inversion.wrap_degrees(diff[k0:k0 + nob])
        fd = diff / (2 * h)
        err = np.abs(J[:, j] - fd).max() / np.abs(fd).max()
        print(f"cell {gn.cells[j]}: max |row| {np.abs(fd).max():.3e}, adjoint vs central difference {err:.1e}", flush=True)
        assert err <= (2e-4 if rank < 4 else 1e-2)
    gn.close()
