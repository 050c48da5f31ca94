"""Pin the oracle's adjoint Jacobian (genjacobian1 restated) by central finite differences and
by a direct-sensitivity solve, with SuperLU solves on the oracle-assembled system."""
import numpy as np
import pytest

from tests.conftest import OMEGA_M2, crust_params, seafloor_sites


def _solve(case, sigma):
    import scipy.sparse.linalg as spla
    from oracle import oracle as O
    o = O.forward_system(case, OMEGA_M2, sigma=sigma)
    A = O.to_scipy(o["INU"], o["IAU"], o["INL"], o["IAL"], o["D"], o["AU"], o["AL"]).tocsc()
    return A, spla.splu(A), o


def _refined_solve(A, lu, b, steps=6):
    """x with A x = b to a forward error far below the system's condition number times eps:
    the factors of a nearby matrix, refinement with residuals in extended precision (x86 80-bit)."""
    Ac = A.tocsr()
    data = Ac.data.astype(np.clongdouble)
    rows = np.repeat(np.arange(Ac.shape[0]), np.diff(Ac.indptr))
    bl = b.astype(np.clongdouble)
    x = lu.solve(b).astype(np.clongdouble)
    for _ in range(steps):
        ax = np.zeros(Ac.shape[0], np.clongdouble)
        np.add.at(ax, rows, data * x[Ac.indices])
        x = x + lu.solve((bl - ax).astype(np.complex128))
    return x


def test_adjoint_jacobian_matches_central_differences(case_tiny):
    from oracle import oracle as O
    c = case_tiny
    sig0 = c.sigma_tet()
    A, lu, o = _solve(c, sig0)
    x = lu.solve(o["b"])
    sites = seafloor_sites(c, 3, 7)
    item, ve, vb, _ = O.prepobscoeff(c, sites)
    rows = [vb[0], vb[2], ve[1]]  # Bx, Bz, Ey(/-i omega) rows
    P = np.zeros((len(rows) * len(sites), c.nline))
    for r, val in enumerate(rows):
        for s in range(len(sites)):
            P[r * len(sites) + s, item[s] - 1] = val[s]
    U = np.array([lu.solve(p.astype(complex)) for p in P])  # A symmetric: u_o = A^-1 P_o^T
    par_ptr, par_tet, _ = crust_params(c, ncell_max=6, seed=3)
    ln10 = np.log(10.0)
    J = O.jacobian(U, x, par_ptr, par_tet, c, OMEGA_M2, ln10)  # d/d(log10 sigma)
    h = 1e-4
    for m in range(len(par_ptr) - 1):
        tets = par_tet[par_ptr[m]:par_ptr[m + 1]]
        d = []
        for sgn in (+1, -1):
            sig = sig0.copy()
            sig[tets] *= 10.0 ** (sgn * h)
            A2, _, o2 = _solve(c, sig)
            # solve the perturbed system with the unperturbed factors plus refinement, so that
            # both sides of the difference carry the same rounding and cancel
            xs = _refined_solve(A2, lu, o2["b"])
            d.append(P.astype(np.longdouble) @ xs)
        fd = ((d[0] - d[1]) / (2 * h)).astype(np.complex128)
        # central differences carry O(h^2 (ln 10)^2) ~ 1e-8 truncation error
        assert np.linalg.norm(J[:, m] - fd) <= 1e-5 * np.linalg.norm(fd)


def test_direct_sensitivity_equals_adjoint(case_tiny):
    """A dx = -(dA/dm) x: P dx equals the adjoint accumulation for every observation."""
    from oracle import oracle as O
    c = case_tiny
    sig0 = c.sigma_tet()
    A, lu, o = _solve(c, sig0)
    x = lu.solve(o["b"])
    sites = seafloor_sites(c, 2, 9)
    item, ve, vb, _ = O.prepobscoeff(c, sites)
    P = np.zeros((2, c.nline))
    for s in range(2):
        P[s, item[s] - 1] = vb[1][s]
    U = np.array([lu.solve(p.astype(complex)) for p in P])
    par_ptr, par_tet, _ = crust_params(c, ncell_max=3, seed=1)
    J = O.jacobian(U, x, par_ptr, par_tet, c, OMEGA_M2, 1.0)  # d/d(ln sigma)
    for m in range(3):
        tets = par_tet[par_ptr[m]:par_ptr[m + 1]]
        sig = np.zeros_like(sig0)
        sig[tets] = sig0[tets]  # dA/d ln(sigma_m) = i omega M(sigma restricted to m)
        INU, IAU, INL, IAL = O.pattern(c.nline, c.n6line)
        Kd, Ku, Md, Mu = O.genmat_split(c, INU, IAU, sigma=sig)
        Ml = O.copy_ul(INU, IAU, INL, IAL, Mu + 0j).real
        dA = 1j * OMEGA_M2 * O.to_scipy(INU, IAU, INL, IAL, Md, Mu, Ml)
        bc = o["bc"]
        dA = dA.tolil()
        dA[np.nonzero(bc)[0], :] = 0
        dA[:, np.nonzero(bc)[0]] = 0
        dx = lu.solve(-(dA.tocsr() @ x))
        assert np.allclose(P @ dx, J[:, m], rtol=1e-9, atol=1e-14)
