"""The benchmark configurations at their own workloads on the GPU (the sizes the parity tests
elsewhere check at small scale):

- 64x64x48 cells (1,407,152 edges), one M2 right-hand side: the complex-FP32 preconditioner
  (AFX_MIXED) to a 1e-8 FP64 true residual, its site fields against the FP64 solve; the corrected
  forms against a 1e-12 solve;
- 32x32x24 cells (179,800 edges): the 5 tidal constituents' forward solves and, for each
  constituent, the 48 adjoint columns of 16 seafloor sites x 3 field components (240 adjoint
  solves), every column at a recomputed 1e-8 true residual and the adjoint identity
  u^T b = P x holding to the residual bound;
- the uneven batch of 16 meshes from 8x8x6 to 40x40x30 cells in one per-system launch, every part
  at a 1e-8 true residual;
- 64x64x48 cells: genjacobian1's accumulation over all 180,224 crust cells against the oracle's on
  the same U and x; on the 32x32x24 mesh, the adjoint Jacobian of 20 seeded crust cells under the
  sites against direct sensitivities (to 1e-5) and central differences (step 1e-4 in log10 sigma,
  to 1e-5 where FP64 solves resolve the difference).
"""
import math

import numpy as np
import pytest

from tests.conftest import crust_params, seafloor_sites

pytestmark = pytest.mark.gpu

OMEGAS = {c: 2 * math.pi / p for c, p in (("M2", 44714.0), ("S2", 43200.0), ("N2", 45570.0), ("K1", 86164.0),
                                        ("O1", 92950.0))}
W0 = OMEGAS["M2"]


def _fields(s, item, ve, vb, x):
    """B (3 components) and E (3 components) at the sites for every column of device x."""
    return [s.obs_apply_device(item, v, x) for v in list(vb) + list(ve)]


def _worst(f, g):
    return max(float(np.abs(a - b).max() / np.abs(b).max()) for a, b in zip(f, g))


@pytest.fixture(scope="module")
def cfg2():
    from actfemtide_amd import afx, mesh
    c = mesh.tide_case(64, 64, 48, seed=1002, spheres=3, sigma_sea=3.3)
    print(f"[cfg2] 64x64x48 system: nline={c.nline}", flush=True)
    s, b, bc = afx.forward_system(c, W0)
    yield c, s, b
    s.close()


def test_cfg2_forms_against_a_tight_solve(cfg2):
    """Every measured form at a 1e-8 FP64 true residual has seafloor site fields within 1e-6 of a
    1e-12 solve of the same system: the bench's NODEBLK+GRAD (bfloat16 block coefficients) and the
    other forms, AFX_MIXED (FP32 edge preconditioner) included."""
    import torch
    from actfemtide_amd import afx
    c, s, b = cfg2
    dB = torch.from_numpy(b).cuda()
    item, ve, vb, _ = s.prepobscoeff(seafloor_sites(c, 96, 31))
    s.precond_setup("SSOR+GRAD")
    dT = torch.zeros_like(dB)
    st, rc = s.solve_device(dB, dT, precond="SSOR+GRAD", tol=1e-12, maxit=20000)
    tight = _fields(s, item, ve, vb, dT)
    print(f"tight SSOR+GRAD: {st[0].iter} iterations, true residual {st[0].true_resid:.2e}", flush=True)
    assert st[0].true_resid <= 1e-11
    for kind, prec in (("NODEBLK+GRAD", afx.FP64), ("DIAG+GRAD", afx.FP64), ("SSOR_BJ+GRAD", afx.FP64),
                       ("SSOR+GRAD", afx.MIXED), ("DIAG+GRAD", afx.MIXED)):
        s.precond_setup(kind)
        dX = torch.zeros_like(dB)
        st, rc = s.solve_device(dB, dX, precond=kind, precision=prec, tol=1e-8, maxit=20000, max_outer=8)
        tr = float(s.true_residual_device(dB, dX)[0])
        err = _worst(_fields(s, item, ve, vb, dX), tight)
        print(f"{kind} {'FP64' if prec == afx.FP64 else 'mixed'}: {st[0].iter} iterations ({st[0].outer + 1} runs), "
              f"true residual {tr:.2e}, site fields vs the 1e-12 solve {err:.1e}", flush=True)
        assert rc == 0 and tr <= 1.05e-8 and err <= 1e-6


@pytest.fixture(scope="module")
def cfg3():
    from actfemtide_amd import afx, mesh
    c = mesh.tide_case(32, 32, 24, seed=1003, spheres=3, sigma_sea=3.3)
    print(f"[cfg3] 32x32x24 system: nline={c.nline}", flush=True)
    s, b, bc = afx.forward_system(c, W0)
    yield c, s, b
    s.close()


def test_cfg3_constituents_and_adjoint_columns(cfg3):
    """5 constituents x (1 forward + 48 adjoint) solves with DIAG+GRAD, several columns per
    launch; each column's true residual recomputed <= 1e-8; u_o^T b = P_o x within the bound
    |u_o|.|b - A x| + |P_o^T - A u_o|.|x| that the two residuals fix."""
    import torch
    from actfemtide_amd import afx
    c, s, b = cfg3
    n = s.n
    om = np.array(list(OMEGAS.values()))
    k = len(om)
    s.precond_setup("DIAG+GRAD")
    s.autotune(k)
    dB = torch.from_numpy(np.ascontiguousarray(np.tile(b[:, None], (1, k)))).cuda()
    dX = torch.zeros_like(dB)
    st, rc = s.solve_device(dB, dX, omega_scale=om / W0, precond="DIAG+GRAD", tol=1e-8, maxit=20000)
    trx = s.true_residual_device(dB, dX, omega_scale=om / W0)
    print(f"forward, 5 constituents: iterations {[x.iter for x in st]}, true residuals max {trx.max():.2e}", flush=True)
    assert rc == 0 and np.all(trx <= 1.05e-8)
    # adjoint right-hand sides: the E rows (x, y, z) of 16 seafloor sites, solvePARDISOinv's P^T
    item, ve, vb, _ = s.prepobscoeff(seafloor_sites(c, 16, 41))
    stack = np.arange(0, 6 * 16 + 1, 6, dtype=np.int32)
    cols = []
    for comp in range(3):
        d = torch.zeros((n, 16), dtype=torch.complex128, device="cuda")
        s.adjoint_rhs_device(stack, item, ve[comp], d)
        cols.append(d)
    dP = torch.cat(cols, dim=1).contiguous()  # (n, 48)
    s.autotune(48)
    worst_tr, worst_id = 0.0, 0.0
    for ci, w in enumerate(om):
        dU = torch.zeros_like(dP)
        sc = np.full(48, w / W0)
        st, rc = s.solve_device(dP, dU, omega_scale=sc, precond="DIAG+GRAD", tol=1e-8, maxit=20000)
        tru = s.true_residual_device(dP, dU, omega_scale=sc)
        assert rc == 0 and np.all(tru <= 1.05e-8)
        worst_tr = max(worst_tr, float(tru.max()))
        # adjoint identity on the host in FP64
        s.set_omega(float(w))
        U = dU.cpu().numpy()                      # (n, 48)
        P = dP.cpu().numpy()
        x = dX[:, ci].contiguous()
        Ax = torch.empty_like(x)
        s.spmv_device(x, Ax)
        rx = b - Ax.cpu().numpy()
        xh = x.cpu().numpy()
        for o in range(48):
            uo = U[:, o].copy()
            duo = torch.from_numpy(uo).cuda()
            Au = torch.empty_like(duo)
            s.spmv_device(duo, Au)
            ru = P[:, o] - Au.cpu().numpy()
            lhs = uo @ b
            px = P[:, o] @ xh
            bound = np.linalg.norm(uo) * np.linalg.norm(rx) + np.linalg.norm(ru) * np.linalg.norm(xh)
            assert abs(lhs - px) <= 1.01 * bound + 1e-13 * abs(px)
            worst_id = max(worst_id, abs(lhs - px) / abs(px))
        s.set_omega(W0)
        print(f"constituent {ci}: 48 adjoint columns, iterations {min(x.iter for x in st)}..{max(x.iter for x in st)},"
              f" true residual max {tru.max():.2e}", flush=True)
    print(f"240 adjoint solves: worst true residual {worst_tr:.2e}, worst |u^T b - P x| / |P x| {worst_id:.1e}")


@pytest.mark.parametrize("pc", ["DIAG+GRAD", "NODEBLK+GRAD"])
def test_cfg3_sixteen_meshes_one_launch(pc):
    """The 16 meshes of 8x8x6 to 40x40x30 cells (seeds 1004 + i, each its own crust) stacked and
    solved in one per-system launch with DIAG+GRAD and with NODEBLK+GRAD (the node blocks of the
    stack are each part's own): every part at a 1e-8 true residual, its iterations within a few of
    its own single solve."""
    import torch
    from actfemtide_amd import afx, mesh
    from actfemtide_amd.meshbatch import MeshBatch
    dims = [(8 + round(32 * i / 15), 8 + round(32 * i / 15), round(0.75 * (8 + round(32 * i / 15)))) for i in range(16)]
    parts, bs, cases, bcs = [], [], [], []
    for i, d in enumerate(dims):
        c = mesh.tide_case(*d, seed=1004 + i)
        s, b, bc = afx.forward_system(c, W0)
        parts.append(s)
        bs.append(b)
        cases.append(c)
        bcs.append(bc)
    own = []
    for s, b in zip(parts, bs):
        s.precond_setup(pc)
        x, st = s.solve(b, precond=pc, tol=1e-8, maxit=40000)
        own.append(st[0].iter)
    mb = MeshBatch(parts, per_system=True)
    mb.set_nodes([c.node for c in cases], [c.line for c in cases], bcs)
    mb.system.precond_setup(pc)
    print(f"stack: {mb.m} systems, {mb.n} edges", flush=True)
    xs, st, tr = mb.solve(bs, tol=1e-8, precond=pc, maxit=40000)
    its = [x.iter for x in st]
    print(f"{pc} one launch: iterations {its}; own solves {own}; worst part residual {tr.max():.2e}", flush=True)
    assert np.all(tr <= 1.05e-8) and all(x.error == 0 for x in st)
    for a, bb in zip(its, own):
        assert abs(a - bb) <= max(3, 0.05 * bb)
    # each part's solution also meets the tolerance on the part's own system
    for s, b, x in zip(parts, bs, xs):
        r = s.true_residual_device(torch.from_numpy(b).cuda(), torch.from_numpy(x).cuda())
        assert float(r[0]) <= 1.05e-8
    mb.close()
    for s in parts:
        s.close()


def test_cfg5_jacobian_accumulation_all_crust_cells(cfg2):
    """genjacobian1's accumulation (afx_jacobian_device) over every crust cell of the 64x64x48 mesh
    (180,224 parameters of 6 tets) for 24 columns of U, against the oracle on the same U and x."""
    import torch
    from oracle import oracle as O
    c, s, b = cfg2
    par_ptr, par_tet, cells = crust_params(c)
    assert len(par_ptr) - 1 == 180224
    rng = np.random.default_rng(12)
    nobs = 24
    U = (rng.standard_normal((nobs, s.n)) + 1j * rng.standard_normal((nobs, s.n))) * 1e-3
    x = rng.standard_normal(s.n) + 1j * rng.standard_normal(s.n)
    dU = torch.from_numpy(np.ascontiguousarray(U.T)).cuda()
    dx = torch.from_numpy(x).cuda()
    dJ = torch.zeros((len(par_ptr) - 1, nobs), dtype=torch.complex128, device="cuda")
    s.jacobian_device(dU, dx, par_ptr, par_tet, c.sigma_tet(), W0, np.log(10.0), dJ)
    J = dJ.cpu().numpy().T
    J_ref = O.jacobian(U, x, par_ptr, par_tet, c, W0, np.log(10.0))
    err = np.abs(J - J_ref).max() / np.abs(J_ref).max()
    print(f"180224 parameters x {nobs} columns: max |J - J_oracle| / max |J| = {err:.1e}", flush=True)
    assert err <= 1e-12


def test_cfg5_jacobian_against_direct_sensitivities_and_differences(cfg3):
    """d(E_x, E_y, B_z at 4 sites)/d log10 sigma of 20 seeded crust cells under the sites, by the
    device adjoint (tight DIAG+GRAD solves for u_o = A^-1 P_o^T, afx_jacobian_device), against
    - the direct sensitivities P dx with A dx = -(dA/dm) x, dA/dm = i omega ln10 M(sigma on cell m)
      (the mass matrix from the oracle's GENMAT split, the solves on the device): to 1e-5;
    - central differences with step 1e-4 in log10 sigma of tight device forward solves (true
      residual ~3e-12): the solves leave about 1e-12 of relative noise in the site fields, which
      the difference quotient turns into 1e-12 / (2e-4 * sensitivity), sensitivity = max |dF/dm|
      / max |F|. So: to 1e-5 for every parameter with sensitivity >= 5e-4 (at least 5 of them),
      and to 1e-5 + 2e-12 / (2e-4 * sensitivity) for the others."""
    import scipy.sparse as sp
    import torch
    from oracle import oracle as O
    from tests.conftest import near_site_params
    c, s, b = cfg3
    n = s.n
    sig0 = c.sigma_tet()
    sites = seafloor_sites(c, 4, 5)
    item, ve, vb, _ = s.prepobscoeff(sites)
    rows = [ve[0], ve[1], vb[2]]
    P = np.zeros((12, n))
    for r, val in enumerate(rows):
        for o in range(4):
            P[4 * r + o, item[o] - 1] += val[o]
    xo, yo, zmin, zmax = c.bounds
    bc = s.genbc(c.line, zmin, zmax, xo, yo)

    def solve(sig, rhs_cols=None, floor=1e-11):
        bb = s.assemble(sig, W0)
        bb = s.set_bc(bc, bb)
        s.precond_setup("DIAG+GRAD")
        cols = bb[:, None] if rhs_cols is None else rhs_cols
        dB = torch.from_numpy(np.ascontiguousarray(cols.astype(np.complex128))).cuda()
        dX = torch.zeros_like(dB)
        st, rc = s.solve_device(dB, dX, precond="DIAG+GRAD", tol=1e-12, maxit=40000, max_outer=3)
        tr = max(x.true_resid for x in st)
        assert tr <= floor, tr
        out = dX.cpu().numpy()
        return out[:, 0] if rhs_cols is None else out

    x = solve(sig0)
    U = solve(sig0, P.T)                          # u_o = A^-1 P_o^T (A symmetric)
    par_ptr, par_tet, cells = near_site_params(c, sites, 20, 20, radius=1.0, layers=2)
    npar = len(par_ptr) - 1
    assert npar == 20
    dJ = torch.zeros((npar, 12), dtype=torch.complex128, device="cuda")
    s.jacobian_device(torch.from_numpy(np.ascontiguousarray(U)).cuda(), torch.from_numpy(x).cuda(), par_ptr, par_tet,
                      sig0, W0, np.log(10.0), dJ)
    J = dJ.cpu().numpy()
    # direct sensitivities: -(dA/dm) x for every parameter, the Dirichlet rows and columns removed
    INU, IAU, INL, IAL = O.pattern(c.nline, c.n6line)
    rhs = np.zeros((n, npar), np.complex128)
    keep = sp.diags((~np.asarray(bc, bool)).astype(float))
    for m in range(npar):
        sg = np.zeros_like(sig0)
        t = par_tet[par_ptr[m]:par_ptr[m + 1]]
        sg[t] = sig0[t]
        Kd, Ku, Md, Mu = O.genmat_split(c, INU, IAU, sigma=sg)
        Ml = O.copy_ul(INU, IAU, INL, IAL, Mu + 0j).real
        dA = 1j * W0 * np.log(10.0) * (keep @ O.to_scipy(INU, IAU, INL, IAL, Md, Mu, Ml) @ keep)
        rhs[:, m] = -(dA @ x)
    # a right-hand side confined to one cell is small against the terms of A x that cancel in
    # it, so FP64 leaves its true residual near 1e-9 (|A| |x| eps / |rhs|)
    DX = solve(sig0, rhs, floor=1e-8)
    dmax = 0.0
    for m in range(npar):
        ds = P @ DX[:, m]
        err = np.abs(J[m] - ds).max() / np.abs(ds).max()
        dmax = max(dmax, err)
        assert err <= 1e-5, (m, err)
    # central differences
    h = 1e-4
    f0 = np.abs(P @ x).max()
    checked, fmax = 0, 0.0
    for m in range(npar):
        tets = par_tet[par_ptr[m]:par_ptr[m + 1]]
        f = []
        for sgn in (1, -1):
            sg = sig0.copy()
            sg[tets] *= 10.0 ** (sgn * h)
            f.append(P @ solve(sg))
        fd = (f[0] - f[1]) / (2 * h)
        sens = np.abs(fd).max() / f0
        err = np.abs(J[m] - fd).max() / np.abs(fd).max()
        print(f"parameter {m} (cell {cells[m]}): sensitivity {sens:.1e}, adjoint vs central difference {err:.1e}",
              flush=True)
        if sens >= 5e-4:
            checked += 1
            fmax = max(fmax, err)
            assert err <= 1e-5, (m, err)
        else:
            assert err <= 1e-5 + 2e-12 / (2 * h * sens), (m, err, sens)
    print(f"{npar} parameters: worst adjoint vs direct sensitivity {dmax:.1e}; {checked} checked by central "
          f"differences, worst {fmax:.1e}", flush=True)
    assert checked >= 5
    solve(sig0)  # leave the shared system at its own conductivity
