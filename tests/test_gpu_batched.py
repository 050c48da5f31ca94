"""Batched, multi-frequency and mixed-precision solves on the GPU against the oracle."""
import math

import numpy as np
import pytest

from tests.conftest import OMEGA_M2

pytestmark = pytest.mark.gpu

OMEGAS = np.array([2 * math.pi / p for p in (44714.0, 43200.0, 45570.0, 86164.0, 92950.0)])


@pytest.fixture(scope="module")
def sysdev(case_small):
    import torch
    from actfemtide_amd import afx
    s, b, bc = afx.forward_system(case_small, OMEGA_M2)
    return s, b, bc, torch.device("cuda:0")


def _oracle_matrix(case, omega):
    from oracle import oracle as O
    o = O.forward_system(case, omega)
    return O.to_scipy(o["INU"], o["IAU"], o["INL"], o["IAL"], o["D"], o["AU"], o["AL"]), o["b"]


def _tidal_rhs(case, seed):
    """The tidal right-hand side of the same mesh for another seeded velocity field."""
    import dataclasses
    from actfemtide_amd import mesh
    from oracle import oracle as O
    v = np.zeros_like(case.vxyz)
    no = case.nodes_ocean
    v[:no] = mesh.fourier_velocity(case.xyz[:no], seed, lx=case.nx * case.h, ly=case.ny * case.h)
    c2 = dataclasses.replace(case, vxyz=v)
    o = O.forward_system(c2, OMEGA_M2)
    return o["b"]


BATCH_MODES = ["columns", "multi", "stack"]


@pytest.mark.parametrize("mode", BATCH_MODES)
def test_batched_columns(case_small, sysdev, mode, monkeypatch):
    """4 independent tidal right-hand sides, one column at a time (the default block width 1), in
    one k-column persistent launch (AFX_BATCH=multi) or as 4 stacked copies of the system in one
    launch (AFX_BATCH=stack): each column reaches 1e-8."""
    import torch
    if mode != "columns":
        monkeypatch.setenv("AFX_BATCH", mode)
    s, b, bc, dev = sysdev
    B = np.stack([b, (1 + 2j) * b, _tidal_rhs(case_small, 5), _tidal_rhs(case_small, 6)], axis=1)
    dB = torch.from_numpy(B).to(dev).contiguous()
    dX = torch.zeros_like(dB)
    st, rc = s.solve_device(dB, dX, precond="SSOR", tol=1e-8, maxit=40000)
    A, _ = _oracle_matrix(case_small, OMEGA_M2)
    X = dX.cpu().numpy()
    for c in range(4):
        assert st[c].error == 0
        assert np.linalg.norm(B[:, c] - A @ X[:, c]) / np.linalg.norm(B[:, c]) <= 1.05e-8
    monkeypatch.delenv("AFX_BATCH", raising=False)
    x1, st1 = s.solve(b, precond="SSOR", tol=1e-8, maxit=40000)
    if mode == "stack":
        # every copy runs its own COCG13_single recurrence with per-system scalars: each column's
        # count is the oracle's (the reference's order) for that right-hand side, up to rounding
        from oracle import oracle as O
        o = O.forward_system(case_small, OMEGA_M2)
        for c in range(4):
            r = O.cocg13(o["INU"], o["IAU"], o["INL"], o["IAL"], o["D"], o["AU"], o["AL"], B[:, c], precond="SSOR",
                         resid=1e-8, maxit=40000)
            print(f"stacked copy {c}: {st[c].iter} iterations, oracle {r['iter']}", flush=True)
            assert st[c].resid <= 1e-8 and abs(st[c].iter - r["iter"]) <= max(3, 0.08 * r["iter"])
    else:
        assert abs(st[0].iter - st1[0].iter) <= max(3, 0.08 * st1[0].iter)
    if mode == "columns":  # the same kernel on the same column: the same iterates
        assert st[0].iter == st1[0].iter and np.array_equal(X[:, 0], x1)


def test_autotuned_block_width(case_small, sysdev, monkeypatch):
    """afx_autotune for 3 columns times the three block widths (one column at a time, the k-column
    kernel, 3 stacked copies) and keeps the fastest; the solve that follows (whichever width it
    uses) reaches 1e-8 in every column."""
    import torch
    monkeypatch.delenv("AFX_BATCH", raising=False)
    s, b, bc, dev = sysdev
    s.autotune(3)
    B = np.stack([b, (0.5 - 1j) * b, _tidal_rhs(case_small, 8)], axis=1)
    dB = torch.from_numpy(B).to(dev).contiguous()
    dX = torch.zeros_like(dB)
    st, rc = s.solve_device(dB, dX, precond="SSOR", tol=1e-8, maxit=40000)
    A, _ = _oracle_matrix(case_small, OMEGA_M2)
    X = dX.cpu().numpy()
    for c in range(3):
        assert st[c].error == 0
        assert np.linalg.norm(B[:, c] - A @ X[:, c]) / np.linalg.norm(B[:, c]) <= 1.05e-8


@pytest.mark.parametrize("mode", BATCH_MODES)
def test_multi_frequency_columns(case_small, sysdev, mode, monkeypatch):
    """5 tidal constituents as 5 columns with their own omega (K + i w M on the fly, or 5 stacked
    copies each holding its omega's values)."""
    import torch
    if mode != "columns":
        monkeypatch.setenv("AFX_BATCH", mode)
    s, b, bc, dev = sysdev
    dB = torch.from_numpy(np.repeat(b[:, None], 5, axis=1)).to(dev).contiguous()
    dX = torch.zeros_like(dB)
    st, rc = s.solve_device(dB, dX, omega_scale=OMEGAS / OMEGA_M2, precond="SSOR", tol=1e-8, maxit=40000)
    X = dX.cpu().numpy()
    for c, w in enumerate(OMEGAS):
        A, bw = _oracle_matrix(case_small, w)
        assert np.allclose(bw, b, rtol=1e-12, atol=0)  # the tidal source does not depend on omega
        r = np.linalg.norm(b - A @ X[:, c]) / np.linalg.norm(b)
        print("omega", w, "iters", st[c].iter, "true", r, flush=True)
        assert st[c].error == 0 and r <= 1.05e-8


def test_stacked_columns_in_groups(case_small, sysdev, monkeypatch):
    """5 constituents stacked at most 2 copies per launch (AFX_STACK_GROUP=2): groups of 2, 2 and
    1 padded with a zero right-hand side; every column reaches 1e-8 at its own omega, and the
    copies are rebuilt for each group's frequencies."""
    import torch
    monkeypatch.setenv("AFX_BATCH", "stack")
    monkeypatch.setenv("AFX_STACK_GROUP", "2")
    s, b, bc, dev = sysdev
    dB = torch.from_numpy(np.repeat(b[:, None], 5, axis=1)).to(dev).contiguous()
    dX = torch.zeros_like(dB)
    st, rc = s.solve_device(dB, dX, omega_scale=OMEGAS / OMEGA_M2, precond="SSOR", tol=1e-8, maxit=40000)
    X = dX.cpu().numpy()
    print("stacked groups: iterations", [x.iter for x in st], flush=True)
    for c, w in enumerate(OMEGAS):
        A, _ = _oracle_matrix(case_small, w)
        r = np.linalg.norm(b - A @ X[:, c]) / np.linalg.norm(b)
        assert st[c].error == 0 and st[c].iter > 0 and r <= 1.05e-8


def test_mixed_precision_refinement(case_small, sysdev):
    """COCG with a complex-FP32 preconditioner (FP64 SpMV, recurrences, true residual) reaches the
    FP64 true residual 1e-8, and its seafloor fields agree with the all-FP64 oracle solution."""
    import torch
    from actfemtide_amd import afx
    s, b, bc, dev = sysdev
    dB = torch.from_numpy(b).to(dev)
    dX = torch.zeros_like(dB)
    st, rc = s.solve_device(dB, dX, precond="SSOR", tol=1e-8, maxit=60000, precision=afx.MIXED)
    A, _ = _oracle_matrix(case_small, OMEGA_M2)
    x = dX.cpu().numpy()
    r = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    print("mixed: iters", st[0].iter, "restarts", st[0].outer, "true", r, flush=True)
    assert rc == 0 and r <= 1.05e-8
    # error against the all-FP64 oracle solution at the seafloor fields
    from oracle import oracle as O
    from tests.conftest import seafloor_sites
    o = O.forward_system(case_small, OMEGA_M2)
    ref = O.cocg13(o["INU"], o["IAU"], o["INL"], o["IAL"], o["D"], o["AU"], o["AL"], o["b"], precond="SSOR",
                   resid=1e-10, maxit=40000)
    item, ve, vb, _ = O.prepobscoeff(case_small, seafloor_sites(case_small, 8, 3))
    for c in range(3):
        f = O.obs_apply(item, vb[c], x)
        fr = O.obs_apply(item, vb[c], ref["x"])
        assert np.max(np.abs(f - fr) / np.abs(fr)) <= 1e-6


def test_warm_start_from_neighbouring_constituent(case_small, sysdev):
    """S2 started from the M2 solution converges in fewer iterations than from zero."""
    import torch
    s, b, bc, dev = sysdev
    dB = torch.from_numpy(b).to(dev)
    dX = torch.zeros_like(dB)
    st0, _ = s.solve_device(dB, dX, precond="SSOR", tol=1e-8, maxit=40000)  # M2
    sc = np.array([OMEGAS[1] / OMEGA_M2])
    dXs = torch.zeros_like(dB)
    st_cold, _ = s.solve_device(dB, dXs, omega_scale=sc, precond="SSOR", tol=1e-8, maxit=40000)
    dXw = dX.clone()
    st_warm, _ = s.solve_device(dB, dXw, omega_scale=sc, precond="SSOR", tol=1e-8, maxit=40000)
    print("S2 cold", st_cold[0].iter, "warm", st_warm[0].iter, flush=True)
    assert st_warm[0].error == 0 and st_warm[0].iter < st_cold[0].iter
    assert s.true_residual_device(dB, dXw, omega_scale=sc)[0] <= 1.05e-8


def test_observation_operator(case_small, sysdev):
    import torch
    from actfemtide_amd import afx
    from oracle import oracle as O
    from tests.conftest import seafloor_sites
    s, b, bc, dev = sysdev
    sites = seafloor_sites(case_small, 16, 5)
    item, ve, vb, iele = s.prepobscoeff(sites)
    item_r, ve_r, vb_r, iele_r = O.prepobscoeff(case_small, sites)
    assert np.array_equal(item, item_r) and np.array_equal(iele, iele_r)
    assert np.allclose(ve, ve_r, rtol=1e-13, atol=1e-15) and np.allclose(vb, vb_r, rtol=1e-13, atol=1e-15)
    rng = np.random.default_rng(8)
    X = rng.standard_normal((s.n, 3)) + 1j * rng.standard_normal((s.n, 3))
    f = afx.calobsebcomp(s, item, ve, vb, torch.from_numpy(X).to(dev).contiguous(), OMEGA_M2)
    for c in range(3):
        assert np.allclose(f[0, :, c], O.obs_apply(item_r, vb_r[0], X[:, c]), rtol=1e-13)
        assert np.allclose(f[3, :, c], -1j * OMEGA_M2 * O.obs_apply(item_r, ve_r[0], X[:, c]), rtol=1e-13)


def test_jacobian_matches_oracle(case_tiny):
    """Device genjacobian1 accumulation equals the oracle's on the same U and x."""
    import torch
    import scipy.sparse.linalg as spla
    from actfemtide_amd import afx
    from oracle import oracle as O
    from tests.conftest import crust_params, seafloor_sites
    c = case_tiny
    s, b, bc = afx.forward_system(c, OMEGA_M2)
    o = O.forward_system(c, OMEGA_M2)
    A = O.to_scipy(o["INU"], o["IAU"], o["INL"], o["IAL"], o["D"], o["AU"], o["AL"]).tocsc()
    lu = spla.splu(A)
    x = lu.solve(o["b"])
    item, ve, vb, _ = O.prepobscoeff(c, seafloor_sites(c, 3, 7))
    P = np.zeros((6, c.nline))
    for k in range(3):
        P[k, item[k] - 1] = vb[0][k]
        P[3 + k, item[k] - 1] = ve[1][k]
    U = np.array([lu.solve(p.astype(complex)) for p in P])
    par_ptr, par_tet, _ = crust_params(c)
    J_ref = O.jacobian(U, x, par_ptr, par_tet, c, OMEGA_M2, np.log(10.0))
    dev = torch.device("cuda:0")
    dU = torch.from_numpy(np.ascontiguousarray(U.T)).to(dev)
    dx = torch.from_numpy(x).to(dev)
    dJ = torch.zeros((len(par_ptr) - 1, U.shape[0]), dtype=torch.complex128, device=dev)
    s.jacobian_device(dU, dx, par_ptr, par_tet, c.sigma_tet(), OMEGA_M2, np.log(10.0), dJ)
    J = dJ.cpu().numpy().T
    assert np.allclose(J, J_ref, rtol=1e-12, atol=1e-14 * np.abs(J_ref).max())
