"""Gauss-Newton model update (getnewmodel + solvebeta, m_jacobian_ap.f90:623-787) and the real
Jacobian rows of genjacobian1 (:338-345) on the GPU, against their closed forms in numpy."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_dgemm_nt_fp64_mfma():
    import torch
    from actfemtide_amd import afx
    rng = np.random.default_rng(0)
    for M, N, K in [(70, 45, 130), (128, 64, 4), (5, 200, 1001)]:
        A = rng.standard_normal((M, K))
        B = rng.standard_normal((N, K))
        C0 = rng.standard_normal((M, N))
        dA = torch.from_numpy(A).cuda()
        dB = torch.from_numpy(B).cuda()
        dC = torch.from_numpy(C0.copy()).cuda()
        afx.dgemm_nt_device(dA, dB, dC, accumulate=True)
        ref = C0 + A @ B.T
        err = np.abs(dC.cpu().numpy() - ref).max() / np.abs(ref).max()
        print(f"dgemm {M}x{N}x{K}: {err:.2e}")
        assert err < 1e-13


def test_jacobian_ampphase_rows():
    import torch
    from actfemtide_amd import afx
    rng = np.random.default_rng(1)
    npar, nobs = 37, 9
    J = rng.standard_normal((npar, nobs)) + 1j * rng.standard_normal((npar, nobs))
    be = rng.standard_normal(nobs) + 1j * rng.standard_normal(nobs)
    dout = torch.zeros((2, nobs, npar), dtype=torch.float64, device="cuda")
    afx.jacobian_ampphase_device(torch.from_numpy(J).cuda(), torch.from_numpy(be).cuda(), dout)
    q = (J / be[None, :]).T
    out = dout.cpu().numpy()
    assert np.allclose(out[0], q.real / np.log(10.0), rtol=1e-13, atol=1e-15)
    assert np.allclose(out[1], q.imag * 180.0 / np.pi, rtol=1e-13, atol=1e-13)


def test_gn_update_matches_closed_form():
    import scipy.sparse as sp
    import torch
    from actfemtide_amd import afx
    rng = np.random.default_rng(2)
    ndat, nmodel = 150, 700
    J = rng.standard_normal((ndat, nmodel)) * 0.1
    # a symmetric positive definite, sparse model covariance (a smoothing stencil on a 1-D chain)
    main = np.full(nmodel, 2.0)
    off = np.full(nmodel - 1, -0.45)
    Cm = sp.diags([off, main, off], [-1, 0, 1], format="csr")
    cd = rng.uniform(0.5, 2.0, ndat)
    alpha = 0.3
    dobs = rng.standard_normal(ndat)
    dcal = rng.standard_normal(ndat)
    m = rng.standard_normal(nmodel)
    mref = rng.standard_normal(nmodel)
    new, beta = afx.gn_update(torch.from_numpy(J).cuda(), Cm, cd, alpha, dobs, dcal, m, mref)
    Cd = Cm.toarray()
    C = J @ Cd @ J.T + alpha * np.diag(cd)
    X = dobs - dcal + J @ (m - mref)
    beta_ref = np.linalg.solve(C, X)
    new_ref = mref + Cd @ J.T @ beta_ref
    eb = np.abs(beta - beta_ref).max() / np.abs(beta_ref).max()
    em = np.abs(new - new_ref).max() / np.abs(new_ref).max()
    print(f"gn update: beta {eb:.2e}, model {em:.2e}")
    assert eb < 1e-10 and em < 1e-10


def test_gn_update_rejects_singular_system():
    import scipy.sparse as sp
    import torch
    from actfemtide_amd import afx
    ndat, nmodel = 4, 6
    J = np.zeros((ndat, nmodel))
    with pytest.raises(afx.AfxError):
        afx.gn_update(torch.from_numpy(J).cuda(), sp.identity(nmodel, format="csr"), np.zeros(ndat), 0.0,
                      np.ones(ndat), np.zeros(ndat), np.zeros(nmodel), np.zeros(nmodel))


def test_spectral_alpha_matches_oracle():
    """alphaspectralradius_v2 + LANCZOSMINMAX (m_spectral.f90:63-179): J Cm J^T on FP64 MFMA, the
    Lanczos recurrence and the tridiagonal's extreme eigenvalues (Sturm bisection), against the
    oracle (numpy, LAPACK eigenvalues), for short and full-length recurrences."""
    import scipy.sparse as sp
    import torch
    from actfemtide_amd import afx
    from oracle import oracle as O
    rng = np.random.default_rng(4)
    ndat, nmodel = 150, 700
    J = rng.standard_normal((ndat, nmodel)) * 0.1
    main = np.full(nmodel, 2.0)
    off = np.full(nmodel - 1, -0.45)
    Cm = sp.diags([off, main, off], [-1, 0, 1], format="csr")
    cd = rng.uniform(0.5, 2.0, ndat)
    dJ = torch.from_numpy(J).cuda()
    for kmax in (1, 5, 30, ndat):
        a, lo, hi = afx.spectral_alpha(dJ, Cm, cd, kmax)
        a_r, lo_r, hi_r = O.spectral_alpha_v2(J, Cm.toarray(), cd, kmax)
        print(f"kmax {kmax}: alpha {a:.12e} (oracle {a_r:.12e}), emin {lo:.6e} ({lo_r:.6e}), emax {hi:.6e}")
        assert abs(a - a_r) <= 1e-10 * a_r
        assert abs(hi - hi_r) <= 1e-10 * hi_r
        assert abs(lo - lo_r) <= 1e-8 * hi_r
    ev = np.linalg.eigvalsh(J @ Cm.toarray() @ J.T)
    assert abs(hi - ev[-1]) <= 1e-9 * ev[-1]


@pytest.mark.parametrize("stabiliser", ["MS", "MSG"])
def test_regularised_gn_step(stabiliser):
    """One model update as n_inv_ap.f90:380-430 runs it: Cm = BMI of the minimum-support (MS) or
    minimum-support-gradient (MSG, with GENDQ's R of a box mesh's crust cells) stabiliser,
    alpha = gamma * rho(J Cm J^T) / max(Cd) on the device, then getnewmodel; checked against the
    closed form in numpy."""
    import scipy.sparse as sp
    import torch
    from actfemtide_amd import afx, mesh
    from actfemtide_amd import regularization as G
    from oracle import oracle as O
    c = mesh.tide_case(6, 5, 8, seed=2)
    e2m, nmodel = G.crust_cell_models(c)
    rng = np.random.default_rng(9)
    ndat, beta, gamma = 40, 0.5, 0.1
    J = rng.standard_normal((ndat, nmodel)) * 0.05
    m_ref = np.full(nmodel, 2.0)
    m = m_ref + 0.3 * rng.standard_normal(nmodel)
    dm = m - m_ref
    if stabiliser == "MS":
        Cm = G.bmi_ms(dm, beta)
    else:
        R = G.model_roughening(c.n4, c.nphys1, e2m, nmodel)
        Cm = sp.csr_matrix(G.bmi_msg(G.ri_msg(R), dm, beta))
    cd = rng.uniform(0.5, 2.0, ndat)
    dJ = torch.from_numpy(J).cuda()
    a, _, _ = afx.spectral_alpha(dJ, Cm, cd, 30)
    alpha = gamma * a
    dobs, dcal = rng.standard_normal(ndat), rng.standard_normal(ndat)
    new, _ = afx.gn_update(dJ, Cm, cd, alpha, dobs, dcal, m, m_ref)
    Cd = Cm.toarray()
    if stabiliser == "MS":  # J Cm J^T symmetric: the Lanczos extreme is its largest eigenvalue
        ev = np.linalg.eigvalsh(J @ Cd @ J.T)
        assert a == pytest.approx(ev[-1] / cd.max(), rel=1e-8)
    # MSG's Cm is not symmetric; the reference runs the same recurrence on J Cm J^T regardless
    a_ref, _, _ = O.spectral_alpha_v2(J, Cd, cd, 30)
    assert a == pytest.approx(a_ref, rel=1e-9)
    Cmat = J @ Cd @ J.T + alpha * np.diag(cd)
    beta_ref = np.linalg.solve(Cmat, dobs - dcal + J @ dm)
    new_ref = m_ref + Cd @ J.T @ beta_ref
    err = np.abs(new - new_ref).max() / np.abs(new_ref).max()
    print(f"{stabiliser}: nmodel {nmodel}, alpha {alpha:.4e}, model error {err:.1e}")
    assert err < 1e-10
