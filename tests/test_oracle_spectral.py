"""The oracle's LANCZOSMINMAX / alphaspectralradius_v2 restatement (src/src_inv_ap/m_spectral.f90)
pinned by what the mathematics fixes:
- Lanczos run to the full dimension reproduces the matrix's extreme eigenvalues (the Krylov
  space is the whole space), for a random symmetric matrix with both signs in its spectrum;
- a well-separated extreme eigenvalue is found to high accuracy in a few steps;
- a start vector that is an exact eigenvector ends the recurrence after one step (beta = 0),
  with that eigenvalue;
- alpha = max(|emin|, |emax|) / max(cd) with J Cm J^T formed as the reference does.
"""
import numpy as np

from oracle import oracle as O


def test_full_lanczos_recovers_extremes():
    rng = np.random.default_rng(0)
    n = 14
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.linspace(-3.0, 7.0, n)
    A = (Q * lam) @ Q.T
    emin, emax, al, be = O.lanczos_minmax(A, n)
    assert abs(emin - (-3.0)) < 1e-9 and abs(emax - 7.0) < 1e-9
    assert len(al) == n


def test_separated_extreme_converges_quickly():
    n = 400
    d = np.linspace(1.0, 2.0, n)
    d[123] = 50.0
    emin, emax, _, _ = O.lanczos_minmax(np.diag(d), 30)
    assert abs(emax - 50.0) < 1e-10
    assert 1.0 <= emin < 1.2


def test_invariant_start_vector_stops_after_one_step():
    # n = 4: q = ones / 2 = 0.5 exactly and A q = 3 q, so the first residual is exactly 0 and the
    # reference's test beta .eq. 0 (m_spectral.f90:149) ends the recurrence
    emin, emax, al, be = O.lanczos_minmax(3.0 * np.eye(4), 30)
    assert len(al) == 1 and be[0] == 0.0
    assert emin == 3.0 and emax == 3.0


def test_alpha_formula():
    rng = np.random.default_rng(3)
    ndat, nmodel = 12, 40
    J = rng.standard_normal((ndat, nmodel))
    Cm = np.diag(rng.uniform(0.5, 2.0, nmodel))
    cd = rng.uniform(0.1, 4.0, ndat)
    alpha, emin, emax = O.spectral_alpha_v2(J, Cm, cd, ndat)
    ev = np.linalg.eigvalsh(J @ Cm @ J.T)
    assert abs(alpha - ev[-1] / cd.max()) < 1e-10 * alpha
    assert abs(emin - ev[0]) < 1e-9 * ev[-1]
