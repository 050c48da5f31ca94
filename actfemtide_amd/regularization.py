"""Model regularisation of the active-source inversion and the choice of its trade-off alpha.

  GENDQ                 the model roughening matrix R (library: afx_model_roughening),
                        src/src_inv_ap/m_modelroughness.f90:575-862
  CALROUGHNESS_SM/MS/MSG the roughness of a model against the reference model (:312-404)
  GENBMI_SM / MS / MSG  the model covariance Cm = BMI of the smoothest-model, minimum-support and
                        minimum-support-gradient stabilisers (:72-252, :466-566); GENRI_MSG's
                        RI = (R + 0.01 I)^-1 (:410-462)
  alphaspectralradius_v2 alpha = rho(J Cm J^T) / max(Cd) (library: afx.spectral_alpha)

Models are the reference's `logrho_model` (log10 resistivity); dm = model - model_ref.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import scipy.sparse as sp

from . import afx


def model_roughening(n4, nphys1: int, ele2model, nmodel: int, iactive=None) -> sp.csr_matrix:
    """GENDQ's R (m_modelroughness.f90:575-862) for land elements nphys1+1..ntet: -1 between models
    whose land elements share a tet face, the number of neighbour models on the diagonal."""
    n4 = afx._i32(n4).reshape(-1, 4)
    e2m = afx._i32(ele2model)
    act = None if iactive is None else afx._i32(iactive)
    nrow, nnz = C.c_int(), C.c_longlong()
    args = (n4.shape[0], afx._p(n4), int(nphys1), afx._p(e2m), int(nmodel), afx._p(act))
    afx._err(afx.lib().afx_model_roughening(*args, C.byref(nrow), C.byref(nnz), None, None, None))
    ptr = np.zeros(nrow.value + 1, np.int32)
    col = np.zeros(nnz.value, np.int32)
    val = np.zeros(nnz.value)
    afx._err(afx.lib().afx_model_roughening(*args, C.byref(nrow), C.byref(nnz), afx._p(ptr), afx._p(col),
                                            afx._p(val)))
    return sp.csr_matrix((val, col, ptr), shape=(nrow.value, nrow.value))


def crust_cell_models(case):
    """One model per crust cell (its 6 tets), numbered by increasing cell id: (ele2model 1-based
    over the nphys2 crust elements, nmodel)."""
    cells = np.asarray(case.cell_of_tet)[case.nphys1:]
    uniq, inv = np.unique(cells, return_inverse=True)
    return (inv + 1).astype(np.int32), int(uniq.size)


def roughness_sm(BM, dm) -> float:
    """CALROUGHNESS_SM (:359-404): dm^T BM dm."""
    dm = np.asarray(dm, np.float64)
    return float(np.dot(dm, BM @ dm))


def roughness_ms(dm, beta: float) -> float:
    """CALROUGHNESS_MS (:312-357): 0.5 sum dm^2 / (dm^2 + beta^2)."""
    m2 = np.asarray(dm, np.float64) ** 2
    return float(0.5 * np.sum(m2 / (m2 + beta ** 2)))


def roughness_msg(R, dm, beta: float) -> float:
    """CALROUGHNESS_MSG (:254-310): 0.5 q^T R q with q = dm / sqrt(dm^2 + beta^2)."""
    dm = np.asarray(dm, np.float64)
    q = dm / np.sqrt(dm ** 2 + beta ** 2)
    return float(0.5 * np.dot(q, R @ q))


def bmi_ms(dm, beta: float) -> sp.csr_matrix:
    """GENBMI_MS (:72-150): the diagonal Cm = BMI, ((dm^2 + beta^2)^2) / beta^2."""
    m2 = np.asarray(dm, np.float64) ** 2
    return sp.diags((m2 + beta ** 2) ** 2 / beta ** 2, format="csr")


# PARDISOphase3 (src/src_inv_ap/m_solveCM_ap.f90:133-178) keeps the entries of the inverse above
# this magnitude (conv_full2ccs, src/common/m_matrix.f90:219-240)
INVERSE_THRESHOLD = 1e-15


def _inverse(A) -> np.ndarray:
    X = np.linalg.inv(A)
    X[np.abs(X) <= INVERSE_THRESHOLD] = 0.0
    return X


def ri_msg(R, eps: float = 1e-2) -> np.ndarray:
    """GENRI_MSG (:410-462): RI = (R + eps I)^-1 (eps = 0.01 on the diagonal), dense, entries of
    magnitude <= 1e-15 dropped as PARDISOphase3 drops them."""
    A = np.asarray(sp.csr_matrix(R).todense(), np.float64) + eps * np.eye(R.shape[0])
    return _inverse(A)


def bm_sm(R) -> sp.csr_matrix:
    """GENBMI_SM's BM = R^T R (GENDQ :843-848, taken before eps is added)."""
    R = sp.csr_matrix(R)
    return (R.T @ R).tocsr()


def bmi_sm(R, eps: float = 1e-2) -> np.ndarray:
    """GENBMI_SM (:466-566): Cm = BMI = (R^T R + eps I)^-1, dense, small entries dropped."""
    A = np.asarray(bm_sm(R).todense(), np.float64) + eps * np.eye(R.shape[0])
    return _inverse(A)


def bm_msg(R, dm, beta: float) -> sp.csr_matrix:
    """GENBMI_MSG's BM for the first iteration (:225-238): W1 R W2 with W1 = beta^2 /
    (dm^2 + beta^2)^(3/2) and W2 = 1 / sqrt(dm^2 + beta^2)."""
    m2 = np.asarray(dm, np.float64) ** 2
    w1 = beta ** 2 / (m2 + beta ** 2) ** 1.5
    w2 = 1.0 / np.sqrt(m2 + beta ** 2)
    return (sp.diags(w1) @ sp.csr_matrix(R) @ sp.diags(w2)).tocsr()


def bmi_msg(RI, dm, beta: float) -> np.ndarray:
    """GENBMI_MSG (:152-252): Cm = BMI = W2I RI W1I with W1I = (dm^2 + beta^2)^(3/2) / beta^2 and
    W2I = sqrt(dm^2 + beta^2) (diagonal)."""
    m2 = np.asarray(dm, np.float64) ** 2
    w1 = (m2 + beta ** 2) ** 1.5 / beta ** 2
    w2 = np.sqrt(m2 + beta ** 2)
    return (w2[:, None] * np.asarray(RI)) * w1[None, :]
