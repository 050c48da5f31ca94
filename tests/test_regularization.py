"""Model regularisation of the active inversion (src/src_inv_ap/m_modelroughness.f90) on the CPU.

What pins it:
- the oracle's GENDQ (the reference's D, Q, Q^T D Q steps) equals, on the box mesh with one
  model per crust cell, the graph Laplacian of face-adjacent crust cells written from the cell
  indices (Kuhn tets: cells that share a face share two tet faces, cells that share only an
  edge share none);
- the library's afx_model_roughening (host code, sorted face keys) equals the oracle on cell
  models, on 2x2x1 blocks of cells and with inactive models, and refuses an isolated model;
- the roughness and Cm formulas reduce to closed forms: zero roughness for dm = 0 and for a
  constant dm under a Laplacian, the MS limit nmodel/2, BMI_MSG = BMI_MS when RI = I, and
  (R + 0.01 I) RI = I.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from actfemtide_amd import afx
from actfemtide_amd import regularization as G
from oracle import oracle as O


@pytest.fixture(scope="module")
def box():
    from actfemtide_amd import mesh
    return mesh.tide_case(6, 5, 8, seed=2)


def _cell_laplacian(c, ele2model, nmodel):
    cells = np.asarray(c.cell_of_tet)[c.nphys1:]
    model_of_cell = {}
    for cell, m in zip(cells, ele2model):
        model_of_cell[int(cell)] = int(m) - 1
    L = np.zeros((nmodel, nmodel))
    for cell, m in model_of_cell.items():
        i, j, k = cell % c.nx, (cell // c.nx) % c.ny, cell // (c.nx * c.ny)
        for di, dj, dk in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
            ii, jj, kk = i + di, j + dj, k + dk
            nb = ii + c.nx * (jj + c.ny * kk)
            if 0 <= ii < c.nx and 0 <= jj < c.ny and 0 <= kk < c.nz and nb in model_of_cell:
                L[m, model_of_cell[nb]] = -1.0
    for m in range(nmodel):
        L[m, m] = -L[m].sum()
    return L


def test_oracle_gendq_is_the_cell_laplacian(box):
    e2m, nmodel = G.crust_cell_models(box)
    R = O.model_roughening(box.n4, box.nphys1, e2m, nmodel)
    assert np.array_equal(R, _cell_laplacian(box, e2m, nmodel))
    assert np.all(R.sum(axis=1) == 0)


def _blocks(c):
    cells = np.asarray(c.cell_of_tet)[c.nphys1:]
    i, j, k = cells % c.nx, (cells // c.nx) % c.ny, cells // (c.nx * c.ny)
    key = (i // 2) + 100 * (j // 2) + 10000 * k
    _, inv = np.unique(key, return_inverse=True)
    return (inv + 1).astype(np.int32), int(inv.max() + 1)


@pytest.mark.parametrize("models", ["cells", "blocks", "inactive"])
def test_library_gendq_matches_oracle(box, models):
    if models == "blocks":
        e2m, nmodel = _blocks(box)
        act = None
    else:
        e2m, nmodel = G.crust_cell_models(box)
        act = None
        if models == "inactive":
            act = np.ones(nmodel, np.int32)
            act[[3, 17, nmodel - 1]] = 0
    R = G.model_roughening(box.n4, box.nphys1, e2m, nmodel, act).toarray()
    R_ref = O.model_roughening(box.n4, box.nphys1, e2m, nmodel, act)
    assert R.shape == R_ref.shape
    assert np.array_equal(R, R_ref)


def test_isolated_model_is_refused(box):
    e2m, nmodel = G.crust_cell_models(box)
    # one model for all land elements: it has no neighbour model
    one = np.ones_like(e2m)
    with pytest.raises(afx.AfxError):
        G.model_roughening(box.n4, box.nphys1, one, 1)
    with pytest.raises(ValueError):
        O.model_roughening(box.n4, box.nphys1, one, 1)


def test_roughness_and_cm_closed_forms(box):
    e2m, nmodel = G.crust_cell_models(box)
    R = G.model_roughening(box.n4, box.nphys1, e2m, nmodel)
    beta = 0.3
    rng = np.random.default_rng(0)
    dm = rng.standard_normal(nmodel)
    assert G.roughness_ms(np.zeros(nmodel), beta) == 0.0
    assert G.roughness_ms(np.full(nmodel, 1e8), beta) == pytest.approx(nmodel / 2, rel=1e-12)
    assert G.roughness_ms(np.r_[beta, np.zeros(nmodel - 1)], beta) == pytest.approx(0.25, rel=1e-15)
    assert abs(G.roughness_msg(R, np.full(nmodel, 0.7), beta)) < 1e-12
    assert G.roughness_msg(R, dm, beta) > 0.0
    BM = (R.T @ R).tocsr()
    assert abs(G.roughness_sm(BM, np.full(nmodel, -1.3))) < 1e-10
    e = np.zeros(nmodel)
    e[5] = 1.0
    deg = R[5, 5]
    assert G.roughness_sm(BM, e) == pytest.approx(deg * deg + deg, rel=1e-15)
    b = G.bmi_ms(np.array([0.0, 1.0, -2.0]), 1.0).diagonal()
    assert np.array_equal(b, [1.0, 4.0, 25.0])
    assert np.allclose(G.bmi_msg(np.eye(nmodel), dm, beta), G.bmi_ms(dm, beta).toarray(), rtol=1e-14, atol=0)
    RI = G.ri_msg(R)
    assert np.allclose((R.toarray() + 0.01 * np.eye(nmodel)) @ RI, np.eye(nmodel), atol=1e-10)
    assert np.all(np.abs(RI[RI != 0]) > G.INVERSE_THRESHOLD)
    BMI = G.bmi_sm(R)
    assert np.allclose((G.bm_sm(R).toarray() + 0.01 * np.eye(nmodel)) @ BMI, np.eye(nmodel), atol=1e-8)
    assert np.array_equal(G.bm_sm(R).toarray(), R.toarray().T @ R.toarray())
    # MSG's first-iteration BM is the inverse of BMI_MSG's diagonal weights around R
    W = G.bm_msg(sp.identity(nmodel, format="csr"), dm, beta).diagonal()
    assert np.allclose(W * G.bmi_ms(dm, beta).diagonal(), 1.0, rtol=1e-14)


def test_roughening_rejects_bad_input(box):
    e2m, nmodel = G.crust_cell_models(box)
    bad = e2m.copy()
    bad[0] = nmodel + 1
    with pytest.raises(afx.AfxError):
        G.model_roughening(box.n4, box.nphys1, bad, nmodel)
    with pytest.raises(afx.AfxError):
        G.model_roughening(box.n4, box.ntet, e2m, nmodel)  # no land elements
    # every model inactive: an empty matrix, not an error
    R = G.model_roughening(box.n4, box.nphys1, e2m, nmodel, np.zeros(nmodel, np.int32))
    assert R.shape == (0, 0)
