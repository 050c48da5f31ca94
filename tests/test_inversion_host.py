"""Host logic of the Gauss-Newton driver (actfemtide_amd.inversion) that needs no GPU: the crust
parameters, model2cond and the phase wrap."""
import numpy as np


def test_crust_parameters_number_cells_as_the_models():
    from actfemtide_amd import inversion, mesh
    from actfemtide_amd import regularization as reg
    c = mesh.tide_case(6, 5, 8, seed=3)
    ptr, tets, cells = inversion.crust_parameters(c)
    e2m, nmodel = reg.crust_cell_models(c)
    assert len(ptr) - 1 == nmodel == cells.size
    assert np.all(np.diff(ptr) == 6)
    # tet t of parameter j is crust element t - nphys1 of model e2m - 1 == j
    for j in range(nmodel):
        t = tets[ptr[j]:ptr[j + 1]]
        assert np.all(e2m[t - c.nphys1] - 1 == j)
        assert np.all(c.cell_of_tet[t] == cells[j])


def test_sigma_of_model_sets_only_the_crust():
    from actfemtide_amd import inversion, mesh
    c = mesh.tide_case(6, 5, 8, seed=3)
    ptr, tets, cells = inversion.crust_parameters(c)
    m = np.linspace(-3.0, 0.0, len(ptr) - 1)
    s = inversion.sigma_of_model(c, ptr, tets, m)
    s0 = c.sigma_tet()
    crust = c.group >= mesh.GROUP_CRUST
    assert np.array_equal(s[~crust], s0[~crust])
    for j in (0, 7, len(m) - 1):
        assert np.allclose(s[tets[ptr[j]:ptr[j + 1]]], 10.0 ** m[j], rtol=1e-15)


def test_wrap_degrees():
    from actfemtide_amd.inversion import wrap_degrees
    d = np.array([0.0, 179.0, 181.0, -181.0, 359.0, -359.0, 540.0])
    assert np.allclose(wrap_degrees(d), [0.0, 179.0, -179.0, 179.0, -1.0, 1.0, -180.0])
