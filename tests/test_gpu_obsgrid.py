"""The seafloor current (IXYHOUT) on the GPU against the oracle's restatement, and the grid
observation path end to end: a forward solve, observation points 1 m above every seafloor
triangle (PREPOBSCOEFF_XYSBTM), their field coefficients from the library's point location
against the oracle's PREPOBSCOEFF, the fields, the currents and the result files."""
import os

import numpy as np
import pytest

from tests.conftest import OMEGA_M2

pytestmark = pytest.mark.gpu


def test_ixyh_device_matches_oracle():
    from actfemtide_amd import afx, obsgrid as G
    from oracle import oracle as O
    rng = np.random.default_rng(4)
    node, nobs = 40, 300
    vxyz = rng.standard_normal((node, 3)) + 1j * rng.standard_normal((node, 3))
    vxyz[:, 2] = 0.0
    fxyz = rng.standard_normal((node, 3)) * 3e4
    lens = rng.integers(0, 5, nobs)
    stack = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    item = rng.integers(1, node + 1, stack[-1]).astype(np.int32)
    val = rng.uniform(0, 1, stack[-1])
    xyz = rng.standard_normal((nobs, 3)) * 5
    exy = rng.standard_normal((nobs, 2)) + 1j * rng.standard_normal((nobs, 2))
    got = afx.ixyh(xyz, exy, stack, item, val, vxyz, fxyz, G.SIGMA_OCEAN_IXYH)
    ref = O.ixyhout(xyz, exy, stack, item, val, vxyz, fxyz, G.SIGMA_OCEAN_IXYH)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)  # the same IEEE operations in the same order


def test_ixyh_rejects_bad_rows():
    from actfemtide_amd import afx
    with pytest.raises(afx.AfxError):
        afx.ixyh(np.zeros((1, 3)), np.zeros((1, 2)), np.array([0, 1]), np.array([5]), np.ones(1),
                 np.zeros((3, 3)), np.zeros((3, 3)), 3.3)


def test_seafloor_currents_end_to_end(case_small, tmp_path):
    from actfemtide_amd import afx, obsgrid as G
    from oracle import oracle as O
    c = case_small
    s, b, bc = afx.forward_system(c, OMEGA_M2)
    try:
        x, st = s.solve(b, precond="SSOR", tol=1e-8, maxit=20000)
        assert st[0].error == 0
        n3, k23 = G.seafloor_mesh(c)
        pts, coeff_vf = G.xysbtm_points(c.xyz, n3, k23)
        item, ve, vb, iele = s.prepobscoeff(pts)
        oitem, ove, ovb, oiele = O.prepobscoeff(c, pts)
        assert np.array_equal(item, oitem) and np.array_equal(iele, oiele)
        assert np.allclose(ve, ove, rtol=0, atol=1e-12 * np.abs(ove).max())
        assert np.allclose(vb, ovb, rtol=0, atol=1e-12 * np.abs(ovb).max())
        import torch
        dX = torch.from_numpy(x).cuda()
        fields = afx.calobsebcomp(s, item, ve, vb, dX, OMEGA_M2)[:, :, 0]  # (5, nobs)
    finally:
        s.close()
    r = G.ixyhout(pts, fields, coeff_vf, c.vxyz, c.fxyz)
    ref = O.ixyhout(pts, fields[3:5].T, *coeff_vf, c.vxyz, c.fxyz, G.SIGMA_OCEAN_IXYH)
    for key, want in zip(("ixyh_e", "ixyh_vF", "ixyh", "vh", "h"), ref):
        assert np.array_equal(r[key], want), key
    # points 1 m above the seafloor of the one-layer sea: h = depth - 1 m
    assert np.allclose(r["h"], c.h - 0.001, rtol=0, atol=1e-12)
    assert np.abs(r["ixyh_vF"]).max() > 0 and np.abs(r["ixyh_e"]).max() > 0
    out = str(tmp_path) + "/"
    G.write_coord(out + "coord_xy2D.dat", pts)
    path = G.ixyh_path(out, 1.0 / 44714.0)
    G.write_ixyh(path, pts, r)
    rows = np.loadtxt(path)
    assert rows.shape == (pts.shape[0], 26)
    assert np.allclose(rows[:, 11], r["ixyh"][:, 0].real, rtol=5e-7, atol=1e-300)
    G.write_outfreqfiles2(out, "xy", 1.0 / 44714.0, pts, fields[None])
    assert os.path.exists(out + "xy0.22E-04_S1.dat")
