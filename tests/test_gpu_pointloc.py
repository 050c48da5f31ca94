"""Point location of PREPOBSCOEFF on the device (the bucket-grid search that replaces the scan of
FINDELEMENT0, m_mesh_type.f90:697-734) against the oracle's FINDELEMENT0 scan in id order:

- random points inside two meshes;
- points where several tets tie -- every mesh node, edge midpoints and face centroids -- where
  the scan's rule (the lowest tet id whose four barycentric coordinates are >= 0) decides;
- points outside the mesh, which the reference stops on ("GEGEGE! No corresponding element",
  :731-733) and the library returns as AFX_ERR_OUTSIDE.
"""
import numpy as np
import pytest

from tests.conftest import OMEGA_M2

pytestmark = pytest.mark.gpu


def _points(case, seed, nrand):
    rng = np.random.default_rng(seed)
    lo, hi = case.xyz.min(axis=0), case.xyz.max(axis=0)
    rand = lo + (hi - lo) * rng.uniform(0.0, 1.0, (nrand, 3))
    n4 = np.asarray(case.n4) - 1
    t = rng.choice(len(n4), 40, replace=False)
    v = case.xyz[n4[t]]                                    # (40, 4, 3)
    edge_mid = 0.5 * (v[:, 0] + v[:, 2])
    face_cen = (v[:, 0] + v[:, 1] + v[:, 3]) / 3.0
    return np.concatenate([rand, case.xyz, edge_mid, face_cen])


@pytest.mark.parametrize("dims", [(8, 8, 6), (20, 16, 12)])
def test_point_location_matches_findelement0(dims):
    from actfemtide_amd import afx, mesh
    from oracle import oracle as O
    c = mesh.tide_case(*dims, seed=5)
    pts = _points(c, seed=dims[0], nrand=300)
    print(f"[pointloc] {dims}: {len(pts)} points ({c.node} nodes), ntet={c.ntet}", flush=True)
    s, _, _ = afx.forward_system(c, OMEGA_M2)
    try:
        item, ve, vb, iele = s.prepobscoeff(pts)
    finally:
        s.close()
    item_r, ve_r, vb_r, iele_r = O.prepobscoeff(c, pts)
    bad = np.nonzero(iele != iele_r)[0]
    kinds = np.array(["random"] * 300 + ["node"] * c.node + ["edge midpoint"] * 40 + ["face centroid"] * 40)
    assert bad.size == 0, [(int(i), kinds[i], int(iele[i]), int(iele_r[i])) for i in bad[:10]]
    assert np.array_equal(item, item_r)
    assert np.allclose(ve, ve_r, rtol=1e-12, atol=1e-12) and np.allclose(vb, vb_r, rtol=1e-12, atol=1e-12)


def test_point_outside_mesh_is_an_error(case_tiny):
    from actfemtide_amd import afx
    s, _, _ = afx.forward_system(case_tiny, OMEGA_M2)
    try:
        hi = case_tiny.xyz.max(axis=0)
        for p in ([hi[0] + 1e-3, 0.0, -1.0], [0.0, 0.0, hi[2] + 50.0], [1e6, -1e6, 0.0]):
            with pytest.raises(afx.AfxError) as e:
                s.prepobscoeff(np.array([[0.0, 0.0, -1.0], p]))
            assert e.value.code == afx.AFX_ERR_OUTSIDE
    finally:
        s.close()
