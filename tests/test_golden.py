"""Pin the CPU oracle (and the host-side numbering / file formats) against golden vectors that
the reference's own Fortran produced (tests/golden/make_golden.py, oracle/_ref)."""
import os

import numpy as np
import pytest

from tests.conftest import OMEGA_M2

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return dict(np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False))


def case_of(g):
    from actfemtide_amd import mesh
    return mesh.tide_case(int(g["nx"]), int(g["ny"]), int(g["nz"]), seed=int(g["seed"]))


def _sums(a):
    a = np.asarray(a, np.int64).ravel()
    return a.size, int(a.sum()), int((a * (np.arange(a.size) % 9973 + 1)).sum())


def _fortran_2d(a, rows, cols):
    """column-major (rows, cols) Fortran array as a C (rows, cols) array"""
    return np.asarray(a).reshape(cols, rows).T


@pytest.fixture(scope="module")
def tiny():
    g = load("tiny_ssor")
    c = case_of(g)
    from oracle import oracle as O
    return g, c, O.forward_system(c, float(g["omega"]))


def test_edge_numbering_matches_mkline_v2(tiny):
    g, c, _ = tiny
    line = _fortran_2d(g["mk_line"], 2, c.nline).T if False else np.asarray(g["mk_line"]).reshape(-1, 2)
    assert np.array_equal(line, c.line)
    n6 = _fortran_2d(g["mk_n6line"], c.ntet, 6)
    assert np.array_equal(n6, c.n6line)


@pytest.mark.parametrize("name", ["small_ssor", "cfg1_ssor"])
def test_edge_numbering_checksums(name):
    g = load(name)
    c = case_of(g)
    n6_f = c.n6line.T.ravel()  # Fortran column-major order of n6line(ntet, 6)
    assert _sums(c.line.ravel()) == (int(g["mk_line_len"]), int(g["mk_line_sum"]), int(g["mk_line_wsum"]))
    assert _sums(n6_f) == (int(g["mk_n6line_len"]), int(g["mk_n6line_sum"]), int(g["mk_n6line_wsum"]))


def test_readers_match_reference(tiny, tmp_path):
    from actfemtide_amd import formats
    g, c, _ = tiny
    p = formats.write_case(c, str(tmp_path))
    v = formats.read_vxyz(p["vxyz"], c.nodes_ocean, c.node)
    f = formats.read_fxyz(p["fxyz"], c.nodes_ocean, c.node)
    assert np.array_equal(v, _fortran_2d(g["vxyz"], 3, c.node).T)
    assert np.array_equal(f, _fortran_2d(g["fxyz"], 3, c.node).T)
    xyz, n4, flag = formats.read_em3d_msh(p["mesh"])
    assert np.array_equal(xyz, c.xyz) and np.array_equal(n4, c.n4) and np.array_equal(flag[:, 1], c.group)
    line, n6 = formats.read_lineinfo(p["line"])
    assert np.array_equal(line, c.line) and np.array_equal(n6, c.n6line)
    _, rho, sig = formats.read_cond(p["cond"])
    assert np.array_equal(sig, c.sigma_crust)


def test_pattern_matches_reference(tiny):
    g, _, o = tiny
    for k in ("INU", "INL", "IAU", "IAL"):
        assert np.array_equal(o[k], g[k]), k


def test_genmat_copy_bc_match_reference(tiny):
    from oracle import oracle as O
    g, c, o = tiny
    D, AU, b = O.genmat(c, float(g["omega"]), o["INU"], o["IAU"])
    scale = np.max(np.abs(g["D_genmat"]))
    assert np.max(np.abs(D - g["D_genmat"])) <= 1e-13 * scale
    assert np.max(np.abs(AU - g["AU_genmat"])) <= 1e-13 * scale
    assert np.max(np.abs(b - g["b_genmat"])) <= 1e-13 * np.max(np.abs(g["b_genmat"]))
    AL = O.copy_ul(o["INU"], o["IAU"], o["INL"], o["IAL"], AU)
    assert np.max(np.abs(AL - g["AL_copy"])) <= 1e-13 * scale
    assert np.array_equal(o["bc"].astype(np.int32), g["line_bc"])
    for k in ("D", "AU", "AL"):
        assert np.max(np.abs(o[k] - g[k])) <= 1e-13 * scale, k
    assert np.max(np.abs(o["b"] - g["b"])) <= 1e-13 * np.max(np.abs(g["b"]))
    # explicit zeros of SET_BC_ICCG stay in the pattern, identically
    assert np.array_equal(o["AU"] == 0, g["AU"] == 0)


@pytest.mark.parametrize("name", ["small_ssor", "cfg1_ssor"])
def test_assembly_samples_match_reference(name):
    from oracle import oracle as O
    g = load(name)
    c = case_of(g)
    o = O.forward_system(c, float(g["omega"]))
    for k in ("INU", "INL", "IAU", "IAL"):
        assert _sums(o[k]) == (int(g[k + "_len"]), int(g[k + "_sum"]), int(g[k + "_wsum"])), k
    assert _sums(o["bc"].astype(np.int64)) == (int(g["line_bc_len"]), int(g["line_bc_sum"]), int(g["line_bc_wsum"]))
    scale = np.max(np.abs(o["D"]))
    for k in ("D", "AU", "AL"):
        assert np.max(np.abs(o[k][g[k + "_idx"]] - g[k + "_val"])) <= 1e-13 * scale, k
        assert abs(np.linalg.norm(o[k]) - float(g[k + "_norm"])) <= 1e-12 * float(g[k + "_norm"])
    assert np.max(np.abs(o["b"][g["b_idx"]] - g["b_val"])) <= 1e-13 * np.max(np.abs(o["b"]))


@pytest.mark.parametrize("name", ["tiny_ssor", "tiny_ic", "tiny_diag_nset2", "small_ssor", "cfg1_ssor"])
def test_cocg13_matches_reference(name):
    """Identical iteration count and residual history; x within 1e-13 relative (the oracle's
    operations follow the reference's order: the measured difference is 0)."""
    from oracle import oracle as O
    g = load(name)
    c = case_of(g)
    o = O.forward_system(c, float(g["omega"]))
    r = O.cocg13(o["INU"], o["IAU"], o["INL"], o["IAL"], o["D"], o["AU"], o["AL"], o["b"],
                 precond=str(g["precond"]), resid=float(g["tol"]), maxit=int(g["maxit"]), nset=int(g["nset"]),
                 history=True)
    it_ref, err_ref = (int(v) for v in g["iter"])
    assert r["error"] == err_ref
    assert abs(r["iter"] - it_ref) <= (0 if name.startswith("tiny") else 2)
    n = min(len(r["hist"]), len(g["hist_resid"]), 200)
    # fort.12 holds 8 significant digits
    assert np.allclose(r["hist"][:n], g["hist_resid"][:n], rtol=2e-6 if str(g["precond"]) != "IC" else 1e-4)
    if err_ref == 0:
        assert abs(r["resid"] - float(g["resid"][0])) <= 0.05 * float(g["resid"][0]) + 1e-12
        x_ref = g["x"]
        print(name, "x rel diff", np.linalg.norm(r["x"] - x_ref) / np.linalg.norm(x_ref))
        assert np.linalg.norm(r["x"] - x_ref) <= 1e-13 * np.linalg.norm(x_ref)


def test_cocg13_call_sequence_matches_reference():
    """Successive COCG13_single calls on the same arrays (NSET 1, 0, 2, 1, 0): the oracle with its
    saved state reproduces the reference's iteration counts, x and returned B of every call --
    the SAVEd DD reused by NSET = 0, SCALE set by NSET = 2 and applied to B by the later NSET = 1
    call, B divided by SCALE after NSET = 0 (COCG13_single.f90:39-62,98-100,326-329)."""
    from oracle import oracle as O
    g = load("tiny_nset_seq")
    c = case_of(g)
    o = O.forward_system(c, float(g["omega"]))
    st = O.Cocg13State(len(o["D"]))
    D, AU, AL = o["D"], o["AU"], o["AL"]
    r = O.cocg13(o["INU"], o["IAU"], o["INL"], o["IAL"], D, AU, AL, o["b"], precond=str(g["precond"]),
                 resid=float(g["tol"]), maxit=int(g["maxit"]), nset=int(g["nset"]), state=st)
    assert r["iter"] == int(g["iter"][0])
    for q, (nset, pc, f) in enumerate(zip(g["seq_nset"], g["seq_precond"], g["seq_bfac"]), start=1):
        D, AU, AL = r["D"], r["AU"], r["AL"]  # what the previous call left (scaled in place by NSET = 2)
        r = O.cocg13(o["INU"], o["IAU"], o["INL"], o["IAL"], D, AU, AL, o["b"] * float(f), precond=str(pc),
                     resid=float(g["tol"]), maxit=int(g["maxit"]), nset=int(nset), state=st)
        it_ref, err_ref = (int(v) for v in g[f"seq_iter{q}"])
        x_ref, b_ref = g[f"seq_x{q}"], g[f"seq_b{q}"]
        dx = np.linalg.norm(r["x"] - x_ref) / np.linalg.norm(x_ref)
        db = np.linalg.norm(r["B"] - b_ref) / np.linalg.norm(b_ref)
        print(f"call {q} nset {nset}: iter {r['iter']} (reference {it_ref}), x {dx:.1e}, B {db:.1e}")
        assert (r["iter"], r["error"]) == (it_ref, err_ref)
        assert dx <= 1e-13 and db <= 1e-14


def test_observation_basis_matches_reference(tiny):
    from oracle import oracle as O
    g, c, _ = tiny
    item, ve, vb, iele = O.prepobscoeff(c, g["sites"])
    assert np.array_equal(iele + 1, g["obs_iele"])
    for o in range(len(iele)):
        t = iele[o]
        sg = np.where(c.n6line[t] < 0, -1.0, 1.0)
        ids = np.abs(c.n6line[t])
        order = np.argsort(ids)
        w_ref = g["obs_w"][o]  # (6 edges, 3 comps)
        wf_ref = g["obs_wf"][o]
        for j in range(6):
            e = order[j]
            for comp in range(3):
                assert ve[comp, o, j] == pytest.approx(w_ref[e, comp] * sg[e], rel=1e-12, abs=1e-14)
                assert vb[comp, o, j] == pytest.approx(wf_ref[e, comp] * sg[e], rel=1e-12, abs=1e-14)
