"""GPU parity of the HIP path against the CPU oracle (same seeded inputs)."""
import numpy as np
import pytest

from tests.conftest import OMEGA_M2

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.fixture(scope="module")
def dev_small(case_small):
    from actfemtide_amd import afx
    print(f"[dev_small] building device system n={case_small.nline}", flush=True)
    return afx.forward_system(case_small, OMEGA_M2)


def test_pattern_identical(case_small, oracle_small, dev_small):
    s, _, _ = dev_small
    INU, IAU, INL, IAL = s.pattern()
    assert np.array_equal(INU, oracle_small["INU"])
    assert np.array_equal(IAU, oracle_small["IAU"])
    assert np.array_equal(INL, oracle_small["INL"])
    assert np.array_equal(IAL, oracle_small["IAL"])


def test_assembly_values(case_small, oracle_small, dev_small):
    s, b, bc = dev_small
    D, AU, AL = s.get_values()
    # matrix values within 1e-12 of the row scale (explicit zeros stay zero)
    scale = np.max(np.abs(oracle_small["D"]))
    assert np.max(np.abs(D - oracle_small["D"])) <= 1e-12 * scale
    assert np.max(np.abs(AU - oracle_small["AU"])) <= 1e-12 * scale
    assert np.max(np.abs(AL - oracle_small["AL"])) <= 1e-12 * scale
    assert np.array_equal(bc, oracle_small["bc"])
    assert _rel(b, oracle_small["b"]) <= 1e-12


def test_spmv(oracle_small, dev_small):
    from oracle import oracle as O
    s, _, _ = dev_small
    rng = np.random.default_rng(3)
    x = rng.standard_normal(s.n) + 1j * rng.standard_normal(s.n)
    o = oracle_small
    y_ref = O.spmv(o["INU"], o["IAU"], o["INL"], o["IAL"], o["D"], o["AU"], o["AL"], x)
    assert _rel(s.spmv(x), y_ref) <= 1e-13


@pytest.mark.parametrize("kind", ["SSOR", "IC", "DIAG"])
def test_precond(oracle_small, dev_small, kind):
    from oracle import oracle as O
    s, _, _ = dev_small
    o = oracle_small
    s.precond_setup(kind)
    DD_ref = O.precond_setup(kind, o["D"], o["INU"], o["AU"], o["INL"], o["IAL"], o["AL"])
    assert _rel(s.precond_diag(), DD_ref) <= 1e-11
    rng = np.random.default_rng(5)
    r = rng.standard_normal(s.n) + 1j * rng.standard_normal(s.n)
    z_ref = O.precond_apply(kind, o["INU"], o["IAU"], o["AU"], o["INL"], o["IAL"], o["AL"], DD_ref, r)
    assert _rel(s.precond_apply(r), z_ref) <= 1e-10


def test_cocg_ssor_matches_oracle(oracle_small, dev_small):
    from oracle import oracle as O
    s, b, _ = dev_small
    o = oracle_small
    ref = O.cocg13(o["INU"], o["IAU"], o["INL"], o["IAL"], o["D"], o["AU"], o["AL"], o["b"], precond="SSOR",
                   resid=1e-8, maxit=20000)
    x, st = s.solve(b, precond="SSOR", tol=1e-8, maxit=20000)
    print("oracle iters", ref["iter"], "gpu iters", st[0].iter, "true", st[0].true_resid, flush=True)
    assert st[0].error == 0
    # Last-bit differences (parallel reductions, FMA) are amplified by the ill-conditioned Krylov
    # recurrence: the count agrees within 8 % (the oracle itself moves 819 -> 860 on this system
    # when only its complex-division rounding changes).
    assert abs(st[0].iter - ref["iter"]) <= max(3, 0.08 * ref["iter"])
    assert st[0].true_resid <= 1e-8 * 1.05
    A = O.to_scipy(o["INU"], o["IAU"], o["INL"], o["IAL"], o["D"], o["AU"], o["AL"])
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) <= 1.05e-8


def _fields(case, x, sites, omega):
    """B (face rows) and E = -i omega (edge rows) at the sites, via the oracle's rows."""
    from oracle import oracle as O
    item, ve, vb, _ = O.prepobscoeff(case, sites)
    return np.concatenate([O.obs_apply(item, vb[c], x) for c in range(3)] +
                          [-1j * omega * O.obs_apply(item, ve[c], x) for c in range(2)])


def test_site_fields_match_oracle_solution(case_small, oracle_small, dev_small):
    """Fields at seafloor sites from the GPU solve agree with the oracle's COCG solution."""
    from oracle import oracle as O
    from tests.conftest import seafloor_sites
    s, b, _ = dev_small
    o = oracle_small
    ref = O.cocg13(o["INU"], o["IAU"], o["INL"], o["IAL"], o["D"], o["AU"], o["AL"], o["b"], precond="SSOR",
                   resid=1e-10, maxit=40000)
    x, st = s.solve(b, precond="SSOR", tol=1e-10, maxit=40000)
    sites = seafloor_sites(case_small, 8, 3)
    f_gpu = _fields(case_small, x, sites, OMEGA_M2)
    f_ref = _fields(case_small, ref["x"], sites, OMEGA_M2)
    assert np.max(np.abs(f_gpu - f_ref) / np.abs(f_ref)) <= 1e-6


def test_golden_cfg1_on_gpu():
    """Config-1 system (18,432 tets): device assembly equals the reference samples; the solve
    reaches 1e-8 in a count close to the reference COCG13_single's."""
    from actfemtide_amd import afx
    from tests.test_golden import case_of, load
    g = load("cfg1_ssor")
    c = case_of(g)
    s, b, bc = afx.forward_system(c, float(g["omega"]))
    D, AU, AL = s.get_values()
    scale = np.max(np.abs(D))
    for k, v in (("D", D), ("AU", AU), ("AL", AL)):
        assert np.max(np.abs(v[g[k + "_idx"]] - g[k + "_val"])) <= 1e-12 * scale, k
    assert np.max(np.abs(b[g["b_idx"]] - g["b_val"])) <= 1e-12 * np.max(np.abs(b))
    x, st = s.solve(b, precond="SSOR", tol=1e-8, maxit=40000)
    it_ref = int(g["iter"][0])
    print("cfg1 gpu iters", st[0].iter, "reference", it_ref, "true", st[0].true_resid, flush=True)
    assert st[0].error == 0 and st[0].true_resid <= 1.05e-8
    assert abs(st[0].iter - it_ref) <= 0.08 * it_ref


@pytest.mark.parametrize("name", ["tiny_ssor", "tiny_diag_nset2"])
def test_dropin_cocg13_single_matches_reference(name):
    """afx_cocg13_single (the COCG13_single mirror) on the reference's own assembled system."""
    from actfemtide_amd import afx
    from tests.test_golden import load
    g = load("tiny_ssor")
    gg = load(name)
    D, AU, AL, B = (np.array(g[k], np.complex128) for k in ("D", "AU", "AL", "b"))
    X = np.zeros_like(B)
    afx.cocg13_single_release()  # a first call, as in the reference's fresh program
    resid, it, err = afx.cocg13_single(D, AL, g["INL"], g["IAL"], AU, g["INU"], g["IAU"], B, X,
                                       precond=str(gg["precond"]), resid=1e-8, iter=20000, nset=int(gg["nset"]))
    afx.cocg13_single_release()
    it_ref = int(gg["iter"][0])
    print(name, "dropin iters", it, "reference", it_ref, flush=True)
    assert err == 0
    assert abs(it - it_ref) <= max(3, 0.08 * it_ref)
    assert np.linalg.norm(X - gg["x"]) <= 1e-5 * np.linalg.norm(gg["x"])
    # B is scaled and restored (NSET = 2 leaves D/AU/AL scaled, as the reference does)
    assert np.allclose(B, g["b"], rtol=1e-14, atol=0)


def test_dropin_keeps_its_system_between_calls():
    """Repeated drop-in calls as a per-frequency caller makes them: new values on the same pattern
    (the kept system) give what a fresh system gives, a changed pattern is noticed, and
    afx_cocg13_single_release frees the kept system."""
    from actfemtide_amd import afx
    from tests.test_golden import load
    g = load("tiny_ssor")
    INL, IAL, INU, IAU = g["INL"], g["IAL"], g["INU"], g["IAU"]

    def call(scale):
        D, AU, AL = (np.array(g[k], np.complex128) * scale for k in ("D", "AU", "AL"))
        B = np.array(g["b"], np.complex128)
        X = np.zeros_like(B)
        r, it, err = afx.cocg13_single(D, AL, INL, IAL, AU, INU, IAU, B, X, precond="SSOR", resid=1e-8, iter=20000)
        return X, it, err

    afx.cocg13_single_release()
    x1, it1, e1 = call(1.0)
    x2, it2, e2 = call(2.0)          # same pattern, kept system, new values
    afx.cocg13_single_release()
    x2f, it2f, e2f = call(2.0)       # the same values on a fresh system
    assert e1 == e2 == e2f == 0
    assert abs(it2 - it2f) <= max(3, 0.08 * it2f)
    assert np.linalg.norm(x2 - x2f) <= 1e-6 * np.linalg.norm(x2f)
    assert np.linalg.norm(2.0 * x2 - x1) <= 1e-6 * np.linalg.norm(x1)  # A -> 2A halves x
    # another mesh (another pattern) through the same entry point, against the oracle
    from actfemtide_amd import mesh
    from oracle import oracle as O
    o = O.forward_system(mesh.tide_case(5, 4, 6, seed=3), 2.0 * np.pi / 44714.0)
    D, AU, AL, B = (np.array(o[k], np.complex128) for k in ("D", "AU", "AL", "b"))
    X = np.zeros_like(B)
    r, it, err = afx.cocg13_single(D, AL, o["INL"], o["IAL"], AU, o["INU"], o["IAU"], B, X, precond="SSOR",
                                   resid=1e-8, iter=20000)
    ref = O.cocg13(o["INU"], o["IAU"], o["INL"], o["IAL"], o["D"], o["AU"], o["AL"], o["b"], precond="SSOR",
                   resid=1e-8, maxit=20000)
    # iteration counts on this mesh are sensitive to rounding: the oracle itself takes 369 to 406
    # iterations when the matrix entries are perturbed by 1e-15 (relative), so the solution and its
    # true residual are compared instead
    A = O.to_scipy(o["INU"], o["IAU"], o["INL"], o["IAL"], o["D"], o["AU"], o["AL"])
    b = np.array(o["b"], np.complex128)
    print("another mesh: drop-in", it, "oracle", ref["iter"], flush=True)
    assert err == 0 and np.linalg.norm(b - A @ X) / np.linalg.norm(b) <= 1e-8
    assert np.linalg.norm(X - ref["x"]) <= 1e-5 * np.linalg.norm(ref["x"])
    afx.cocg13_single_release()


def test_dropin_ic_tracks_reference_history():
    """'IC' (D-ILU) stagnates on this system in the reference too: ERROR = -100 after MAXIT,
    ITER = MAXIT + 1, and the first residuals follow the reference's fort.12 history."""
    from actfemtide_amd import afx
    from tests.test_golden import load
    g = load("tiny_ssor")
    gi = load("tiny_ic")
    D, AU, AL, B = (np.array(g[k], np.complex128) for k in ("D", "AU", "AL", "b"))
    X = np.zeros_like(B)
    resid, it, err = afx.cocg13_single(D, AL, g["INL"], g["IAL"], AU, g["INU"], g["IAU"], B, X, precond="IC",
                                       resid=1e-8, iter=int(gi["maxit"]), nset=1)
    assert err == -100 and it == int(gi["maxit"]) + 1
    assert resid > 1e-3


def test_zero_rhs_and_maxit(dev_small):
    s, b, _ = dev_small
    x, st = s.solve(np.zeros_like(b), np.ones_like(b), precond="SSOR")
    assert st[0].iter == 0 and st[0].resid == 0 and np.all(x == 0)
    x, st = s.solve(b, precond="SSOR", tol=1e-30, maxit=7)
    assert st[0].error == -100 and st[0].iter == 8


def test_solve_cg_mirror(oracle_small):
    """solveCG: SSOR, RESID = 1e-4 (single-precision literal), ITER = 3000, X0 = 0."""
    from actfemtide_amd import afx
    from oracle import oracle as O
    o = oracle_small
    x = afx.solve_cg(o["D"], o["AL"], o["INL"], o["IAL"], o["AU"], o["INU"], o["IAU"], o["b"])
    A = O.to_scipy(o["INU"], o["IAU"], o["INL"], o["IAL"], o["D"], o["AU"], o["AL"])
    r = np.linalg.norm(o["b"] - A @ x) / np.linalg.norm(o["b"])
    assert 1e-6 < r <= 1.0000000474974513e-4 * 1.01


def test_dropin_call_sequence_matches_reference():
    """Successive drop-in calls on the same arrays, NSET 1, 0, 2, 1, 0, as the reference ran them
    (tests/golden/tiny_nset_seq.npz): NSET = 0 reuses the kept preconditioner, NSET = 2 scales
    D/AU/AL in place and saves SCALE, the next NSET = 1 call scales B by it, NSET = 0 divides B by it
    on exit (COCG13_single.f90:39-62,98-100,326-329). Each call's x against the reference's, its
    iteration count within 8%, and the returned B."""
    from actfemtide_amd import afx
    from tests.test_golden import load
    g0 = load("tiny_ssor")
    g = load("tiny_nset_seq")
    D, AU, AL = (np.array(g0[k], np.complex128) for k in ("D", "AU", "AL"))
    b = np.array(g0["b"], np.complex128)
    INL, IAL, INU, IAU = g0["INL"], g0["IAL"], g0["INU"], g0["IAU"]
    afx.cocg13_single_release()
    calls = [(int(g["nset"]), str(g["precond"]), 1.0, g["x"], b, g["iter"])]
    calls += [(int(k), str(pc), float(f), g[f"seq_x{q}"], g[f"seq_b{q}"], g[f"seq_iter{q}"])
              for q, (k, pc, f) in enumerate(zip(g["seq_nset"], g["seq_precond"], g["seq_bfac"]), start=1)]
    for q, (nset, pc, f, x_ref, b_ref, it_ref) in enumerate(calls):
        B = b * f
        X = np.zeros_like(B)
        resid, it, err = afx.cocg13_single(D, AL, INL, IAL, AU, INU, IAU, B, X, precond=pc, resid=1e-8,
                                           iter=20000, nset=nset)
        dx = np.linalg.norm(X - x_ref) / np.linalg.norm(x_ref)
        db = np.linalg.norm(B - b_ref) / np.linalg.norm(b_ref)
        print(f"call {q} nset {nset}: iter {it} (reference {int(it_ref[0])}), x {dx:.1e}, B {db:.1e}", flush=True)
        assert err == 0 and abs(it - int(it_ref[0])) <= max(3, 0.08 * int(it_ref[0]))
        assert dx <= 1e-5 and db <= 1e-14
    # NSET = 0 with no saved setup, and with another preconditioner than the saved one: ERROR = 101
    afx.cocg13_single_release()
    B = b.copy()
    X = np.zeros_like(B)
    assert afx.cocg13_single(D, AL, INL, IAL, AU, INU, IAU, B, X, precond="SSOR", nset=0)[2] == 101
    assert np.array_equal(B, b)
    afx.cocg13_single(D, AL, INL, IAL, AU, INU, IAU, B, X, precond="SSOR", resid=1e-8, iter=20000, nset=1)
    assert afx.cocg13_single(D, AL, INL, IAL, AU, INU, IAU, b.copy(), np.zeros_like(b), precond="DIAG",
                             nset=0)[2] == 101
    afx.cocg13_single_release()


def test_raised_error_flag_ends_the_solve_promptly(dev_small, monkeypatch):
    """The abort path of the persistent kernel's bounded waits: with the device error flag raised
    by the host before the first launch, every block leaves at its first grid barrier and the
    solve returns AFX_ERR_HIP at once instead of iterating to maxit; the next solve is normal."""
    import time

    from actfemtide_amd import afx
    s, b, _ = dev_small
    monkeypatch.setenv("AFX_TEST_RAISE_ERR", "1")
    t0 = time.perf_counter()
    with pytest.raises(afx.AfxError) as e:
        s.solve(b, precond="SSOR", tol=1e-30, maxit=100000, chunk=100000)
    dt = time.perf_counter() - t0
    print(f"raised flag: {e.value} after {dt:.3f}s", flush=True)
    assert e.value.code == afx.AFX_ERR_HIP and dt < 10.0
    monkeypatch.delenv("AFX_TEST_RAISE_ERR")
    x, st = s.solve(b, precond="SSOR", tol=1e-8, maxit=20000)
    assert st[0].error == 0 and st[0].true_resid <= 1.05e-8
