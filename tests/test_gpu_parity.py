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
    assert abs(st[0].iter - ref["iter"]) <= max(2, ref["iter"] // 100)
    assert st[0].true_resid <= 1e-8 * 1.05
    A = O.to_scipy(o["INU"], o["IAU"], o["INL"], o["IAL"], o["D"], o["AU"], o["AL"])
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) <= 1.05e-8
