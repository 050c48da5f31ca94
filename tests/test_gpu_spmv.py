"""The standalone device SpMV (afx_spmv_device, q = A p of COCG13_single.f90:270-290) against the
oracle's SpMV (orc_spmv, the reference's row loop) and a dense product, on:

- the assembled edge-element system (rows of 6-19 entries: 4 lanes per row);
- random symmetric patterns whose longest rows select 8 lanes per row (21-24 entries) and
  16 lanes per row with the loop for entries past the unrolled 32 (rows of up to 60 entries);
- a 1x1 system, a system whose size is not a multiple of a block's rows, and rows holding only
  the diagonal.
"""
import numpy as np
import pytest

from tests.conftest import OMEGA_M2

pytestmark = pytest.mark.gpu


def _random_symmetric(n, max_upper, seed, long_rows=()):
    """Split-triangle arrays (INU, IAU, INL, IAL 1-based, D, AU, AL) of a random complex-symmetric
    pattern: up to max_upper upper entries per row, and the rows in long_rows given as many more."""
    rng = np.random.default_rng(seed)
    upper = []
    for i in range(n):
        m = min(n - 1 - i, int(rng.integers(0, max_upper + 1)))
        if i in long_rows:
            m = n - 1 - i
        upper.append(np.sort(rng.choice(np.arange(i + 1, n), size=m, replace=False)) if m else np.zeros(0, int))
    INU = np.zeros(n + 1, np.int32)
    for i in range(n):
        INU[i + 1] = INU[i] + len(upper[i])
    IAU = (np.concatenate(upper) + 1).astype(np.int32) if INU[-1] else np.zeros(0, np.int32)
    AU = rng.standard_normal(INU[-1]) + 1j * rng.standard_normal(INU[-1])
    lower = [[] for _ in range(n)]
    for i in range(n):
        for q in range(INU[i], INU[i + 1]):
            lower[IAU[q] - 1].append((i, AU[q]))
    INL = np.zeros(n + 1, np.int32)
    for i in range(n):
        INL[i + 1] = INL[i] + len(lower[i])
    IAL = np.array([j + 1 for i in range(n) for j, _ in sorted(lower[i])], np.int32)
    AL = np.array([v for i in range(n) for _, v in sorted(lower[i])], np.complex128)
    D = 4.0 + rng.standard_normal(n) + 1j * rng.standard_normal(n)
    return INU, IAU, INL, IAL, D, AU, AL


def _dense(INU, IAU, INL, IAL, D, AU, AL):
    n = len(D)
    A = np.diag(D).astype(np.complex128)
    for i in range(n):
        for q in range(INU[i], INU[i + 1]):
            A[i, IAU[q] - 1] += AU[q]
        for q in range(INL[i], INL[i + 1]):
            A[i, IAL[q] - 1] += AL[q]
    return A


def _device_spmv(s, x):
    import torch
    dx = torch.from_numpy(np.ascontiguousarray(x, np.complex128)).cuda()
    dy = torch.full((s.n,), complex(np.nan, np.nan), dtype=torch.complex128, device="cuda")
    ms = s.spmv_device(dx, dy, timed=True)
    return dy.cpu().numpy(), ms


def test_spmv_device_edge_system(case_small, oracle_small):
    from actfemtide_amd import afx
    from oracle import oracle as O
    o = oracle_small
    s, _, _ = afx.forward_system(case_small, OMEGA_M2)
    try:
        rng = np.random.default_rng(7)
        x = rng.standard_normal(s.n) + 1j * rng.standard_normal(s.n)
        y, ms = _device_spmv(s, x)
        y_ref = O.spmv(o["INU"], o["IAU"], o["INL"], o["IAL"], o["D"], o["AU"], o["AL"], x)
        err = np.abs(y - y_ref).max() / np.abs(y_ref).max()
        print(f"edge system n={s.n}: relative difference {err:.1e}, {1e3 * ms:.1f} us", flush=True)
        assert err <= 1e-13 and ms > 0
    finally:
        s.close()


@pytest.mark.parametrize("n,max_upper,long_rows", [
    (1, 0, ()),
    (37, 3, ()),           # longest row <= 20: 4 lanes per row; n not a multiple of 128
    (301, 7, ()),          # longest row 23 entries: 8 lanes per row
    (300, 11, ()),         # longest row 32: 16 lanes per row, unrolled
    (500, 6, (3, 77, 250, 499)),  # rows of up to ~500 entries: 16 lanes and the tail loop
])
def test_spmv_device_ragged_patterns(n, max_upper, long_rows):
    from actfemtide_amd import afx
    from oracle import oracle as O
    INU, IAU, INL, IAL, D, AU, AL = _random_symmetric(n, max_upper, seed=n, long_rows=long_rows)
    rows = np.diff(INU) + np.diff(INL) + 1
    print(f"[spmv ragged] n={n} longest row {rows.max()}, diagonal-only rows {int((rows == 1).sum())}", flush=True)
    s = afx.System.from_pattern(INU, IAU, INL, IAL)
    try:
        s.set_values(D, AU, AL)
        rng = np.random.default_rng(n + 1)
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        y, _ = _device_spmv(s, x)
        y_orc = O.spmv(INU, IAU, INL, IAL, D, AU, AL, x)
        y_dense = _dense(INU, IAU, INL, IAL, D, AU, AL) @ x
        scale = np.abs(y_dense).max()
        assert np.abs(y_orc - y_dense).max() <= 1e-13 * scale
        assert np.abs(y - y_orc).max() <= 1e-13 * scale
        assert (rows.max() <= 20) == (n in (1, 37))
    finally:
        s.close()


def test_spmv_device_rejects_aliasing(case_tiny):
    import torch
    from actfemtide_amd import afx
    s, _, _ = afx.forward_system(case_tiny, OMEGA_M2)
    try:
        d = torch.zeros(s.n, dtype=torch.complex128, device="cuda")
        with pytest.raises(afx.AfxError):
            s.spmv_device(d, d)
        with pytest.raises(ValueError):
            s.spmv_device(d[:-1], torch.zeros(s.n - 1, dtype=torch.complex128, device="cuda"))
    finally:
        s.close()
