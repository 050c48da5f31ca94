"""Several meshes in one persistent launch (actfemtide_amd.meshbatch) on the GPU, with per-system
scalars (afx_create_stacked + afx_solve_stacked_device) and with shared scalars:

- the stacked system kept on the device holds exactly block_diag_split's pattern and values;
- three meshes of different sizes, seeds and frequencies (M2, S2, K1), right-hand sides scaled
  far apart: every part's FP64 true residual, recomputed with scipy from the part's own values,
  is <= 1e-8, and the library's per-part figure agrees with it;
- each part's magnetic field B = curl x at 100 interior points (PREPOBSCOEFF's face-basis rows)
  agrees with the part solved on its own to 1e-6. The solution vectors themselves are not
  compared: a 1e-8 residual leaves the discrete-gradient components in air unresolved (the
  oracle on these meshes: x of the stacked and of the own solve 3-7% apart, and each as far from
  the direct solution, while B agrees with the direct solution to 5e-9);
- per system, each part stops on its own recurrence residual (<= tol) at its own iteration. The
  counts are not compared with the part's own solve: on this operator they move with the
  rounding of the reductions (part 1 here: 1,682 iterations in the oracle's row order, 935 in
  its own GPU solve, 1,692 stacked), while the residuals and B do not;
- a zero right-hand side gives that part x = 0 and does not count in the joint tolerance.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OMEGAS = (2 * math.pi / 44714.0, 2 * math.pi / 43200.0, 2 * math.pi / 86164.0)


@pytest.fixture(scope="module")
def batch():
    from actfemtide_amd import afx, mesh
    from actfemtide_amd.meshbatch import MeshBatch
    parts, bs, pts = [], [], []
    rng = np.random.default_rng(9)
    for i, dims in enumerate([(5, 4, 6), (8, 8, 6), (10, 9, 8)]):
        c = mesh.tide_case(*dims, seed=60 + i)
        s, b, _ = afx.forward_system(c, OMEGAS[i])
        parts.append(s)
        bs.append(b * (1.0, 1e4, 1e-3)[i])
        lo, hi = c.xyz.min(axis=0), c.xyz.max(axis=0)
        pts.append(lo + (hi - lo) * rng.uniform(0.05, 0.95, (100, 3)))
    print(f"[meshbatch] parts n={[p.n for p in parts]}", flush=True)
    batches = {mode: MeshBatch(parts, per_system=(mode == "per_system")) for mode in ("per_system", "shared")}
    for mb in batches.values():
        mb.points = pts
    yield batches, parts, bs
    for mb in batches.values():
        mb.close()
    for p in parts:
        p.close()


def _scipy(s):
    from oracle import oracle as O
    return O.to_scipy(*s.pattern(), *s.get_values())


@pytest.mark.parametrize("mode", ["per_system", "shared"])
def test_stacked_system_is_block_diagonal(batch, mode):
    from actfemtide_amd.meshbatch import block_diag_split
    batches, parts, bs = batch
    mb = batches[mode]
    ref = block_diag_split([p.pattern() for p in parts], [p.get_values() for p in parts])
    got = mb.system.pattern() + mb.system.get_values()
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)
    assert mb.system.nsys == (3 if mode == "per_system" else 1)


def _b_field(s, pts, x):
    item, val_e, val_b, iele = s.prepobscoeff(pts)
    return np.stack([(val_b[c] * x[item - 1]).sum(axis=1) for c in range(3)])


@pytest.mark.parametrize("mode", ["per_system", "shared"])
def test_joint_solve_meets_tolerance_per_part(batch, mode):
    batches, parts, bs = batch
    mb = batches[mode]
    xs, st, tr = mb.solve(bs, tol=1e-8, precond="SSOR", maxit=100000)
    print(f"[meshbatch] {mode}: iterations {[x.iter for x in st]}, errors {[x.error for x in st]}, "
          f"true residuals {tr}", flush=True)
    assert all(x.error == 0 for x in st)
    for s, (p, b, x) in enumerate(zip(parts, bs, xs)):
        A = _scipy(p)
        r = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
        assert r <= 1e-8
        assert tr[s] == pytest.approx(r, rel=1e-3)  # two FP64 evaluations of a residual near 1e-9
        x_own, st_own = p.solve(b, precond="SSOR", tol=1e-8, maxit=100000)
        B, B_own = _b_field(p, mb.points[s], x), _b_field(p, mb.points[s], x_own)
        d = np.linalg.norm(B - B_own) / np.linalg.norm(B_own)
        print(f"[meshbatch] part {s}: residual {r:.2e}, own solve {st_own[0].iter} iterations, "
              f"B difference {d:.1e}", flush=True)
        assert d <= 1e-6
        if mode == "per_system":
            assert st[s].resid <= 1e-8 and st[s].iter > 0
    if mode == "per_system":
        assert len({x.iter for x in st}) > 1  # the parts stopped at different iterations


@pytest.mark.parametrize("mode", ["per_system", "shared"])
def test_zero_right_hand_side_part(batch, mode):
    batches, parts, bs = batch
    mb = batches[mode]
    bz = [bs[0], np.zeros(parts[1].n, np.complex128), bs[2]]
    xs, st, tr = mb.solve(bz, tol=1e-8, precond="SSOR", maxit=100000)
    assert all(x.error == 0 for x in st)
    if mode == "per_system":
        assert st[1].iter == 0 and st[1].resid == 0.0
    assert not np.any(xs[1])
    assert tr[0] <= 1e-8 and tr[2] <= 1e-8 and tr[1] == 0.0


def test_frequency_stack_matches_each_frequency():
    """One system stacked at three frequencies (MeshBatch.frequency_stack): copy c is the system
    at omega_c (values equal afx_get_values after afx_set_omega), each copy ends below 1e-8 and
    its B field agrees with the system solved at omega_c on its own."""
    from actfemtide_amd import afx, mesh
    from actfemtide_amd.meshbatch import MeshBatch
    c = mesh.tide_case(8, 8, 6, seed=70)
    s, b, _ = afx.forward_system(c, OMEGAS[0])
    try:
        mb = MeshBatch.frequency_stack(s, OMEGAS)
        assert s.omega == OMEGAS[0] and mb.m == 3 and mb.system.nsys == 3
        D, AU, AL = mb.system.get_values()
        rng = np.random.default_rng(3)
        lo, hi = c.xyz.min(axis=0), c.xyz.max(axis=0)
        pts = lo + (hi - lo) * rng.uniform(0.05, 0.95, (100, 3))
        xs, st, tr = mb.solve([b, b, b], tol=1e-8, precond="SSOR", maxit=100000)
        print(f"[meshbatch] frequency stack: iterations {[x.iter for x in st]}, true residuals {tr}", flush=True)
        for q, w in enumerate(OMEGAS):
            s.set_omega(w)
            Dq, AUq, ALq = s.get_values()
            o0, o1 = mb.offsets[q], mb.offsets[q + 1]
            assert np.array_equal(D[o0:o1], Dq)
            A = _scipy(s)
            r = np.linalg.norm(b - A @ xs[q]) / np.linalg.norm(b)
            assert st[q].error == 0 and r <= 1e-8
            x_own, _ = s.solve(b, precond="SSOR", tol=1e-8, maxit=100000)
            B, B_own = _b_field(s, pts, xs[q]), _b_field(s, pts, x_own)
            assert np.linalg.norm(B - B_own) <= 1e-6 * np.linalg.norm(B_own)
        mb.close()
    finally:
        s.close()


def test_one_stacked_system_is_the_plain_solve(case_tiny):
    """afx_create_stacked with a single system keeps the plain partition and kernel path: the same
    iterates, bit for bit, as afx_solve_device on the same arrays."""
    import torch
    from actfemtide_amd import afx
    from tests.conftest import OMEGA_M2
    s, b, _ = afx.forward_system(case_tiny, OMEGA_M2)
    t = afx.System.from_stacked([0, s.n], *s.pattern())
    try:
        t.set_values(*s.get_values())
        assert t.nsys == 1 and s.nsys == 1
        dB = torch.from_numpy(b).cuda()
        x1, x2 = torch.zeros_like(dB), torch.zeros_like(dB)
        st1, _ = s.solve_device(dB, x1, precond="SSOR", tol=1e-8, maxit=20000)
        st2, _ = t.solve_stacked_device(dB, x2, precond="SSOR", tol=1e-8, maxit=20000)
        assert len(st2) == 1 and st2[0].iter == st1[0].iter and st2[0].error == 0
        assert torch.equal(x1, x2)
    finally:
        t.close()
        s.close()
